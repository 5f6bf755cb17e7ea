// Standalone FEM seams of one model in the reference's layouts (fem.hip).
#pragma once

#include "model.h"

namespace mpf {

// A = S(gamma) + alpha R(gamma) of one frame as a SparseMatrix: 32 slots per row,
// entry (row, slot) at row*32 + slot, empty slots with column id 0xFFFFFFFF; rows and
// column ids in the caller's node numbering, slots in first-insertion order.
void fem_assemble(Model& M, const float* d_gamma, double alpha, uint32_t* d_ids, float* d_values);
// U (ceil16(M) x D, column-major) (+)= W Phi[:, 0:D]
void fem_measure(Model& M, const float* d_phi, int ldphi, int additive, float* d_U);
// J (ceil16(M) ceil16(D) x E, column-major) (+)= calcJacobianKernel of Phi
void fem_jacobian(Model& M, const float* d_phi, int ldphi, const float* d_gamma, int additive, float* d_J);

// Fixed (voltage) sources, see MeshSetup::fixed: element contributions of the eliminated
// couplings (rows outside D) and of the full rows of D, CSR over rows; Fsh = B_clip Pi
struct FixDev {
    const int32_t* cpl_off; const int32_t* cpl_node; const int32_t* cpl_elem;
    const double* cpl_k; const double* cpl_m;
    const int32_t* drow_off; const int32_t* drow_node; const int32_t* drow_elem;
    const double* drow_k; const double* drow_m;
    const double* Fsh;
    int np, Rp, C;
};
// every job's per-slice excitation F_s = [b_D; -A_s,ID b_D] (after the assembly)
void launch_fixed_rhs(const JobDev* d_jobs, int njobs, const FixDev& fd, hipStream_t st);
// every job's electrode currents cur_s = A_s Phi_s on the rows of D (after the solve)
void launch_fixed_currents(const JobDev* d_jobs, int njobs, const FixDev& fd, int use64, hipStream_t st);

}  // namespace mpf
