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

}  // namespace mpf
