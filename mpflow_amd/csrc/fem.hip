// Standalone FEM seams of one model, in the reference's own layouts and node numbering:
//   fem_assemble  Equation::update (src/fem/equation.cpp:261-274; updateMatrixKernel and
//                 updateSystemMatrixKernel, src/fem/equation_kernel.cu:60-157): the
//                 system matrix A = S(gamma) + alpha R(gamma) of one frame as a
//                 SparseMatrix (32 slots per row, sparse_matrix.cpp:48-114)
//   fem_measure   EIT::applyMeasurementPattern (src/models/eit.cpp:220-245):
//                 U (+)= W Phi[:, 0:D]
//   fem_jacobian  Equation::calcJacobian (src/fem/equation.cpp:282-310; calcJacobianKernel,
//                 src/fem/equation_kernel.cu:197-245)
// The model keeps its mesh in the device numbering (rows renumbered for locality);
// these entry points translate rows and column ids back through `node_of_row`.
#include "fem.h"

namespace mpf {

constexpr int kRefSlots = 32;
constexpr uint32_t kInvalidIndex = 0xFFFFFFFFu;

__device__ __forceinline__ double mat_param(int logarithmic, double ref, float g) {
    // matParam (src/fem/equation_kernel.cu:35-43)
    return logarithmic ? ref * exp(2.302585092994046 * (double)g / 10.0) : ref * (double)g;
}

__device__ __forceinline__ double mat_param_deriv(int logarithmic, double ref, float g) {
    // matParamDeriv (src/fem/equation_kernel.cu:45-53)
    return logarithmic ? (2.302585092994046 / 10.0) * ref * exp(2.302585092994046 * (double)g / 10.0) : ref;
}

// one thread per (row, slot), slot fastest: coalesced row-major output.  Entries are
// FP64 sums of sigma_e (K_e + alpha M_e) over the gather plan, rounded once.
__global__ void __launch_bounds__(256) k_fem_export(int n, int np, int wpad, const int32_t* __restrict__ ids,
                                                    const int32_t* __restrict__ diag, const int32_t* __restrict__ off,
                                                    const int32_t* __restrict__ elem, const double* __restrict__ ck,
                                                    const double* __restrict__ cm, const float* __restrict__ gamma,
                                                    int logarithmic, double ref, double alpha,
                                                    const int32_t* __restrict__ node_of_row,
                                                    uint32_t* __restrict__ out_ids, float* __restrict__ out_val) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * kRefSlots) return;
    const int row = (int)(t / kRefSlots), q = (int)(t % kRefSlots);
    uint32_t cid = kInvalidIndex;
    float v = 0.0f;
    if (q < wpad) {
        const size_t ent = (size_t)q * np + row;
        const int col = ids[ent];
        // padding slots repeat the row's own id outside its diagonal slot
        if (col != row || q == diag[row]) {
            double a = 0.0;
            for (int i = off[ent]; i < off[ent + 1]; ++i) {
                const int el = elem[i];
                const double sg = el < 0 ? 1.0 : mat_param(logarithmic, ref, gamma[el]);
                a += sg * (ck[i] + alpha * cm[i]);
            }
            cid = (uint32_t)node_of_row[col];
            v = (float)a;
        }
    }
    const size_t o = (size_t)node_of_row[row] * kRefSlots + q;
    if (out_ids) out_ids[o] = cid;
    out_val[o] = v;
}

// U[m + d*Mp] (+)= sum_q W[m, node_q] Phi[node_q, d] for d < D, m < Mp (rows m >= M are 0)
__global__ void __launch_bounds__(256) k_fem_measure(int D, int M, int Mp, const int32_t* __restrict__ w_off,
                                                     const int32_t* __restrict__ w_node,
                                                     const double* __restrict__ w_val,
                                                     const int32_t* __restrict__ node_of_row,
                                                     const float* __restrict__ phi, int ldphi, int additive,
                                                     float* __restrict__ U) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Mp * D) return;
    const int m = i % Mp, d = i / Mp;
    double v = 0.0;
    if (m < M)
        for (int q = w_off[m]; q < w_off[m + 1]; ++q)
            v += w_val[q] * (double)phi[(size_t)node_of_row[w_node[q]] + (size_t)d * ldphi];
    const size_t o = (size_t)m + (size_t)d * Mp;
    U[o] = additive ? U[o] + (float)v : (float)v;
}

// One block per element: J[d*Mp + m, e] (+)= sigma'(gamma_e) sum_ij Phi[n_i, d] K_e[i][j] Phi[n_j, D + m];
// rows with d >= D or m >= M are 0 (the reference zeroes those potentials).
__global__ void __launch_bounds__(256) k_fem_jacobian(int E, int k, int D, int M, int Dp, int Mp,
                                                      const int32_t* __restrict__ elements,
                                                      const float* __restrict__ ek,
                                                      const int32_t* __restrict__ node_of_row,
                                                      const float* __restrict__ phi, int ldphi,
                                                      const float* __restrict__ gamma, int logarithmic, double ref,
                                                      int additive, float* __restrict__ J) {
    const int e = blockIdx.x;
    __shared__ float dphi[6][64];
    __shared__ float T[6][64];
    __shared__ float K[36];
    __shared__ int nodes[6];
    if (threadIdx.x < k) nodes[threadIdx.x] = node_of_row[elements[(size_t)e * k + threadIdx.x]];
    for (int i = threadIdx.x; i < k * k; i += blockDim.x) K[i] = ek[(size_t)e * k * k + i];
    __syncthreads();
    __shared__ float mphi[6][64];
    for (int i = threadIdx.x; i < k * (D + M); i += blockDim.x) {
        const int a = i / (D + M), c = i % (D + M);
        const float v = phi[(size_t)nodes[a] + (size_t)c * ldphi];
        if (c < D) dphi[a][c] = v; else mphi[a][c - D] = v;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < k * M; i += blockDim.x) {
        const int a = i / M, m = i % M;
        float t = 0.0f;
        for (int b = 0; b < k; ++b) t += K[a * k + b] * mphi[b][m];
        T[a][m] = t;
    }
    __syncthreads();
    const float sd = (float)mat_param_deriv(logarithmic, ref, gamma[e]);
    const int rows = Mp * Dp;
    for (int idx = threadIdx.x; idx < rows; idx += blockDim.x) {
        const int d = idx / Mp, m = idx % Mp;
        float v = 0.0f;
        if (d < D && m < M)
            for (int a = 0; a < k; ++a) v += dphi[a][d] * T[a][m];
        v *= sd;
        const size_t o = (size_t)idx + (size_t)e * rows;
        J[o] = additive ? J[o] + v : v;
    }
}

// Voltage sources (applyMixedBoundaryCondition, eit.cpp:89-91,159-194): the device matrix
// holds identity rows on D and no columns of D, so the excitation of slice s is b_D on D
// and -A_s,iD b_D elsewhere, from the kept coupling contributions at this frame's sigma.
__global__ void __launch_bounds__(256) k_fixed_rhs(const JobDev* __restrict__ jobs, FixDev fd) {
    const JobDev& J = jobs[blockIdx.z];
    const int s = blockIdx.y;
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= fd.np * fd.Rp) return;
    const int row = idx / fd.Rp, u = idx % fd.Rp;
    double v = 0.0;
    if (fd.drow_off[row + 1] > fd.drow_off[row]) {
        v = fd.Fsh[idx];
    } else {
        const double al = J.alpha64[s];
        for (int q = fd.cpl_off[row]; q < fd.cpl_off[row + 1]; ++q) {
            const int el = fd.cpl_elem[q];
            const double sg = el < 0 ? 1.0 : mat_param(J.logarithmic, J.ref, J.gamma[el]);
            v -= sg * (fd.cpl_k[q] + al * fd.cpl_m[q]) * fd.Fsh[(size_t)fd.cpl_node[q] * fd.Rp + u];
        }
    }
    const_cast<double*>(J.F)[s * J.fstride + idx] = v;
}

// electrode currents (eit.cpp:185-190): cur_s = A_s Phi_s on the rows of D with the full
// (uneliminated) rows; the other rows of cur stay 0
__global__ void __launch_bounds__(256) k_fixed_currents(const JobDev* __restrict__ jobs, FixDev fd, int use64) {
    const JobDev& J = jobs[blockIdx.z];
    const int s = blockIdx.y;
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= fd.np * fd.Rp) return;
    const int row = idx / fd.Rp, u = idx % fd.Rp;
    const int q0 = fd.drow_off[row], q1 = fd.drow_off[row + 1];
    if (q0 == q1) return;
    const double al = J.alpha64[s];
    double a = 0.0;
    for (int q = q0; q < q1; ++q) {
        const int el = fd.drow_elem[q];
        const double sg = el < 0 ? 1.0 : mat_param(J.logarithmic, J.ref, J.gamma[el]);
        const size_t i = vec_index(J, s, fd.drow_node[q], u);
        a += sg * (fd.drow_k[q] + al * fd.drow_m[q]) * (use64 ? J.x64[i] : (double)J.x[i]);
    }
    J.cur[vec_index(J, s, row, u)] = a;
}

void launch_fixed_rhs(const JobDev* d_jobs, int njobs, const FixDev& fd, hipStream_t st) {
    dim3 grid((unsigned)((fd.np * fd.Rp + 255) / 256), fd.C, njobs);
    k_fixed_rhs<<<grid, 256, 0, st>>>(d_jobs, fd);
    MPF_HIP(hipGetLastError());
}

void launch_fixed_currents(const JobDev* d_jobs, int njobs, const FixDev& fd, int use64, hipStream_t st) {
    dim3 grid((unsigned)((fd.np * fd.Rp + 255) / 256), fd.C, njobs);
    k_fixed_currents<<<grid, 256, 0, st>>>(d_jobs, fd, use64);
    MPF_HIP(hipGetLastError());
}

void fem_assemble(Model& M, const float* d_gamma, double alpha, uint32_t* d_ids, float* d_values) {
    const MeshSetup& S = M.S;
    MPF_REQUIRE(S.wpad <= kRefSlots, "fem_assemble: rows wider than the SparseMatrix's 32 slots");
    const size_t n = (size_t)S.N * kRefSlots;
    k_fem_export<<<(unsigned)((n + 255) / 256), 256, 0, M.ctx->stream>>>(
        S.N, S.np, S.wpad, M.ids.p, M.diag_slot.p, M.ctb_off.p, M.ctb_elem.p, M.ctb_k.p, M.ctb_m.p, d_gamma,
        S.logarithmic, S.ref, alpha, M.node_of_row.p, d_ids, d_values);
    MPF_HIP(hipGetLastError());
}

void fem_measure(Model& M, const float* d_phi, int ldphi, int additive, float* d_U) {
    const MeshSetup& S = M.S;
    MPF_REQUIRE(!S.cem_coupled, "mpf_eit_measure: the coupled electrode model has no per-harmonic measurement");
    MPF_REQUIRE(ldphi >= S.N, "fem_measure: ldphi < rows of the fields");
    const int n = S.Mp * S.D;
    k_fem_measure<<<(n + 255) / 256, 256, 0, M.ctx->stream>>>(S.D, S.M, S.Mp, M.w_off.p, M.w_node.p, M.w_val.p,
                                                             M.node_of_row.p, d_phi, ldphi, additive, d_U);
    MPF_HIP(hipGetLastError());
}

void fem_jacobian(Model& M, const float* d_phi, int ldphi, const float* d_gamma, int additive, float* d_J) {
    MPF_REQUIRE(!M.S.cem_coupled, "mpf_jacobian: the coupled electrode model has no per-harmonic Jacobian seam");
    const MeshSetup& S = M.S;
    MPF_REQUIRE(ldphi >= S.N, "fem_jacobian: ldphi < rows of the fields");
    MPF_REQUIRE(S.k <= 6 && S.D <= 64 && S.M <= 64, "fem_jacobian: unsupported shape");
    k_fem_jacobian<<<S.E, 256, 0, M.ctx->stream>>>(S.E, S.k, S.D, S.M, S.Dp, S.Mp, M.elements.p, M.ek.p,
                                                   M.node_of_row.p, d_phi, ldphi, d_gamma, S.logarithmic, S.ref,
                                                   additive, d_J);
    MPF_HIP(hipGetLastError());
}

}  // namespace mpf
