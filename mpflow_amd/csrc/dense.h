// Launch interface of the dense kernels (dense.hip).
#pragma once

#include "cplx.h"
#include "jobs.h"

namespace mpf {

// measurement map W (CSR over measurement patterns), shared by all jobs of a mesh
struct MeasDev {
    const int32_t* w_off; const int32_t* w_node; const double* w_val;
    const int32_t* colmap; const double* colscale;   // pattern column -> solved lane, scale
    int D, M, Dp, Mp;
};

// element data for the Jacobian, shared by all jobs of a mesh
struct JacDev {
    const int32_t* elements;     // [E][k]
    const float* ek; const float* em;   // [E][k*k]
    const float* gamma;          // [jobs][E]
    const int32_t* colmap; const double* colscale;   // pattern column -> solved lane, scale
    int E, k, D, M, Dp, Mp, exact, logarithmic;
    double ref;
    float alpha[kMaxSlices];
    float weight[kMaxSlices];
    // nullable: elements in a locality order (neighbouring elements share nodes); XCD x
    // processes the x-th contiguous eighth of it, so node fields are re-read from its L2
    const int32_t* order;
};

void launch_voltages(const JobDev* d_jobs, int njobs, const MeasDev& md, int use64, float* U32,
                     double* U64, hipStream_t st);
// Jb (nullable): a BF16 copy of J for launch_gram_bf16
void launch_jacobian(const JobDev* d_jobs, int njobs, const JacDev& jd, int use64, float* J,
                     hipStream_t st, uint16_t* Jb = nullptr);

// G_b = A_b A_b^T for `batch` column-major A_b (n x nk, lda; n, lda multiples of 4) on
// MFMA with 128 x 128 tiles; G_b full (both triangles), column-major n x n
bool gram_nt_applies(int n, int lda);
// max_slabs > 0 caps the contraction slabs (and so the workspace) at that many
size_t gram_nt_workspace_bytes(int n, int nk, int batch, int max_slabs = 0);
void launch_gram_nt(const float* A, size_t strideA, int n, int nk, int lda, float* G, size_t strideG, int batch,
                    void* work, hipStream_t st, int max_slabs = 0);
// the same Gram from a BF16 copy of A (raw bfloat16 bits) with 256 x 256 tiles; n and
// lda multiples of 8
bool gram_bf16_applies(int n, int lda);
size_t gram_bf16_workspace_bytes(int n, int nk, int batch);
void launch_gram_bf16(const uint16_t* A, size_t strideA, int n, int nk, int lda, float* G, size_t strideG,
                      int batch, void* work, hipStream_t st);
// b[i] = bf16(a[i]) (round to nearest even)
void launch_to_bf16(size_t n, const float* a, uint16_t* b, hipStream_t st);
int gram_kslab(int n, int nk);
size_t gram_workspace_bytes(int n, int nk);
void launch_gram(const float* A, int trans, int n, int nk, int lda, float* G, void* work,
                 hipStream_t st);

void launch_gemv_t(int rows, int cols, const float* A, int lda, size_t strideA, const double* x,
                   size_t stridex, double* y, size_t stridey, int batch, hipStream_t st);
size_t gemv_n_workspace(int rows, int cols, int batch);
// out = rhs - (A A^T x + shift x) in one pass over A (work: gemv_n_workspace bytes);
// applies for rows <= 1024 with rows, lda, strideA multiples of 4 and A 16-byte aligned
bool gemv_tn_applies(int rows, int lda, size_t strideA, const float* A);
void launch_gemv_tn(int rows, int cols, const float* A, int lda, size_t strideA, const double* x,
                    size_t stridex, double shift, const double* rhs, size_t striderhs, double* out,
                    size_t strideout, int batch, double* work, hipStream_t st);
// y_b += dy_b; ratio[b] = max|dy_b| / max|y_b| for the batch's n-vectors
void launch_add_to_ratio(int n, int batch, double* y, const double* dy, double* ratio, hipStream_t st);
void launch_gemv_n(int rows, int cols, const float* A, int lda, size_t strideA, const double* x,
                   size_t stridex, double shift, const double* z, size_t stridez, const double* rhs,
                   size_t striderhs, double* out, size_t strideout, int batch, double* work,
                   hipStream_t st);

// device int2 entries launch_cholesky needs for its update list (filled on the device)
size_t cholesky_pairs(int n);
void launch_cholesky(int n, const float* G, size_t strideG, double shift, double* A, size_t strideA,
                     int batch, int* d_info, int2* d_pairs, hipStream_t st);
void launch_potrs(int n, const double* A, size_t strideA, const double* b, size_t strideb, double* y,
                  size_t stridey, int batch, hipStream_t st);
// complex symmetric (not Hermitian) A = L L^T in place and its solves, FP64; A holds
// both triangles of the already shifted matrices; d_pairs as for launch_cholesky
void launch_cholesky_z(int n, zd* A, size_t strideA, int batch, int* d_info, int2* d_pairs, hipStream_t st);
void launch_potrs_z(int n, const zd* A, size_t strideA, const zd* b, size_t strideb, zd* y, size_t stridey,
                    int batch, hipStream_t st);

void launch_sub(int n, const double* a, const double* b, double* out, hipStream_t st);
void launch_add_to(int n, double* y, const double* dy, hipStream_t st);
void launch_gamma_update(int n, float* g, const double* d, hipStream_t st);
// out[b][i] = a[b][i] - u[i]
void launch_sub_bcast(int n, int batch, const double* a, const double* u, double* out, hipStream_t st);
// d[e] = sum_i A[i + e lda]^2
void launch_col_sq(int rows, int cols, const float* A, int lda, double* d, hipStream_t st);
// A_b[:, e] *= w[e]
void launch_scale_cols(int rows, int cols, float* A, int lda, size_t strideA, int batch, const float* w,
                       hipStream_t st);
// a[i] *= b[i], i < n
void launch_mul(size_t n, float* a, const float* b, hipStream_t st);
// E-space CG pieces of the laplacian-regularised step (batch columns of n entries)
// out[b] = u[b] + c (x[b] + B x[b]), B[e] = sum over k < 3 of val[e][k] x[nb[e][k]] (nb < 0: none)
void launch_lap_combine(int n, int batch, const double* u, double c, const double* x, const int32_t* nb,
                        const double* val, double* out, hipStream_t st);
// z = a (r - u)
void launch_axmy(size_t n, double a, const double* r, const double* u, double* z, hipStream_t st);
// out[b * ostride] = a[b] . c[b] (fixed-order FP64 sums)
void launch_dots(int n, int batch, const double* a, const double* c, double* out, int ostride, hipStream_t st);
// x[b] += alpha[b] p[b], r[b] -= alpha[b] q[b]
void launch_cg_xr(int n, int batch, const double* alpha, double* x, const double* p, double* r,
                  const double* q, hipStream_t st);
// p[b] = z[b] + beta[b] p[b]
void launch_cg_p(int n, int batch, const double* beta, double* p, const double* z, hipStream_t st);
// v[b][e] *= w[e]
void launch_scale_vec(int n, int batch, double* v, const float* w, hipStream_t st);
// g[b][e] = g0[e] + d[b][e]
void launch_gamma_from(int n, int batch, const float* g0, const double* d, float* g, hipStream_t st);

}  // namespace mpf
