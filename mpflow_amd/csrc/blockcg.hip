// Block preconditioned conjugate gradients: one Krylov space shared by all right-hand
// sides of one matrix.  ConjugateGradient::solve (src/numeric/conjugate_gradient.cpp:
// 61-161) iterates the columns of B independently, with per-column scalars; here the
// columns share their search directions, so every column is minimised over the sum of
// all columns' Krylov spaces (O'Leary's block CG).  No counterpart in the reference.
//
// Iteration, with the Jacobi preconditioner M = D^-1 (preconditioner.cpp:59-76):
//   R = B, Z = M R, P = orth(Z)
//   repeat:  Q = A P;  alpha = (P^T Q)^-1 P^T R;  X += P alpha;  R -= Q alpha
//            stop once every column meets the stop rule
//            Z = M R;  beta = -(P^T Q)^-1 Q^T Z;  P = orth(Z + P beta)
// orth(W) = W T: W^T W = L L^T by a Cholesky factorisation with pivoting on the column
// that keeps the largest fraction of its norm; columns whose fraction falls below
// kBcgDrop are dropped (linearly dependent directions: the drive patterns of an EIT model
// sum to zero, so their block is rank deficient from the start), and T = L^-T on the kept
// columns.  The kept directions occupy lanes 0 .. k-1 of P; lanes k .. Rp-1 are zero.
// With dcFree the operator is A + 1 1^T: its rank-one part is carried in the small
// products through the column sums of P and Z instead of being added to Q.
//
// Vectors are FP32 [np][Rp] (row-interleaved lanes, as the Krylov driver's); the Gram
// matrices and the small dense algebra are FP64.  Every product over the rows is summed
// in a fixed order (per-chunk partials, then the chunks in order), so a solve is
// deterministic.  Per iteration: the SpMV, four row kernels (the three that form Gram
// matrices write per-block partials) and three reductions whose last block runs the
// dense step.
#include <cfloat>
#include <cstring>

#include "blockcg.h"
#include "numeric.h"

namespace mpf {

constexpr int kBcgThreads = 256;
constexpr int kBcgChunk = 64;           // rows per chunk of the row kernels
constexpr int kBcgMaxParts = 256;       // blocks of the row kernels (fixed partition of the chunks)
constexpr int kBcgNG = 2;               // Gram matrices per row kernel (at most)
constexpr int kBcgNV = 3;               // column vectors per row kernel: r.r, r.z, 1^T z (or 1^T w)
constexpr double kBcgDrop = 1e-10;      // squared-norm fraction below which a direction is dropped

enum { BF_DONE = 0, BF_BROKE, BF_ITER, BF_K, BF_COUNT, BF_CAPPED, BF_NFLAGS };
enum { RM_INIT = 0, RM_GRAM, RM_UPD1, RM_UPD2, RM_UPD3 };
enum { DM_ORTH0 = 0, DM_ALPHA, DM_BETA, DM_ORTH };

struct BcgDev {
    int n, np, R, dc, rule, max_iter, nparts, cpb;
    double tol2;
    const float* dinv;
    const float* rhs;
    double *x, *r, *z;      // FP64 [np][Rp]; z also holds W
    float *p, *q;           // FP32 [np][Rp]
    const double* bb_ref;   // null: the start residual's column norms
    double* gpart;          // [nparts][kBcgNG][Rp * Rp]
    double* vpart;          // [nparts][kBcgNV][Rp]
    double* gsum;           // [kBcgNG][Rp * Rp]
    double* vsum;           // [kBcgNV][Rp]
    double* L;              // [Rp * Rp] Cholesky factor of P^T A' P (lower, row-major)
    double* cP;             // [Rp] column sums of P
    double* bb;             // [Rp] b.b per column
    double* mat;            // [Rp * Rp] alpha, beta or T for the next row kernel
    double* d;              // [Rp] (1^T P) alpha: the rank-one part of Q alpha in every row
    int* flags;             // BF_*
};

// ---------------------------------------------------------------------------
// row kernels
// ---------------------------------------------------------------------------

// tile [kBcgChunk][Rp + 1] <- rows row0 .. row0 + 63 of a [np][Rp] vector (128-bit loads)
template <int Rp, class T>
__device__ inline void bcg_stage(T* tile, const T* __restrict__ src, int row0) {
    constexpr int V = 16 / sizeof(T), CV = Rp / V, RS = Rp + 1;
    const uint4* s4 = reinterpret_cast<const uint4*>(src + (size_t)row0 * Rp);
    for (int e = threadIdx.x; e < kBcgChunk * CV; e += kBcgThreads) {
        const uint4 raw = s4[e];
        T v[V];
        __builtin_memcpy(v, &raw, sizeof(raw));
        T* o = tile + (e / CV) * RS + V * (e % CV);
#pragma unroll
        for (int k = 0; k < V; ++k) o[k] = v[k];
    }
}

// g[a][b] += sum over the chunk's rows of U[row][gi*BI + a] * V[row][gj*BI + b] (FP64)
template <int Rp, class TU, class TV>
__device__ inline void bcg_gram(double (&g)[Rp / 16][Rp / 16], const TU* U, const TV* V, int gi, int gj) {
    constexpr int BI = Rp / 16, RS = Rp + 1;
    for (int row = 0; row < kBcgChunk; ++row) {
        double u[BI], v[BI];
#pragma unroll
        for (int a = 0; a < BI; ++a) u[a] = (double)U[row * RS + gi * BI + a];
#pragma unroll
        for (int b = 0; b < BI; ++b) v[b] = (double)V[row * RS + gj * BI + b];
#pragma unroll
        for (int a = 0; a < BI; ++a)
#pragma unroll
            for (int b = 0; b < BI; ++b) g[a][b] = fma(u[a], v[b], g[a][b]);
    }
}

// Block b runs chunks [b*cpb, min((b+1)*cpb, np/64)).  Thread (j, rg) = (tid % Rp, tid / Rp)
// owns lane j of rows rg, rg + 256/Rp, ...; thread (gi, gj) = (tid / 16, tid % 16) owns
// the Gram entries (gi*BI + a, gj*BI + b).  P and Q are FP32 (the SpMV's operands), X, R
// and Z (which also holds W) FP64.
//   RM_INIT: x = 0, r = rhs, z = M r;             G0 = Z^T Z;  v = r.r, r.z, 1^T z
//   RM_GRAM: G0 = P^T Q, G1 = P^T R;              v[2] = 1^T p
//   RM_UPD1: x += P alpha, r -= Q alpha + d, z = M r;  G0 = Q^T Z;  v = r.r, r.z, 1^T z
//   RM_UPD2: z <- w = z + P beta;                  G0 = W^T W;  v[2] = 1^T w
//   RM_UPD3: p = W T
template <int LG, int MODE>
__global__ void __launch_bounds__(kBcgThreads) k_bcg_rows(BcgDev D) {
    constexpr int Rp = 1 << LG, RS = Rp + 1, RG = kBcgThreads / Rp, RPT = kBcgChunk / RG, BI = Rp / 16;
    constexpr bool GRAM = MODE != RM_UPD3;
    static_assert(Rp >= 16 && Rp <= 64, "16 to 64 lanes");
    if (D.flags[BF_DONE]) return;
    __shared__ float ta[kBcgChunk * RS], tb[kBcgChunk * RS];
    __shared__ double tc[kBcgChunk * RS];
    __shared__ double red[kBcgNV][RG][Rp];
    // The row products P beta and W T for all of a thread's rows of a chunk, the matrix
    // index outermost (each entry of column j read once per chunk); P alpha and Q alpha
    // row by row from alpha in LDS (the x and r loads of a row then overlap the next row's
    // products).  alpha and beta in FP64 (block CG's orthogonality relations need the
    // exact step), T rounded to FP32 (any combination of the W columns is a valid
    // direction block).
    constexpr bool PROD = MODE == RM_UPD2 || MODE == RM_UPD3;
    __shared__ double ms[MODE == RM_UPD1 ? Rp * Rp : 1];
    const int j = threadIdx.x % Rp, rg = threadIdx.x / Rp;
    const int gi = threadIdx.x / 16, gj = threadIdx.x % 16;
    double acc1[PROD ? RPT : 1];
    if constexpr (MODE == RM_UPD1) {
        for (int e = threadIdx.x; e < Rp * Rp; e += kBcgThreads) ms[e] = D.mat[e];
    }
    const double dj = (MODE == RM_UPD1) ? D.d[j] : 0.0;
    double g0[BI][BI], g1[BI][BI];
#pragma unroll
    for (int a = 0; a < BI; ++a)
#pragma unroll
        for (int b = 0; b < BI; ++b) g0[a][b] = g1[a][b] = 0.0;
    double v0 = 0.0, v1 = 0.0, v2 = 0.0;
    const int nchunks = D.np / kBcgChunk;
    const int c0 = blockIdx.x * D.cpb, c1 = min(c0 + D.cpb, nchunks);
    for (int c = c0; c < c1; ++c) {
        const int row0 = c * kBcgChunk;
        __syncthreads();   // the previous chunk's tiles have been read
        if constexpr (MODE == RM_GRAM) {
            bcg_stage<Rp>(ta, D.p, row0);
            bcg_stage<Rp>(tb, D.q, row0);
            bcg_stage<Rp>(tc, D.r, row0);
        } else if constexpr (MODE == RM_UPD1) {
            bcg_stage<Rp>(ta, D.p, row0);
            bcg_stage<Rp>(tb, D.q, row0);
        } else if constexpr (MODE == RM_UPD2) {
            bcg_stage<Rp>(ta, D.p, row0);
        } else if constexpr (MODE == RM_UPD3) {
            bcg_stage<Rp>(tc, D.z, row0);
        }
        __syncthreads();
        if constexpr (PROD) {
#pragma unroll
            for (int t = 0; t < RPT; ++t) acc1[t] = 0.0;
#pragma unroll 4
            for (int i = 0; i < Rp; ++i) {
                const double m = MODE == RM_UPD3 ? (double)(float)D.mat[i * Rp + j] : D.mat[i * Rp + j];
#pragma unroll
                for (int t = 0; t < RPT; ++t) {
                    const int lr = rg + RG * t;
                    if constexpr (MODE == RM_UPD3) acc1[t] = fma(tc[lr * RS + i], m, acc1[t]);
                    else acc1[t] = fma((double)ta[lr * RS + i], m, acc1[t]);
                }
            }
        }
#pragma unroll(MODE == RM_UPD1 ? 2 : RPT)
        for (int t = 0; t < RPT; ++t) {
            const int lr = rg + RG * t, row = row0 + lr;
            const size_t e = (size_t)row * Rp + j;
            if constexpr (MODE == RM_INIT) {
                const double r = (double)D.rhs[e];
                const double z = (double)D.dinv[row] * r;
                D.x[e] = 0.0;
                D.r[e] = r;
                D.z[e] = z;
                tc[lr * RS + j] = z;
                v0 += r * r; v1 += r * z; v2 += z;
            } else if constexpr (MODE == RM_GRAM) {
                v2 += (double)ta[lr * RS + j];
            } else if constexpr (MODE == RM_UPD1) {
                double px = 0.0, qa = 0.0;
#pragma unroll
                for (int i = 0; i < Rp; ++i) {
                    const double m = ms[i * Rp + j];
                    px = fma((double)ta[lr * RS + i], m, px);
                    qa = fma((double)tb[lr * RS + i], m, qa);
                }
                const double r = D.r[e] - qa - (row < D.n ? dj : 0.0);
                const double z = (double)D.dinv[row] * r;
                D.x[e] += px;
                D.r[e] = r;
                D.z[e] = z;
                tc[lr * RS + j] = z;
                v0 += r * r; v1 += r * z; v2 += z;
            } else if constexpr (MODE == RM_UPD2) {
                const double w = D.z[e] + acc1[t];
                D.z[e] = w;
                tc[lr * RS + j] = w;
                v2 += w;
            } else if constexpr (MODE == RM_UPD3) {
                D.p[e] = (float)acc1[t];
            }
        }
        if constexpr (GRAM) {
            __syncthreads();
            if constexpr (MODE == RM_GRAM) {
                bcg_gram<Rp>(g0, ta, tb, gi, gj);
                bcg_gram<Rp>(g1, ta, tc, gi, gj);
            } else if constexpr (MODE == RM_UPD1) {
                bcg_gram<Rp>(g0, tb, tc, gi, gj);
            } else {
                bcg_gram<Rp>(g0, tc, tc, gi, gj);
            }
        }
    }
    if constexpr (GRAM) {
        double* gp = D.gpart + (size_t)blockIdx.x * kBcgNG * Rp * Rp;
#pragma unroll
        for (int a = 0; a < BI; ++a)
#pragma unroll
            for (int b = 0; b < BI; ++b) {
                const int i = gi * BI + a, jj = gj * BI + b;
                gp[i * Rp + jj] = g0[a][b];
                if constexpr (MODE == RM_GRAM) gp[Rp * Rp + i * Rp + jj] = g1[a][b];
            }
        red[0][rg][j] = v0;
        red[1][rg][j] = v1;
        red[2][rg][j] = v2;
        __syncthreads();
        if ((int)threadIdx.x < Rp) {
            double sv[kBcgNV] = {0.0, 0.0, 0.0};
            for (int g = 0; g < RG; ++g)
#pragma unroll
                for (int v = 0; v < kBcgNV; ++v) sv[v] += red[v][g][threadIdx.x];
            double* vp = D.vpart + (size_t)blockIdx.x * kBcgNV * Rp;
#pragma unroll
            for (int v = 0; v < kBcgNV; ++v) vp[v * Rp + threadIdx.x] = sv[v];
        }
    }
}

// x (FP32 [np][Rp]) = X
__global__ void k_bcg_export(size_t count, const double* __restrict__ X, float* __restrict__ x) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) x[i] = (float)X[i];
}

// ---------------------------------------------------------------------------
// reductions and the dense steps (the last block of the reduction)
// ---------------------------------------------------------------------------

// the stop rule on the column sums r.r, r.z (the reference rule, conjugate_gradient.cpp:
// 145-151: sqrt(|r.z| / b.b) < FLT_EPSILON on the first column)
template <int Rp>
__device__ inline bool bcg_converged(const BcgDev& D, const double* rr, const double* rz) {
    if (D.rule == MPF_STOP_REFERENCE) {
        const double bb = D.bb[0];
        return bb <= 0.0 || sqrt(fabs(rz[0]) / bb) < (double)FLT_EPSILON;
    }
    for (int j = 0; j < D.R; ++j)
        if (!(rr[j] <= D.tol2 * D.bb[j])) return false;
    return true;
}

// sum over the four consecutive lanes of a column group (every lane gets the same value)
__device__ inline double bcg_sum4(double s) {
    s += __shfl_xor(s, 1);
    return s + __shfl_xor(s, 2);
}

// sum over k = k0, k0 + 4, ... < k1 of U[k] * V[k] with four independent partial sums
// (the LDS loads of four terms in flight), combined in a fixed order
__device__ inline double bcg_dot4(const double* U, int us, const double* V, int vs, int k0, int k1) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int k = k0;
    for (; k + 12 < k1; k += 16) {
        s0 = fma(U[k * us], V[k * vs], s0);
        s1 = fma(U[(k + 4) * us], V[(k + 4) * vs], s1);
        s2 = fma(U[(k + 8) * us], V[(k + 8) * vs], s2);
        s3 = fma(U[(k + 12) * us], V[(k + 12) * vs], s3);
    }
    for (; k < k1; k += 4) s0 = fma(U[k * us], V[k * vs], s0);
    return (s0 + s1) + (s2 + s3);
}

// X (Rp x Rp, LDS, row stride LS) <- L^-T L^-1 X for the lower-triangular L (LDS, row
// stride LS).  Four lanes per column (tid / 4) split each dot product; all four write the
// same value, so each lane only reads entries it wrote itself.
template <int Rp, int LS>
__device__ inline void bcg_chol_solve(const double* L, double* X) {
    static_assert(4 * Rp <= kBcgThreads, "four lanes per column");
    const int c = threadIdx.x >> 2, part = threadIdx.x & 3;
    if (c < Rp) {
        for (int i = 0; i < Rp; ++i) {
            const double s = bcg_sum4(bcg_dot4(L + i * LS, 1, X + c, LS, part, i));
            X[i * LS + c] = (X[i * LS + c] - s) / L[i * LS + i];
        }
        for (int i = Rp - 1; i >= 0; --i) {
            const double s = bcg_sum4(bcg_dot4(L + i, LS, X + c, LS, i + 1 + part, Rp));
            X[i * LS + c] = (X[i * LS + c] - s) / L[i * LS + i];
        }
    }
    __syncthreads();
}

// The dense step of the last reduction block.  LDS matrices use the row stride Rp + 1
// (rows start on different banks).
template <int Rp, int MODE>
__device__ void bcg_dense(const BcgDev& D) {
    constexpr int LS = Rp + 1;
    __shared__ double A[Rp * LS], B[Rp * LS], F[Rp * LS];
    __shared__ double vec[kBcgNV][Rp], cp[Rp], lcol[Rp], orig[Rp];
    __shared__ int piv[Rp], used[Rp], sk, sstop;
    const int tid = threadIdx.x;
    constexpr int NPT = Rp * Rp / kBcgThreads;   // matrix entries per thread
    static_assert(Rp * Rp % kBcgThreads == 0, "whole matrix rows per block");
    // global loads issued together (one latency, not NPT)
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
        const int e = tid + r * kBcgThreads, i = e / Rp, c = e % Rp;
        A[i * LS + c] = D.gsum[e];
        if (MODE == DM_ALPHA) B[i * LS + c] = D.gsum[Rp * Rp + e];
    }
    for (int e = tid; e < kBcgNV * Rp; e += kBcgThreads) vec[e / Rp][e % Rp] = D.vsum[e];
    if (tid < Rp) cp[tid] = D.cP[tid];
    __syncthreads();
    const int k = D.flags[BF_K];
    if constexpr (MODE == DM_ALPHA) {
        // the column sums of the current P (1^T P, summed by the Gram kernel)
        if (tid < Rp) {
            cp[tid] = vec[2][tid];
            D.cP[tid] = vec[2][tid];
        }
        __syncthreads();
        // F = P^T Q symmetrised (+ cP cP^T), unit diagonal on the empty lanes
        for (int e = tid; e < Rp * Rp; e += kBcgThreads) {
            const int i = e / Rp, c = e % Rp;
            double v = 0.5 * (A[i * LS + c] + A[c * LS + i]);
            if (D.dc) v += cp[i] * cp[c];
            if (i >= k || c >= k) v = (i == c) ? 1.0 : 0.0;
            F[i * LS + c] = v;
        }
        __syncthreads();
        // Cholesky F = L L^T with one barrier per step: the trailing update uses the
        // unscaled column (F[i][c] -= F[i][m] F[c][m] / F[m][m]); the columns are scaled at
        // the end
        bool ok = true;
        for (int m = 0; m < Rp; ++m) {
            const double dmm = F[m * LS + m];
            if (!(dmm > 0.0)) { ok = false; break; }   // uniform: every thread reads the same value
            const double inv = 1.0 / dmm;
            const int nt = Rp - 1 - m;
#pragma unroll 4
            for (int e = tid; e < nt * nt; e += kBcgThreads) {
                const int i = m + 1 + e / nt, c = m + 1 + e % nt;
                if (c <= i) F[i * LS + c] -= F[i * LS + m] * F[c * LS + m] * inv;
            }
            __syncthreads();
        }
        if (!ok) {   // P^T A' P not positive definite: breakdown
            if (tid == 0) { D.flags[BF_BROKE] = 1; D.flags[BF_DONE] = 1; }
            return;
        }
        for (int e = tid; e < Rp * Rp; e += kBcgThreads) {
            const int i = e / Rp, c = e % Rp;
            A[i * LS + c] = c > i ? 0.0 : F[i * LS + c] / sqrt(F[c * LS + c]);
        }
        __syncthreads();
        bcg_chol_solve<Rp, LS>(A, B);   // alpha = (P^T A' P)^-1 P^T R
        for (int e = tid; e < Rp * Rp; e += kBcgThreads) {
            const int i = e / Rp, c = e % Rp;
            D.L[e] = A[i * LS + c];
            D.mat[e] = B[i * LS + c];
        }
        if (tid < Rp) {
            double s = 0.0;
            if (D.dc)
                for (int i = 0; i < Rp; ++i) s += cp[i] * B[i * LS + tid];
            D.d[tid] = s;
        }
    } else if constexpr (MODE == DM_BETA) {
        // stop test on the updated residual, then beta = -(P^T A' P)^-1 Q'^T Z
        if (tid == 0) {
            const int it = D.flags[BF_ITER] + 1;
            D.flags[BF_ITER] = it;
            sstop = 0;
            if (bcg_converged<Rp>(D, vec[0], vec[1])) {
                sstop = 1;
            } else if (it >= D.max_iter) {
                sstop = 1;
                D.flags[BF_CAPPED] = 1;
            }
            if (sstop) D.flags[BF_DONE] = 1;
        }
        __syncthreads();
        if (sstop) return;
#pragma unroll
        for (int r = 0; r < NPT; ++r) {
            const int e = tid + r * kBcgThreads, i = e / Rp, c = e % Rp;
            B[i * LS + c] = -(A[i * LS + c] + (D.dc ? cp[i] * vec[2][c] : 0.0));
            A[i * LS + c] = D.L[e];
        }
        __syncthreads();
        bcg_chol_solve<Rp, LS>(A, B);
        for (int e = tid; e < Rp * Rp; e += kBcgThreads) D.mat[e] = B[(e / Rp) * LS + e % Rp];
    } else {
        // DM_ORTH0 / DM_ORTH: A = W^T W
        if constexpr (MODE == DM_ORTH0) {
            if (tid < Rp) D.bb[tid] = D.bb_ref ? D.bb_ref[tid] : vec[0][tid];
            __syncthreads();
            if (tid == 0) {
                sstop = bcg_converged<Rp>(D, vec[0], vec[1]);
                if (sstop) D.flags[BF_DONE] = 1;
            }
            __syncthreads();
            if (sstop) return;
        }
        // pivoted Cholesky: orig keeps the original diagonal, A is updated in place as the
        // Schur complement of the remaining columns, row m of F is factor column m indexed
        // by the original column
        if (tid < Rp) {
            orig[tid] = A[tid * LS + tid];
            used[tid] = 0;
        }
        if (tid == 0) sk = 0;
        __syncthreads();
        for (int m = 0; m < Rp; ++m) {
            // the remaining column keeping the largest fraction of its norm (first of equal
            // ones), by a butterfly over the first wavefront
            if (tid < 64) {
                double f = -1.0;
                int idx = tid;
                if (tid < Rp && !used[tid] && orig[tid] > 0.0) f = A[tid * LS + tid] / orig[tid];
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    const double of = __shfl_xor(f, off);
                    const int oi = __shfl_xor(idx, off);
                    if (of > f || (of == f && oi < idx)) { f = of; idx = oi; }
                }
                if (tid == 0) {
                    sstop = !(f > kBcgDrop);
                    if (!sstop) {
                        piv[m] = idx;
                        used[idx] = 1;
                        sk = m + 1;
                    }
                }
            }
            __syncthreads();
            if (sstop) break;
            const int p = piv[m];
            const double lpp = sqrt(A[p * LS + p]);
            if (tid < Rp) lcol[tid] = used[tid] && tid != p ? 0.0 : A[tid * LS + p] / lpp;
            __syncthreads();
            // Schur complement of the remaining columns; factor column m
#pragma unroll 4
            for (int e = tid; e < Rp * Rp; e += kBcgThreads) {
                const int i = e / Rp, c = e % Rp;
                if (!used[i] && !used[c]) A[i * LS + c] -= lcol[i] * lcol[c];
            }
            if (tid < Rp) F[m * LS + tid] = lcol[tid];
            __syncthreads();
        }
        const int kk = sk;
        // L (kk x kk, pivot order) in A: L[a][m] = factor column m at original column piv[a]
        __syncthreads();
        for (int e = tid; e < Rp * Rp; e += kBcgThreads) {
            const int a = e / Rp, m = e % Rp;
            A[a * LS + m] = (a < kk && m < kk && m <= a) ? F[m * LS + piv[a]] : 0.0;
        }
        __syncthreads();
        // T = L^-T (upper triangular) in B: L^T T = I, four lanes per column as in
        // bcg_chol_solve
        if ((tid >> 2) < kk) {
            const int m = tid >> 2, part = tid & 3;
            for (int a = kk - 1; a >= 0; --a) {
                const double s = bcg_sum4(bcg_dot4(A + a, LS, B + m, LS, a + 1 + part, kk));
                B[a * LS + m] = ((a == m ? 1.0 : 0.0) - s) / A[a * LS + a];
            }
        }
        __syncthreads();
        // mat[i][m] = T[a][m] for i = piv[a]
        for (int e = tid; e < Rp * Rp; e += kBcgThreads) D.mat[e] = 0.0;
        __syncthreads();
        for (int e = tid; e < Rp * Rp; e += kBcgThreads) {
            const int a = e / Rp, m = e % Rp;
            if (a < kk && m < kk) D.mat[piv[a] * Rp + m] = B[a * LS + m];
        }
        if (tid == 0) {
            D.flags[BF_K] = kk;
            if (kk == 0) { D.flags[BF_BROKE] = 1; D.flags[BF_DONE] = 1; }
        }
    }
}

// gsum / vsum = the partials summed over the row kernel's blocks in block order; the
// last block to finish runs the dense step
template <int LG, int MODE>
__global__ void __launch_bounds__(kBcgThreads) k_bcg_reduce(BcgDev D) {
    constexpr int Rp = 1 << LG;
    constexpr int NG = MODE == DM_ALPHA ? 2 : 1;
    constexpr int NE = NG * Rp * Rp + kBcgNV * Rp;
    if (D.flags[BF_DONE]) return;
    const int e = blockIdx.x * kBcgThreads + threadIdx.x;
    if (e < NE) {
        // the partials in block order, 8 loads in flight
        const bool gram = e < NG * Rp * Rp;
        const double* src = gram ? D.gpart + e : D.vpart + (e - NG * Rp * Rp);
        const size_t stride = gram ? (size_t)kBcgNG * Rp * Rp : (size_t)kBcgNV * Rp;
        double s = 0.0;
        int b = 0;
        for (; b + 8 <= D.nparts; b += 8) {
            double t[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) t[u] = src[(size_t)(b + u) * stride];
#pragma unroll
            for (int u = 0; u < 8; ++u) s += t[u];
        }
        for (; b < D.nparts; ++b) s += src[(size_t)b * stride];
        if (gram) D.gsum[e] = s;
        else D.vsum[e - NG * Rp * Rp] = s;
    }
    __threadfence();
    __syncthreads();
    __shared__ int last;
    if (threadIdx.x == 0) last = atomicAdd(&D.flags[BF_COUNT], 1) == (int)gridDim.x - 1;
    __syncthreads();
    if (!last) return;
    __threadfence();
    if (threadIdx.x == 0) D.flags[BF_COUNT] = 0;
    bcg_dense<Rp, MODE>(D);
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------

template <int LG>
static void bcg_rows(int mode, const BcgDev& D, hipStream_t st) {
    const dim3 g(D.nparts), b(kBcgThreads);
    switch (mode) {
        case RM_INIT: k_bcg_rows<LG, RM_INIT><<<g, b, 0, st>>>(D); break;
        case RM_GRAM: k_bcg_rows<LG, RM_GRAM><<<g, b, 0, st>>>(D); break;
        case RM_UPD1: k_bcg_rows<LG, RM_UPD1><<<g, b, 0, st>>>(D); break;
        case RM_UPD2: k_bcg_rows<LG, RM_UPD2><<<g, b, 0, st>>>(D); break;
        default: k_bcg_rows<LG, RM_UPD3><<<g, b, 0, st>>>(D); break;
    }
}

template <int LG>
static void bcg_reduce(int mode, const BcgDev& D, hipStream_t st) {
    constexpr int Rp = 1 << LG;
    const int ng = mode == DM_ALPHA ? 2 : 1;
    const dim3 g((ng * Rp * Rp + kBcgNV * Rp + kBcgThreads - 1) / kBcgThreads), b(kBcgThreads);
    switch (mode) {
        case DM_ORTH0: k_bcg_reduce<LG, DM_ORTH0><<<g, b, 0, st>>>(D); break;
        case DM_ALPHA: k_bcg_reduce<LG, DM_ALPHA><<<g, b, 0, st>>>(D); break;
        case DM_BETA: k_bcg_reduce<LG, DM_BETA><<<g, b, 0, st>>>(D); break;
        default: k_bcg_reduce<LG, DM_ORTH><<<g, b, 0, st>>>(D); break;
    }
}

template <int LG>
static void bcg_iteration(const BcgDev& D, const BcgArgs& a, hipStream_t st) {
    ell_spmv_product(a.ids, a.s, a.zero, a.np, a.wpad, 1 << LG, D.p, D.q, a.bad, st);
    bcg_rows<LG>(RM_GRAM, D, st);
    bcg_reduce<LG>(DM_ALPHA, D, st);
    bcg_rows<LG>(RM_UPD1, D, st);
    bcg_reduce<LG>(DM_BETA, D, st);
    bcg_rows<LG>(RM_UPD2, D, st);
    bcg_reduce<LG>(DM_ORTH, D, st);
    bcg_rows<LG>(RM_UPD3, D, st);
}

template <int LG>
static BcgResult bcg_run(const BcgArgs& a, int rule, double tol, int max_iter, int check_every, hipStream_t st) {
    constexpr int Rp = 1 << LG;
    MPF_REQUIRE(a.np % kTileRows == 0 && a.R >= 1 && a.R <= Rp, "block PCG: bad shape");
    const int nchunks = a.np / kBcgChunk;
    const int cpb = (nchunks + kBcgMaxParts - 1) / kBcgMaxParts;
    const int nparts = (nchunks + cpb - 1) / cpb;
    DBuf<double> gpart((size_t)nparts * kBcgNG * Rp * Rp), vpart((size_t)nparts * kBcgNV * Rp);
    DBuf<double> small((size_t)kBcgNG * Rp * Rp + kBcgNV * Rp + 2 * Rp * Rp + 3 * Rp);
    DBuf<int> flags(BF_NFLAGS);
    const size_t V = (size_t)a.np * Rp;
    DBuf<double> X(V), Rv(V), Z(V);
    BcgDev D;
    std::memset(&D, 0, sizeof(D));
    D.n = a.n; D.np = a.np; D.R = a.R; D.dc = a.dc; D.rule = rule; D.max_iter = max_iter;
    D.nparts = nparts; D.cpb = cpb; D.tol2 = tol * tol;
    D.dinv = a.dinv; D.rhs = a.rhs; D.x = X.p; D.r = Rv.p; D.z = Z.p; D.p = a.p; D.q = a.q;
    D.bb_ref = a.bb_ref;
    D.gpart = gpart.p; D.vpart = vpart.p;
    D.gsum = small.p;
    D.vsum = D.gsum + kBcgNG * Rp * Rp;
    D.L = D.vsum + kBcgNV * Rp;
    D.cP = D.L + Rp * Rp;
    D.bb = D.cP + Rp;
    D.d = D.bb + Rp;
    D.mat = D.d + Rp;
    D.flags = flags.p;
    MPF_HIP(hipMemsetAsync(flags.p, 0, flags.bytes(), st));
    MPF_HIP(hipMemsetAsync(small.p, 0, small.bytes(), st));
    bcg_rows<LG>(RM_INIT, D, st);
    bcg_reduce<LG>(DM_ORTH0, D, st);
    bcg_rows<LG>(RM_UPD3, D, st);
    MPF_HIP(hipGetLastError());
    const int every = check_every > 0 ? check_every : 16;
    int h[BF_NFLAGS] = {};
    // the device stops at max_iter; the host bound only guards against a stop that never comes
    long long launched = 0;
    for (;;) {
        MPF_HIP(hipMemcpyAsync(h, flags.p, sizeof(h), hipMemcpyDeviceToHost, st));
        MPF_HIP(hipStreamSynchronize(st));
        if (h[BF_DONE]) break;
        if (launched > (long long)max_iter + every) {
            h[BF_CAPPED] = 1;
            break;
        }
        for (int i = 0; i < every; ++i) bcg_iteration<LG>(D, a, st);
        launched += every;
        MPF_HIP(hipGetLastError());
    }
    k_bcg_export<<<(unsigned)((V + 255) / 256), 256, 0, st>>>(V, X.p, a.x);
    MPF_HIP(hipGetLastError());
    BcgResult res;
    res.steps = h[BF_ITER];
    res.capped = h[BF_CAPPED] != 0;
    res.broke = h[BF_BROKE] != 0;
    res.rank = h[BF_K];
    return res;
}

BcgResult block_pcg(const BcgArgs& a, int rule, double tol, int max_iter, int check_every, hipStream_t st) {
    MPF_REQUIRE(rule == MPF_STOP_REFERENCE || rule == MPF_STOP_TOLERANCE, "block PCG: unknown stop rule");
    MPF_REQUIRE(max_iter >= 1, "block PCG: max_iter < 1");
    switch (a.Rp) {
        case 16: return bcg_run<4>(a, rule, tol, max_iter, check_every, st);
        case 32: return bcg_run<5>(a, rule, tol, max_iter, check_every, st);
        case 64: return bcg_run<6>(a, rule, tol, max_iter, check_every, st);
        default: throw Error(MPF_EINVAL, "block PCG: 16, 32 or 64 lanes");
    }
}

}  // namespace mpf
