// Dense kernels of the EIT path on CDNA4 (gfx950):
//   electrode voltages, Jacobian assembly from drive/measurement fields, the Gram
//   matrix of the Jacobian on MFMA (BF16 in, FP32 accumulate), FP64 matrix-vector
//   products and a blocked FP64 Cholesky for the Tikhonov normal equations.
#include <algorithm>
#include <cmath>

#include "dense.h"

namespace mpf {

// --------------------------------------------------------------------------------
// voltages U[d*Mp + m] = sum_c sum_n W[m,n] Phi_c[n, d]
// (EIT::applyMeasurementPattern, src/models/eit.cpp:220-245; W = Mpat^T B^T)
// --------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_voltages(const JobDev* __restrict__ jobs, MeasDev md,
                                                  int use64, float* __restrict__ U32,
                                                  double* __restrict__ U64) {
    const int j = blockIdx.x;
    const JobDev& J = jobs[j];
    const int mrows = md.Mp * md.Dp;
    for (int idx = threadIdx.x; idx < mrows; idx += blockDim.x) {
        const int d = idx / md.Mp, m = idx % md.Mp;
        double v = 0.0;
        if (d < md.D && m < md.M) {
            for (int s = 0; s < J.C; ++s)
                for (int q = md.w_off[m]; q < md.w_off[m + 1]; ++q) {
                    const size_t i = vec_index(J, s, md.w_node[q], md.colmap[d]);
                    // voltage sources measure the electrode currents A_c Phi
                    const double phi = J.cur ? J.cur[i] : (use64 ? J.x64[i] : (double)J.x[i]);
                    v += md.w_val[q] * md.colscale[d] * phi;
                }
        }
        if (U32) U32[(size_t)j * mrows + idx] = (float)v;
        if (U64) U64[(size_t)j * mrows + idx] = v;
    }
}

// --------------------------------------------------------------------------------
// Jacobian (calcJacobianKernel, src/fem/equation_kernel.cu:197-245 with the scaling of
// src/models/eit.cpp:180-182, 202-205):
//   J[d*Mp + m, e] = -sigma'_e sum_c w_c Phi_c[el, d]^T Kc_e Phi_c[el, D+m]
//   reference: Kc = K_e, w_c = height;   exact: Kc = K_e + alpha_c M_e, w_c = 1/beta_c
// One wavefront per (element, job), four per block, each independent of the others: the
// work of an element is a chain of dependent steps (node ids, gathered fields, K_e Phi,
// the column), so the kernel is latency-bound and wants as many chains in flight per CU
// as possible; the wavefronts synchronise only among their own lanes.  The column J[:, e]
// is written once, contiguously (and its BF16 copy for the Gram when asked).
// --------------------------------------------------------------------------------
constexpr int kJacMaxRows = 4096;   // Dp * Mp <= 64 * 64
constexpr int kJacWaves = 4;        // elements per 256-thread block

// LDS writes of a wavefront made visible to its own lanes
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// PER: column rows per lane (mrows <= 64 PER); FIXEDM: Mp divides 64, so a lane's
// measurement column is the same for all of its rows; K: nodes per element (3 or 6), a
// compile-time constant so that the per-node loops unroll and the lane's column of T
// stays in registers
template <int PER, bool FIXEDM, int K>
__global__ void __launch_bounds__(256) k_jacobian(const JobDev* __restrict__ jobs, JacDev jd,
                                                  int use64, float* __restrict__ Jout,
                                                  __bf16* __restrict__ Jbf) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int E = jd.E;
    // the grid's x extent is a multiple of 8, so block (x, frame) runs on XCD x % 8 for
    // every frame (blocks go to the XCDs round-robin in linear order)
    int pos = blockIdx.x * kJacWaves + wave;
    if (jd.order) {
        // give XCD x the x-th eighth of the locality order
        const int nb = gridDim.x;
        const int base = nb / 8, x = blockIdx.x % 8;
        pos = (x * base + blockIdx.x / 8) * kJacWaves + wave;
    }
    if (pos >= E) return;   // the whole wavefront leaves; no block barrier follows
    const int e = jd.order ? jd.order[pos] : pos;
    const int j = blockIdx.y;
    const JobDev& Jb = jobs[j];
    constexpr int k = K;
    const int D = jd.D, M = jd.M, Mp = jd.Mp, Dp = jd.Dp, R = D + M;
    const int mrows = Mp * Dp;
    __shared__ float phi_s[kJacWaves][6][64];
    __shared__ float T_s[kJacWaves][6][64];
    __shared__ float Kc_s[kJacWaves][36];
    __shared__ int nodes_s[kJacWaves][6];
    __shared__ int cmap_s[kJacWaves][64];
    __shared__ double cscale_s[kJacWaves][64];
    float (*phi)[64] = phi_s[wave];
    float (*T)[64] = T_s[wave];
    float* Kc = Kc_s[wave];
    int* nodes = nodes_s[wave];
    if (lane < k) nodes[lane] = jd.elements[(size_t)e * k + lane];
    // the column map in LDS: the field loads then depend on LDS only
    if (lane < R) {
        cmap_s[wave][lane] = jd.colmap[lane];
        cscale_s[wave][lane] = jd.colscale[lane];
    }
    float acc[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) acc[q] = 0.0f;
    const int m0 = lane % Mp, d0 = lane / Mp, dstep = 64 / Mp;
    wave_sync();
    for (int s = 0; s < Jb.C; ++s) {
        const float al = jd.exact ? jd.alpha[s] : 0.0f;
        const float w = jd.weight[s];
        for (int i = lane; i < k * k; i += 64)
            Kc[i] = jd.ek[(size_t)e * k * k + i] + al * jd.em[(size_t)e * k * k + i];
        for (int i = lane; i < k * R; i += 64) {
            const int a = i / R, c = i % R;
            const size_t vi = vec_index(Jb, s, nodes[a], cmap_s[wave][c]);
            const double sc = cscale_s[wave][c];
            phi[a][c] = use64 ? (float)(sc * Jb.x64[vi]) : (float)(sc * (double)Jb.x[vi]);
        }
        wave_sync();
        for (int i = lane; i < k * M; i += 64) {
            const int a = i / M, m = i % M;
            float t = 0.0f;
            for (int b = 0; b < k; ++b) t += Kc[a * k + b] * phi[b][D + m];
            T[a][m] = w * t;
        }
        wave_sync();
        if constexpr (FIXEDM) {
            // Mp divides 64: the lane's measurement column m is the same for every q, so
            // its k entries of T stay in registers and d = d0 + q (64 / Mp)
            float tm[K];
#pragma unroll
            for (int a = 0; a < K; ++a) tm[a] = m0 < M ? T[a][m0] : 0.0f;
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                const int d = d0 + q * dstep;
                if (d < D && m0 < M && lane + 64 * q < mrows) {
                    float v = 0.0f;
#pragma unroll
                    for (int a = 0; a < K; ++a) v += phi[a][d] * tm[a];
                    acc[q] += v;
                }
            }
        } else {
#pragma unroll
            for (int q = 0; q < PER; ++q) {
                const int idx = lane + 64 * q;
                const int d = idx / Mp, m = idx % Mp;
                if (idx < mrows && d < D && m < M) {
                    float v = 0.0f;
                    for (int a = 0; a < k; ++a) v += phi[a][d] * T[a][m];
                    acc[q] += v;
                }
            }
        }
        // every lane is done with this harmonic's phi and T before they are overwritten
        wave_sync();
    }
    const float g = jd.gamma[(size_t)j * jd.E + e];
    const double sd = jd.logarithmic ? (2.302585092994046 / 10.0) * jd.ref * exp(2.302585092994046 * (double)g / 10.0)
                                     : jd.ref;
    const float scale = (float)(-sd);
    float* col = Jout + ((size_t)j * jd.E + e) * mrows;
    __bf16* colb = Jbf ? Jbf + ((size_t)j * jd.E + e) * mrows : nullptr;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int idx = lane + 64 * q;
        if (idx < mrows) {
            const float v = scale * acc[q];
            col[idx] = v;
            // the Gram's BF16 operand, rounded once here instead of at every panel read
            if (colb) colb[idx] = (__bf16)v;
        }
    }
}

// --------------------------------------------------------------------------------
// Gram matrix on MFMA: G = A A^T (trans 0) or A^T A (trans 1), BF16 operands, FP32
// accumulation with v_mfma_f32_32x32x16_bf16.  Block = 4 waves = 64x64 tile of G,
// wave w owns the 32x32 sub-tile (w/2, w%2).  The contraction is split over
// gridDim.z slabs whose partial tiles are summed in slab order by k_gram_reduce.
// Only tiles with I <= J are computed; the reduction mirrors them.
// --------------------------------------------------------------------------------
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

constexpr int kGT = 64;      // G tile
constexpr int kGK = 32;      // contraction step staged in LDS
constexpr int kGLds = kGK + 8;

__device__ __forceinline__ float gram_load(const float* A, int trans, int lda, int nrow, int nk,
                                           int r, int kk) {
    if (r >= nrow || kk >= nk) return 0.0f;
    return trans ? A[(size_t)r * lda + kk] : A[r + (size_t)kk * lda];
}

__global__ void __launch_bounds__(256) k_gram_mfma(const float* __restrict__ A, int trans, int nrow,
                                                   int nk, int lda, const int2* __restrict__ tiles,
                                                   int kslab, float* __restrict__ part) {
    const int2 tile = tiles[blockIdx.x];
    const int I0 = tile.x * kGT, J0 = tile.y * kGT;
    const int k0 = blockIdx.z * kslab, k1 = min(nk, k0 + kslab);
    __shared__ __attribute__((aligned(16))) __bf16 sa[kGT][kGLds];
    __shared__ __attribute__((aligned(16))) __bf16 sb[kGT][kGLds];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
    f32x16_t acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
    for (int kb = k0; kb < k1; kb += kGK) {
        // stage 64 x 32 of each operand, converted to BF16 (round to nearest even)
        for (int i = tid; i < kGT * kGK; i += 256) {
            int r, kk;
            if (trans) { r = i / kGK; kk = i % kGK; }     // k contiguous in memory
            else       { r = i % kGT; kk = i / kGT; }     // rows contiguous in memory
            const int kg = kb + kk;
            const bool ok = kg < k1;
            sa[r][kk] = (__bf16)(ok ? gram_load(A, trans, lda, nrow, nk, I0 + r, kg) : 0.0f);
            sb[r][kk] = (__bf16)(ok ? gram_load(A, trans, lda, nrow, nk, J0 + r, kg) : 0.0f);
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < kGK; ks += 16) {
            const int kq = ks + (lane >> 5) * 8;
            const bf16x8_t a = *reinterpret_cast<const bf16x8_t*>(&sa[wi + (lane & 31)][kq]);
            const bf16x8_t b = *reinterpret_cast<const bf16x8_t*>(&sb[wj + (lane & 31)][kq]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // C layout: lane -> column (lane & 31); register v -> row 8*(v/4) + 4*(lane/32) + v%4
    float* out = part + ((size_t)blockIdx.z * gridDim.x + blockIdx.x) * kGT * kGT;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int row = wi + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        const int col = wj + (lane & 31);
        out[row * kGT + col] = acc[v];
    }
}

// G = A A^T for a column-major A (nrow x nk, lda; nrow and lda multiples of 4): the
// Gauss-Newton data-space Gram J J^T.  Block = 4 waves = one 128 x 128 tile (I <= J) of
// G over one contraction slab; wave w owns the 64 x 64 quarter (w/2, w%2) as 2 x 2
// v_mfma_f32_32x32x16_bf16 tiles.  Each 32-deep stage of both operand panels is read
// with 128-bit loads (a column of A holds consecutive rows), converted to BF16 and stored
// [row][k] in LDS; the next stage's loads are issued before the current stage's MFMAs
// (two LDS buffers).  Partial tiles go to part[batch][slab][tile]; k_gram_reduce128 sums
// them in slab order.  The (I, J) pair of a block is derived from its index.
constexpr int kGT2 = 128;
constexpr int kGK2 = 32;
constexpr int kGL2 = kGK2 + 8;   // LDS row stride (BF16): 16-byte aligned 8-element reads

__device__ __forceinline__ int2 upper_tile(int b, int nt) {
    int i = 0;
    while (b >= nt - i) { b -= nt - i; ++i; }
    return make_int2(i, i + b);
}

// LDS panel layout of the MFMA Gram kernels: rows of `ld` BF16 (16-byte aligned) with 16
// more bytes after every 8 rows.  The staging writes of a wavefront go to rows 4 or 8
// apart, which a plain row stride puts on the same LDS banks (every 16-byte-aligned
// stride is a multiple of 4 dwords, so rows 8 apart are 32 dwords apart: one bank); the
// extra 16 bytes per 8 rows spread them over 8 bank groups.  The MFMA operand reads (8
// consecutive rows per pass) stay conflict-free.
__device__ __forceinline__ int panel_off(int r, int ld) { return r * ld + (r >> 3) * 8; }
constexpr int panel_elems(int rows, int ld) { return rows * ld + (rows / 8) * 8; }

__global__ void __launch_bounds__(256) k_gram_nt(const float* __restrict__ A, size_t strideA, int nrow, int nk,
                                                 int lda, int nt, int kslab, float* __restrict__ part) {
    const int ntiles = nt * (nt + 1) / 2;
    const int2 tile = upper_tile(blockIdx.x, nt);
    const int I0 = tile.x * kGT2, J0 = tile.y * kGT2;
    const int k0 = blockIdx.z * kslab, k1 = min(nk, k0 + kslab);
    const float* Ab = A + blockIdx.y * strideA;
    __shared__ __attribute__((aligned(16))) __bf16 sa[2][panel_elems(kGT2, kGL2)];
    __shared__ __attribute__((aligned(16))) __bf16 sb[2][panel_elems(kGT2, kGL2)];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wi = (wave >> 1) * 64, wj = (wave & 1) * 64;
    f32x16_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.0f;
    // stage loads: each thread owns a 4-row x 4-k block of each operand panel (row quad
    // tid & 31, k quad tid >> 5): four 128-bit loads (one per k, consecutive rows), then
    // per row the 4 k values packed into one 64-bit LDS write
    float4 ra[4], rb[4];
    const int rq = (tid & 31) * 4, kq4 = (tid >> 5) * 4;
    auto load = [&](int kb) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int kg = kb + kq4 + u;
            const bool kin = kg < k1;
            ra[u] = (kin && I0 + rq < nrow) ? *reinterpret_cast<const float4*>(Ab + (size_t)kg * lda + I0 + rq)
                                            : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            rb[u] = (kin && J0 + rq < nrow) ? *reinterpret_cast<const float4*>(Ab + (size_t)kg * lda + J0 + rq)
                                            : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    };
    typedef __bf16 bf16x4_t __attribute__((ext_vector_type(4)));
    auto store = [&](int buf) {
        bf16x4_t a0 = {(__bf16)ra[0].x, (__bf16)ra[1].x, (__bf16)ra[2].x, (__bf16)ra[3].x};
        bf16x4_t a1 = {(__bf16)ra[0].y, (__bf16)ra[1].y, (__bf16)ra[2].y, (__bf16)ra[3].y};
        bf16x4_t a2 = {(__bf16)ra[0].z, (__bf16)ra[1].z, (__bf16)ra[2].z, (__bf16)ra[3].z};
        bf16x4_t a3 = {(__bf16)ra[0].w, (__bf16)ra[1].w, (__bf16)ra[2].w, (__bf16)ra[3].w};
        bf16x4_t b0 = {(__bf16)rb[0].x, (__bf16)rb[1].x, (__bf16)rb[2].x, (__bf16)rb[3].x};
        bf16x4_t b1 = {(__bf16)rb[0].y, (__bf16)rb[1].y, (__bf16)rb[2].y, (__bf16)rb[3].y};
        bf16x4_t b2 = {(__bf16)rb[0].z, (__bf16)rb[1].z, (__bf16)rb[2].z, (__bf16)rb[3].z};
        bf16x4_t b3 = {(__bf16)rb[0].w, (__bf16)rb[1].w, (__bf16)rb[2].w, (__bf16)rb[3].w};
        *reinterpret_cast<bf16x4_t*>(&sa[buf][panel_off(rq, kGL2) + kq4]) = a0;
        *reinterpret_cast<bf16x4_t*>(&sa[buf][panel_off(rq + 1, kGL2) + kq4]) = a1;
        *reinterpret_cast<bf16x4_t*>(&sa[buf][panel_off(rq + 2, kGL2) + kq4]) = a2;
        *reinterpret_cast<bf16x4_t*>(&sa[buf][panel_off(rq + 3, kGL2) + kq4]) = a3;
        *reinterpret_cast<bf16x4_t*>(&sb[buf][panel_off(rq, kGL2) + kq4]) = b0;
        *reinterpret_cast<bf16x4_t*>(&sb[buf][panel_off(rq + 1, kGL2) + kq4]) = b1;
        *reinterpret_cast<bf16x4_t*>(&sb[buf][panel_off(rq + 2, kGL2) + kq4]) = b2;
        *reinterpret_cast<bf16x4_t*>(&sb[buf][panel_off(rq + 3, kGL2) + kq4]) = b3;
    };
    const int nst = (k1 - k0 + kGK2 - 1) / kGK2;
    if (nst > 0) {
        load(k0);
        store(0);
    }
    __syncthreads();
    for (int st = 0; st < nst; ++st) {
        const int buf = st & 1;
        if (st + 1 < nst) load(k0 + (st + 1) * kGK2);
#pragma unroll
        for (int ks = 0; ks < kGK2; ks += 16) {
            const int kq = ks + (lane >> 5) * 8;
            bf16x8_t av[2], bv[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                av[a] = *reinterpret_cast<const bf16x8_t*>(&sa[buf][panel_off(wi + 32 * a + (lane & 31), kGL2) + kq]);
                bv[a] = *reinterpret_cast<const bf16x8_t*>(&sb[buf][panel_off(wj + 32 * a + (lane & 31), kGL2) + kq]);
            }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[a], bv[b], acc[a][b], 0, 0, 0);
        }
        if (st + 1 < nst) store(buf ^ 1);
        __syncthreads();
    }
    // C layout: lane -> column (lane & 31); register v -> row 8*(v/4) + 4*(lane/32) + v%4
    float* out = part + (((size_t)blockIdx.y * gridDim.z + blockIdx.z) * ntiles + blockIdx.x) * kGT2 * kGT2;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = wi + 32 * a + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                const int col = wj + 32 * b + (lane & 31);
                out[row * kGT2 + col] = acc[a][b][v];
            }
}

// G_b = A_b A_b^T from a BF16 copy of A (column-major n x nk, lda; the Jacobian's layout
// [element][row]), 256 x 256 tiles of the upper triangle per block of 8 wavefronts (each
// a 128 x 64 sub-tile: 4 x 2 v_mfma_f32_32x32x16_bf16 accumulators), contraction in
// stages of 32 staged through two LDS buffers: every tile reads its two 256-row panels
// once per slab, a quarter of the 64-row tiles' panel traffic and half of the FP32 source's
// bytes.  Partial tiles go to part[batch][slab][tile]; k_gram_reduce256 sums them in
// slab order.  Diagonal tiles stage one panel.
constexpr int kGT3 = 256;
constexpr int kGK3 = 32;
constexpr int kGL3 = kGK3 + 8;   // LDS row stride (BF16): 16-byte aligned rows, staggered banks

__global__ void __launch_bounds__(512) k_gram_bf16(const __bf16* __restrict__ A, size_t strideA, int nrow, int nk,
                                                   int lda, int nt, int kslab, float* __restrict__ part) {
    const int ntiles = nt * (nt + 1) / 2;
    const int2 tile = upper_tile(blockIdx.x, nt);
    const int I0 = tile.x * kGT3, J0 = tile.y * kGT3;
    const bool diag = tile.x == tile.y;
    const int k0 = blockIdx.z * kslab, k1 = min(nk, k0 + kslab);
    const __bf16* Ab = A + blockIdx.y * strideA;
    __shared__ __attribute__((aligned(16))) __bf16 sa[2][panel_elems(kGT3, kGL3)];
    __shared__ __attribute__((aligned(16))) __bf16 sb[2][panel_elems(kGT3, kGL3)];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wi = (wave >> 2) * 128, wj = (wave & 3) * 64;
    f32x16_t acc[4][2];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.0f;
    // staging: threads 0..255 stage panel I, 256..511 panel J; each thread owns 8
    // consecutive rows x 4 consecutive k: four 128-bit loads (one per k), then per row the
    // 4 k values as one 64-bit LDS write
    typedef __bf16 bf16x4_t __attribute__((ext_vector_type(4)));
    const int half = tid >> 8, t8 = tid & 255;
    const int ro = (t8 & 31) * 8, kq = (t8 >> 5) * 4;
    const int P0 = half ? J0 : I0;
    const bool active = !(half && diag);
    bf16x8_t r[4];
    auto load = [&](int kb) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int kg = kb + kq + u;
            bf16x8_t z;
#pragma unroll
            for (int q = 0; q < 8; ++q) z[q] = (__bf16)0.0f;
            r[u] = (active && kg < k1 && P0 + ro < nrow)
                       ? *reinterpret_cast<const bf16x8_t*>(Ab + (size_t)kg * lda + P0 + ro) : z;
        }
    };
    auto store = [&](int buf) {
        if (!active) return;
        __bf16* dst = half ? sb[buf] : sa[buf];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            bf16x4_t v = {r[0][q], r[1][q], r[2][q], r[3][q]};
            *reinterpret_cast<bf16x4_t*>(&dst[panel_off(ro + q, kGL3) + kq]) = v;
        }
    };
    const int nst = (k1 - k0 + kGK3 - 1) / kGK3;
    if (nst > 0) {
        load(k0);
        store(0);
    }
    __syncthreads();
    for (int st = 0; st < nst; ++st) {
        const int buf = st & 1;
        if (st + 1 < nst) load(k0 + (st + 1) * kGK3);
        const __bf16* pb = diag ? sa[buf] : sb[buf];
#pragma unroll
        for (int ks = 0; ks < kGK3; ks += 16) {
            const int kk = ks + (lane >> 5) * 8;
            bf16x8_t av[4], bv[2];
#pragma unroll
            for (int a = 0; a < 4; ++a)
                av[a] = *reinterpret_cast<const bf16x8_t*>(&sa[buf][panel_off(wi + 32 * a + (lane & 31), kGL3) + kk]);
#pragma unroll
            for (int b = 0; b < 2; ++b)
                bv[b] = *reinterpret_cast<const bf16x8_t*>(&pb[panel_off(wj + 32 * b + (lane & 31), kGL3) + kk]);
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[a], bv[b], acc[a][b], 0, 0, 0);
        }
        if (st + 1 < nst) store(buf ^ 1);
        __syncthreads();
    }
    // C layout: lane -> column (lane & 31); register v -> row 8*(v/4) + 4*(lane/32) + v%4
    float* out = part + (((size_t)blockIdx.y * gridDim.z + blockIdx.z) * ntiles + blockIdx.x) * kGT3 * kGT3;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = wi + 32 * a + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                const int col = wj + 32 * b + (lane & 31);
                out[row * kGT3 + col] = acc[a][b][v];
            }
}

__global__ void __launch_bounds__(256) k_gram_reduce256(const float* __restrict__ part, int nt, int nslab, int n,
                                                        float* __restrict__ G, size_t strideG) {
    const int ntiles = nt * (nt + 1) / 2;
    const int2 tile = upper_tile(blockIdx.x, nt);
    const float* pb = part + (size_t)blockIdx.z * nslab * ntiles * kGT3 * kGT3;
    float* Gb = G + blockIdx.z * strideG;
    for (int i = blockIdx.y * 256 + threadIdx.x; i < kGT3 * kGT3; i += 256 * gridDim.y) {
        float s = 0.0f;
        for (int z = 0; z < nslab; ++z) s += pb[((size_t)z * ntiles + blockIdx.x) * kGT3 * kGT3 + i];
        const int r = tile.x * kGT3 + i / kGT3, c = tile.y * kGT3 + i % kGT3;
        if (r < n && c < n && (tile.x != tile.y || r <= c)) {
            Gb[r + (size_t)c * n] = s;
            Gb[c + (size_t)r * n] = s;
        }
    }
}

__global__ void __launch_bounds__(256) k_gram_reduce128(const float* __restrict__ part, int nt, int nslab, int n,
                                                        float* __restrict__ G, size_t strideG) {
    const int ntiles = nt * (nt + 1) / 2;
    const int2 tile = upper_tile(blockIdx.x, nt);
    const float* pb = part + (size_t)blockIdx.z * nslab * ntiles * kGT2 * kGT2;
    float* Gb = G + blockIdx.z * strideG;
    for (int i = blockIdx.y * 256 + threadIdx.x; i < kGT2 * kGT2; i += 256 * gridDim.y) {
        float s = 0.0f;
        for (int z = 0; z < nslab; ++z) s += pb[((size_t)z * ntiles + blockIdx.x) * kGT2 * kGT2 + i];
        const int r = tile.x * kGT2 + i / kGT2, c = tile.y * kGT2 + i % kGT2;
        // on diagonal tiles the upper entry of each pair is written to both places
        if (r < n && c < n && (tile.x != tile.y || r <= c)) {
            Gb[r + (size_t)c * n] = s;
            Gb[c + (size_t)r * n] = s;
        }
    }
}

__global__ void __launch_bounds__(256) k_gram_reduce(const float* __restrict__ part, const int2* __restrict__ tiles,
                                                     int ntiles, int nslab, int n, float* __restrict__ G) {
    const int b = blockIdx.x;
    const int2 tile = tiles[b];
    for (int i = threadIdx.x; i < kGT * kGT; i += 256) {
        float s = 0.0f;
        for (int z = 0; z < nslab; ++z) s += part[((size_t)z * ntiles + b) * kGT * kGT + i];
        const int r = tile.x * kGT + i / kGT, c = tile.y * kGT + i % kGT;
        if (r < n && c < n && (tile.x != tile.y || r <= c)) {
            G[r + (size_t)c * n] = s;
            G[c + (size_t)r * n] = s;
        }
    }
}

// --------------------------------------------------------------------------------
// FP64-accumulating matrix-vector products with a column-major FP32 matrix
// --------------------------------------------------------------------------------
// y[c] = sum_r A[r + c*lda] x[r]: one wavefront per column
__global__ void __launch_bounds__(256) k_gemv_t(int rows, int cols, const float* __restrict__ A, int lda,
                                                size_t strideA, const double* __restrict__ x, size_t stridex,
                                                double* __restrict__ y, size_t stridey, double ybeta) {
    const int col = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int b = blockIdx.y;
    if (col >= cols) return;
    const float* a = A + b * strideA + (size_t)col * lda;
    const double* xv = x + b * stridex;
    double s = 0.0;
    for (int r = lane; r < rows; r += 64) s += (double)a[r] * xv[r];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) {
        double* yp = y + b * stridey + col;
        *yp = (ybeta != 0.0 ? ybeta * *yp : 0.0) + s;
    }
}

// k_gemv_t for rows % 4 == 0 (and lda, strideA multiples of 4): 128-bit loads, a
// wavefront reads 1 KiB of its column per instruction
__global__ void __launch_bounds__(256) k_gemv_t4(int rows, int cols, const float* __restrict__ A, int lda,
                                                 size_t strideA, const double* __restrict__ x, size_t stridex,
                                                 double* __restrict__ y, size_t stridey) {
    const int col = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int b = blockIdx.y;
    if (col >= cols) return;
    const float* a = A + b * strideA + (size_t)col * lda;
    const double* xv = x + b * stridex;
    double s = 0.0;
    for (int r = 4 * lane; r < rows; r += 256) {
        const float4 v = *(const float4*)(a + r);
        s += (double)v.x * xv[r] + (double)v.y * xv[r + 1] + (double)v.z * xv[r + 2] + (double)v.w * xv[r + 3];
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) y[b * stridey + col] = s;
}

// partial[slab][r] = sum_{c in slab} A[r + c*lda] x[c]
constexpr int kNSlab = 256;

// partial[slab][r] = sum_{c in slab} A[r + c*lda] (sum_r' A[r' + c*lda] x[r']): A (A^T x)
// in one pass over A (rows <= 1024, rows, lda and strideA multiples of 4).  A block owns
// the kNSlab columns of its slab, kTnCols at a time: thread t loads rows 4t..4t+3 of each
// column (128-bit), the block reduces the column dots in FP64, and the same registers
// then give the rows' partial sums, so every element of A is read once.
constexpr int kTnCols = 8;
__global__ void __launch_bounds__(256) k_gemv_tn_part(int rows, int cols, const float* __restrict__ A, int lda,
                                                      size_t strideA, const double* __restrict__ x,
                                                      size_t stridex, double* __restrict__ part) {
    const int slab = blockIdx.x, b = blockIdx.y;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r0 = 4 * threadIdx.x;
    const bool act = r0 < rows;
    const float* a = A + b * strideA;
    double xv[4] = {0.0, 0.0, 0.0, 0.0};
    if (act)
        for (int i = 0; i < 4; ++i) xv[i] = x[b * stridex + r0 + i];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    __shared__ double red[4][kTnCols];
    const int c0 = slab * kNSlab, c1 = min(cols, c0 + kNSlab);
    for (int cb = c0; cb < c1; cb += kTnCols) {
        float4 p[kTnCols];
#pragma unroll
        for (int j = 0; j < kTnCols; ++j)
            p[j] = (act && cb + j < c1) ? *(const float4*)(a + (size_t)(cb + j) * lda + r0)
                                        : make_float4(0.f, 0.f, 0.f, 0.f);
        double d[kTnCols];
#pragma unroll
        for (int j = 0; j < kTnCols; ++j) {
            d[j] = (double)p[j].x * xv[0] + (double)p[j].y * xv[1] + (double)p[j].z * xv[2] + (double)p[j].w * xv[3];
            for (int off = 32; off > 0; off >>= 1) d[j] += __shfl_xor(d[j], off, 64);
        }
        if (lane == 0)
#pragma unroll
            for (int j = 0; j < kTnCols; ++j) red[wave][j] = d[j];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kTnCols; ++j) {
            const double v = (red[0][j] + red[1][j]) + (red[2][j] + red[3][j]);
            acc[0] += (double)p[j].x * v; acc[1] += (double)p[j].y * v;
            acc[2] += (double)p[j].z * v; acc[3] += (double)p[j].w * v;
        }
        __syncthreads();   // red is rewritten by the next group of columns
    }
    if (act)
        for (int i = 0; i < 4; ++i) part[((size_t)b * gridDim.x + slab) * rows + r0 + i] = acc[i];
}
__global__ void __launch_bounds__(256) k_gemv_n_part(int rows, int cols, const float* __restrict__ A, int lda,
                                                     size_t strideA, const double* __restrict__ x, size_t stridex,
                                                     double* __restrict__ part) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    const int slab = blockIdx.y, b = blockIdx.z;
    const int c0 = slab * kNSlab, c1 = min(cols, c0 + kNSlab);
    if (r >= rows) return;
    const float* a = A + b * strideA;
    const double* xv = x + b * stridex;
    double s = 0.0;
    for (int c = c0; c < c1; ++c) s += (double)a[r + (size_t)c * lda] * xv[c];
    part[((size_t)b * gridDim.y + slab) * rows + r] = s;
}

// y = part-sum (+ shift * z) ; out = rhs - y when rhs given
__global__ void __launch_bounds__(256) k_gemv_n_reduce(int rows, int nslab, const double* __restrict__ part,
                                                       double shift, const double* __restrict__ z, size_t stridez,
                                                       const double* __restrict__ rhs, size_t striderhs,
                                                       double* __restrict__ out, size_t strideout) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    const int b = blockIdx.y;
    if (r >= rows) return;
    double s = 0.0;
    for (int q = 0; q < nslab; ++q) s += part[((size_t)b * nslab + q) * rows + r];
    if (z) s += shift * z[b * stridez + r];
    out[b * strideout + r] = rhs ? rhs[b * striderhs + r] - s : s;
}

// --------------------------------------------------------------------------------
// blocked FP64 Cholesky (lower), column-major n x n, batch over blockIdx.z
// --------------------------------------------------------------------------------
constexpr int kCB = 32;

// A64 = G + shift I (and zero the strictly upper triangle is not needed)
__global__ void __launch_bounds__(256) k_to64_shift(int n, const float* __restrict__ G, size_t strideG,
                                                    double shift, double* __restrict__ A, size_t strideA) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const int b = blockIdx.y;
    if (i >= (size_t)n * n) return;
    const int r = (int)(i % n), c = (int)(i / n);
    A[b * strideA + i] = (double)G[b * strideG + i] + (r == c ? shift : 0.0);
}

// The blocked L L^T kernels below are templates over the scalar T: double for the real
// symmetric positive definite normal equations, zd for the complex symmetric ones of the
// complex-conductivity model (the same formulas without conjugation).
//
// factor the diagonal block kb in place
template <class T>
__global__ void __launch_bounds__(256) k_potrf_diag(int n, T* __restrict__ A, size_t strideA, int kb,
                                                    int* __restrict__ info) {
    T* a = A + blockIdx.z * strideA;
    __shared__ T L[kCB][kCB + 1];
    const int o = kb * kCB, nb = min(kCB, n - o);
    for (int i = threadIdx.x; i < kCB * kCB; i += 256) {
        const int r = i % kCB, c = i / kCB;
        L[r][c] = (r < nb && c < nb) ? a[(o + r) + (size_t)(o + c) * n] : Scal<T>::zero();
    }
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
        if (threadIdx.x == 0) {
            T d = L[j][j];
            if (!Scal<T>::pivot_ok(d)) { atomicExch(info, 1); d = Scal<T>::tiny(); }
            L[j][j] = Scal<T>::sqrt_(d);
        }
        __syncthreads();
        const T djj = L[j][j];
        for (int r = j + 1 + threadIdx.x; r < nb; r += 256) L[r][j] /= djj;
        __syncthreads();
        // trailing update of the block: L[r][c] -= L[r][j] L[c][j], c in (j, r]
        for (int i = threadIdx.x; i < kCB * kCB; i += 256) {
            const int r = i % kCB, c = i / kCB;
            if (c > j && r >= c && r < nb) L[r][c] -= L[r][j] * L[c][j];
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < kCB * kCB; i += 256) {
        const int r = i % kCB, c = i / kCB;
        if (r < nb && c < nb && r >= c) a[(o + r) + (size_t)(o + c) * n] = L[r][c];
    }
}

// panel: A[ib, kb] <- A[ib, kb] L_kk^-T for every row block ib > kb (blockIdx.x)
template <class T>
__global__ void __launch_bounds__(256) k_potrf_panel(int n, T* __restrict__ A, size_t strideA, int kb) {
    T* a = A + blockIdx.z * strideA;
    const int o = kb * kCB;
    const int ib = kb + 1 + blockIdx.x;
    const int ro = ib * kCB;
    if (ro >= n) return;
    const int nr = min(kCB, n - ro), nb = min(kCB, n - o);
    __shared__ T L[kCB][kCB + 1];
    __shared__ T X[kCB][kCB + 1];
    for (int i = threadIdx.x; i < kCB * kCB; i += 256) {
        const int r = i % kCB, c = i / kCB;
        L[r][c] = (r < nb && c < nb) ? a[(o + r) + (size_t)(o + c) * n] : Scal<T>::zero();
        X[r][c] = (r < nr && c < nb) ? a[(ro + r) + (size_t)(o + c) * n] : Scal<T>::zero();
    }
    __syncthreads();
    // each row r of X solves x L^T = a  (forward substitution over columns)
    if (threadIdx.x < nr) {
        const int r = threadIdx.x;
        for (int c = 0; c < nb; ++c) {
            T s = X[r][c];
            for (int t = 0; t < c; ++t) s -= X[r][t] * L[c][t];
            X[r][c] = s / L[c][c];
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kCB * kCB; i += 256) {
        const int r = i % kCB, c = i / kCB;
        if (r < nr && c < nb) a[(ro + r) + (size_t)(o + c) * n] = X[r][c];
    }
}

// trailing update: A[ib, jb] -= L[ib, kb] L[jb, kb]^T for kb < jb <= ib
template <class T>
__global__ void __launch_bounds__(256) k_potrf_update(int n, T* __restrict__ A, size_t strideA, int kb,
                                                      const int2* __restrict__ pairs) {
    T* a = A + blockIdx.z * strideA;
    const int2 pr = pairs[blockIdx.x];
    const int ib = pr.x, jb = pr.y;
    const int o = kb * kCB, ro = ib * kCB, co = jb * kCB;
    const int nb = min(kCB, n - o), nr = min(kCB, n - ro), nc = min(kCB, n - co);
    __shared__ T Li[kCB][kCB + 1];
    __shared__ T Lj[kCB][kCB + 1];
    for (int i = threadIdx.x; i < kCB * kCB; i += 256) {
        const int r = i % kCB, c = i / kCB;
        Li[r][c] = (r < nr && c < nb) ? a[(ro + r) + (size_t)(o + c) * n] : Scal<T>::zero();
        Lj[r][c] = (r < nc && c < nb) ? a[(co + r) + (size_t)(o + c) * n] : Scal<T>::zero();
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kCB * kCB; i += 256) {
        const int r = i % kCB, c = i / kCB;
        if (r < nr && c < nc && (ib != jb || r >= c)) {
            T s = Scal<T>::zero();
            for (int t = 0; t < nb; ++t) s += Li[r][t] * Lj[c][t];
            a[(ro + r) + (size_t)(co + c) * n] -= s;
        }
    }
}

// solve L L^T y = b (one block per batch entry), blocked substitution.  Each 32x32
// diagonal block is staged in LDS and solved by wavefront 0 with shuffles; the forward
// update reads rows of L coalesced across threads, the backward update lets each
// wavefront read contiguous column segments of L and reduce them with shuffles.
constexpr int kTrsvThreads = 1024;

template <class T>
__global__ void __launch_bounds__(kTrsvThreads) k_potrs(int n, const T* __restrict__ A, size_t strideA,
                                                        const T* __restrict__ b, size_t strideb,
                                                        T* __restrict__ y, size_t stridey) {
    const T* a = A + blockIdx.x * strideA;
    T* yv = y + blockIdx.x * stridey;
    const T* bv = b + blockIdx.x * strideb;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwave = kTrsvThreads / 64;
    __shared__ T Lb[kCB][kCB + 1];
    __shared__ T z[kCB];
    for (int i = tid; i < n; i += kTrsvThreads) yv[i] = bv[i];
    __syncthreads();
    const int nblk = (n + kCB - 1) / kCB;
    // forward: L z = b
    for (int kb = 0; kb < nblk; ++kb) {
        const int o = kb * kCB, nb = min(kCB, n - o);
        for (int i = tid; i < kCB * kCB; i += kTrsvThreads) {
            const int r = i % kCB, c = i / kCB;
            Lb[r][c] = (r < nb && c < nb) ? a[(o + r) + (size_t)(o + c) * n] : (r == c ? Scal<T>::one() : Scal<T>::zero());
        }
        __syncthreads();
        if (wave == 0) {
            T v = lane < nb ? yv[o + lane] : Scal<T>::zero();
            for (int r = 0; r < nb; ++r) {
                const T zr = Scal<T>::shfl(v, r) / Lb[r][r];
                if (lane == r) v = zr;
                else if (lane > r && lane < nb) v -= Lb[lane][r] * zr;
            }
            if (lane < nb) { z[lane] = v; yv[o + lane] = v; }
        }
        __syncthreads();
        for (int r = o + nb + tid; r < n; r += kTrsvThreads) {
            T s = Scal<T>::zero();
            for (int t = 0; t < nb; ++t) s += a[r + (size_t)(o + t) * n] * z[t];
            yv[r] -= s;
        }
        __syncthreads();
    }
    // backward: L^T y = z
    for (int kb = nblk - 1; kb >= 0; --kb) {
        const int o = kb * kCB, nb = min(kCB, n - o);
        for (int i = tid; i < kCB * kCB; i += kTrsvThreads) {
            const int r = i % kCB, c = i / kCB;
            Lb[r][c] = (r < nb && c < nb) ? a[(o + r) + (size_t)(o + c) * n] : (r == c ? Scal<T>::one() : Scal<T>::zero());
        }
        __syncthreads();
        if (wave == 0) {
            T v = lane < nb ? yv[o + lane] : Scal<T>::zero();
            for (int r = nb - 1; r >= 0; --r) {
                const T zr = Scal<T>::shfl(v, r) / Lb[r][r];
                if (lane == r) v = zr;
                else if (lane < r) v -= Lb[r][lane] * zr;
            }
            if (lane < nb) { z[lane] = v; yv[o + lane] = v; }
        }
        __syncthreads();
        // y[c] -= sum_t L[o+t][c] z[t] for c < o: column segment a[o..o+nb, c] is contiguous
        for (int c = wave; c < o; c += nwave) {
            T p = lane < nb ? a[(o + lane) + (size_t)c * n] * z[lane] : Scal<T>::zero();
            for (int off = 32; off > 0; off >>= 1) p += Scal<T>::shfl_xor(p, off);
            if (lane == 0) yv[c] -= p;
        }
        __syncthreads();
    }
}

// residual r = U_meas - U_calc (FP64), gamma += delta
__global__ void __launch_bounds__(256) k_sub(int n, const double* __restrict__ a, const double* __restrict__ b,
                                             double* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < (size_t)n) out[i] = a[i] - b[i];
}

__global__ void __launch_bounds__(256) k_add_to(int n, double* __restrict__ y, const double* __restrict__ dy) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < (size_t)n) y[i] += dy[i];
}

// y_b += dy_b and ratio[b] = max|dy_b| / max|y_b| (after the update): one block per vector
__global__ void __launch_bounds__(256) k_add_to_ratio(int n, double* __restrict__ y, const double* __restrict__ dy,
                                                      double* __restrict__ ratio) {
    const int b = blockIdx.x;
    double* yb = y + (size_t)b * n;
    const double* db = dy + (size_t)b * n;
    double md = 0.0, my = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        const double v = yb[i] + db[i];
        yb[i] = v;
        md = fmax(md, fabs(db[i]));
        my = fmax(my, fabs(v));
    }
    __shared__ double sd[256], sy[256];
    sd[threadIdx.x] = md;
    sy[threadIdx.x] = my;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sd[threadIdx.x] = fmax(sd[threadIdx.x], sd[threadIdx.x + s]);
            sy[threadIdx.x] = fmax(sy[threadIdx.x], sy[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) ratio[b] = sy[0] > 0.0 ? sd[0] / sy[0] : (sd[0] > 0.0 ? 1.0 : 0.0);
}

__global__ void __launch_bounds__(256) k_gamma_update(int n, float* __restrict__ g, const double* __restrict__ d) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < (size_t)n) g[i] = (float)((double)g[i] + d[i]);
}

// out[b][i] = a[b][i] - u[i]  (every frame against one reference voltage vector)
__global__ void __launch_bounds__(256) k_sub_bcast(int n, int batch, const double* __restrict__ a,
                                                   const double* __restrict__ u, double* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < (size_t)n * batch) out[i] = a[i] - u[i % n];
}

// d_e = sum_i A[i + e*lda]^2 (column squares, FP64), one wavefront per column
__global__ void __launch_bounds__(256) k_col_sq(int rows, int cols, const float* __restrict__ A, int lda,
                                                double* __restrict__ d) {
    const int col = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (col >= cols) return;
    const float* a = A + (size_t)col * lda;
    double s = 0.0;
    for (int r = lane; r < rows; r += 64) s += (double)a[r] * (double)a[r];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) d[col] = s;
}

// A_b[i + e*lda] *= w[e] for every batch entry
__global__ void __launch_bounds__(256) k_scale_cols(int rows, int cols, float* __restrict__ A, int lda,
                                                    size_t strideA, const float* __restrict__ w) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)rows * cols) return;
    const size_t e = i / rows, r = i % rows;
    A[blockIdx.y * strideA + r + e * lda] *= w[e];
}

// out[b][e] = u[b][e] + c (x[b][e] + sum_k val[e][k] x[b][nb[e][k]])
__global__ void __launch_bounds__(256) k_lap_combine(int n, const double* __restrict__ u, double c,
                                                     const double* __restrict__ x, const int32_t* __restrict__ nb,
                                                     const double* __restrict__ val, double* __restrict__ out) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const size_t o = (size_t)blockIdx.y * n;
    double bx = x[o + e];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int f = nb[(size_t)e * 3 + k];
        if (f >= 0) bx += val[(size_t)e * 3 + k] * x[o + f];
    }
    out[o + e] = u[o + e] + c * bx;
}

__global__ void __launch_bounds__(256) k_axmy(size_t n, double a, const double* __restrict__ r,
                                              const double* __restrict__ u, double* __restrict__ z) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) z[i] = a * (r[i] - u[i]);
}

// one block per batch column: strided FP64 partial sums, then a fixed tree
__global__ void __launch_bounds__(256) k_dots(int n, const double* __restrict__ a, const double* __restrict__ c,
                                              double* __restrict__ out, int ostride) {
    __shared__ double red[256];
    const size_t o = (size_t)blockIdx.x * n;
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += a[o + i] * c[o + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[(size_t)blockIdx.x * ostride] = red[0];
}

__global__ void __launch_bounds__(256) k_cg_xr(int n, const double* __restrict__ alpha, double* __restrict__ x,
                                               const double* __restrict__ p, double* __restrict__ r,
                                               const double* __restrict__ q) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const size_t i = (size_t)blockIdx.y * n + e;
    const double al = alpha[blockIdx.y];
    x[i] += al * p[i];
    r[i] -= al * q[i];
}

__global__ void __launch_bounds__(256) k_cg_p(int n, const double* __restrict__ beta, double* __restrict__ p,
                                              const double* __restrict__ z) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const size_t i = (size_t)blockIdx.y * n + e;
    p[i] = z[i] + beta[blockIdx.y] * p[i];
}

// a[i] *= b[i]
__global__ void __launch_bounds__(256) k_mul(size_t n, float* __restrict__ a, const float* __restrict__ b) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) a[i] *= b[i];
}

// v[b][e] *= w[e]
__global__ void __launch_bounds__(256) k_scale_vec(int n, int batch, double* __restrict__ v,
                                                   const float* __restrict__ w) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < (size_t)n * batch) v[i] *= (double)w[i % n];
}

// g[b][e] = g0[e] + d[b][e]
__global__ void __launch_bounds__(256) k_gamma_from(int n, int batch, const float* __restrict__ g0,
                                                    const double* __restrict__ d, float* __restrict__ g) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < (size_t)n * batch) g[i] = (float)((double)g0[i % n] + d[i]);
}

// --------------------------------------------------------------------------------
// launch wrappers
// --------------------------------------------------------------------------------
static inline void check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error(MPF_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
}

void launch_voltages(const JobDev* d_jobs, int njobs, const MeasDev& md, int use64, float* U32,
                     double* U64, hipStream_t st) {
    k_voltages<<<njobs, 256, 0, st>>>(d_jobs, md, use64, U32, U64);
    check_launch();
}

void launch_jacobian(const JobDev* d_jobs, int njobs, const JacDev& jd, int use64, float* J,
                     hipStream_t st, uint16_t* Jb) {
    MPF_REQUIRE(jd.Mp * jd.Dp <= kJacMaxRows, "jacobian: too many rows");
    MPF_REQUIRE(jd.D + jd.M <= 64 && jd.k <= 6, "jacobian: unsupported shape");
    dim3 grid(((jd.E + kJacWaves - 1) / kJacWaves + 7) / 8 * 8, njobs);
    const int mrows = jd.Mp * jd.Dp;
    __bf16* jb = reinterpret_cast<__bf16*>(Jb);
    const bool fixedm = 64 % jd.Mp == 0;
    auto go = [&](auto per) {
        constexpr int P = decltype(per)::value;
        if (jd.k == 3) {
            if (fixedm) k_jacobian<P, true, 3><<<grid, 256, 0, st>>>(d_jobs, jd, use64, J, jb);
            else k_jacobian<P, false, 3><<<grid, 256, 0, st>>>(d_jobs, jd, use64, J, jb);
        } else {
            if (fixedm) k_jacobian<P, true, 6><<<grid, 256, 0, st>>>(d_jobs, jd, use64, J, jb);
            else k_jacobian<P, false, 6><<<grid, 256, 0, st>>>(d_jobs, jd, use64, J, jb);
        }
    };
    MPF_REQUIRE(jd.k == 3 || jd.k == 6, "jacobian: 3 or 6 nodes per element");
    if (mrows <= 256) go(std::integral_constant<int, 4>());
    else if (mrows <= 1024) go(std::integral_constant<int, 16>());
    else go(std::integral_constant<int, 64>());
    check_launch();
}

size_t gram_workspace_bytes(int n, int nk) {
    const int nt = (n + kGT - 1) / kGT;
    const int ntiles = nt * (nt + 1) / 2;
    const int kslab = gram_kslab(n, nk);
    const int nslab = (nk + kslab - 1) / kslab;
    return (size_t)ntiles * nslab * kGT * kGT * sizeof(float) + (size_t)ntiles * sizeof(int2);
}

int gram_kslab(int n, int nk) {
    const int nt = (n + kGT - 1) / kGT;
    const int ntiles = nt * (nt + 1) / 2;
    // aim for >= 1024 blocks, slabs a multiple of kGK and at least 256 deep
    int nslab = (1024 + ntiles - 1) / ntiles;
    int kslab = (nk + nslab - 1) / nslab;
    kslab = ((kslab + kGK - 1) / kGK) * kGK;
    if (kslab < 256) kslab = 256;
    return kslab;
}

// slabs of the 128-tile kernel: about 512 blocks per launch over all batch entries, at
// most max_slabs (> 0) of them
static int gram_nt_slabs(int n, int nk, int batch, int max_slabs) {
    const int nt = (n + kGT2 - 1) / kGT2;
    const int ntiles = nt * (nt + 1) / 2;
    int nslab = std::max(1, (512 + ntiles * batch - 1) / (ntiles * batch));
    nslab = std::min(nslab, std::max(1, (nk + 255) / 256));
    if (max_slabs > 0) nslab = std::min(nslab, max_slabs);
    return nslab;
}

bool gram_nt_applies(int n, int lda) { return n % 4 == 0 && lda % 4 == 0; }

size_t gram_nt_workspace_bytes(int n, int nk, int batch, int max_slabs) {
    const int nt = (n + kGT2 - 1) / kGT2;
    const int ntiles = nt * (nt + 1) / 2;
    return (size_t)batch * gram_nt_slabs(n, nk, batch, max_slabs) * ntiles * kGT2 * kGT2 * sizeof(float);
}

void launch_gram_nt(const float* A, size_t strideA, int n, int nk, int lda, float* G, size_t strideG, int batch,
                    void* work, hipStream_t st, int max_slabs) {
    MPF_REQUIRE(gram_nt_applies(n, lda), "gram: rows and leading dimension must be multiples of 4");
    const int nt = (n + kGT2 - 1) / kGT2;
    const int ntiles = nt * (nt + 1) / 2;
    const int nslab = gram_nt_slabs(n, nk, batch, max_slabs);
    int kslab = (nk + nslab - 1) / nslab;
    kslab = std::max(kGK2, (kslab + kGK2 - 1) / kGK2 * kGK2);
    const int used = std::max(1, (nk + kslab - 1) / kslab);
    float* part = (float*)work;
    k_gram_nt<<<dim3(ntiles, batch, used), 256, 0, st>>>(A, strideA, n, nk, lda, nt, kslab, part);
    k_gram_reduce128<<<dim3(ntiles, 16, batch), 256, 0, st>>>(part, nt, used, n, G, strideG);
    check_launch();
}

__global__ void __launch_bounds__(256) k_to_bf16(size_t n, const float* __restrict__ a, __bf16* __restrict__ b) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) b[i] = (__bf16)a[i];
}

void launch_to_bf16(size_t n, const float* a, uint16_t* b, hipStream_t st) {
    if (!n) return;
    k_to_bf16<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, a, reinterpret_cast<__bf16*>(b));
    check_launch();
}

// slabs of the 256-tile kernel: about 512 blocks over all batch entries, slabs of at
// least 1024 contraction steps
static int gram_bf16_slabs(int n, int nk, int batch) {
    const int nt = (n + kGT3 - 1) / kGT3;
    const int ntiles = nt * (nt + 1) / 2;
    int nslab = std::max(1, (512 + ntiles * batch - 1) / (ntiles * batch));
    return std::min(nslab, std::max(1, nk / 1024));
}

bool gram_bf16_applies(int n, int lda) { return n % 8 == 0 && lda % 8 == 0; }

size_t gram_bf16_workspace_bytes(int n, int nk, int batch) {
    const int nt = (n + kGT3 - 1) / kGT3;
    const int ntiles = nt * (nt + 1) / 2;
    return (size_t)batch * gram_bf16_slabs(n, nk, batch) * ntiles * kGT3 * kGT3 * sizeof(float);
}

void launch_gram_bf16(const uint16_t* A, size_t strideA, int n, int nk, int lda, float* G, size_t strideG,
                      int batch, void* work, hipStream_t st) {
    MPF_REQUIRE(gram_bf16_applies(n, lda), "gram: rows and leading dimension must be multiples of 8");
    const int nt = (n + kGT3 - 1) / kGT3;
    const int ntiles = nt * (nt + 1) / 2;
    const int nslab = gram_bf16_slabs(n, nk, batch);
    int kslab = (nk + nslab - 1) / nslab;
    kslab = std::max(kGK3, (kslab + kGK3 - 1) / kGK3 * kGK3);
    const int used = std::max(1, (nk + kslab - 1) / kslab);
    float* part = (float*)work;
    k_gram_bf16<<<dim3(ntiles, batch, used), 512, 0, st>>>(reinterpret_cast<const __bf16*>(A), strideA, n, nk, lda,
                                                            nt, kslab, part);
    k_gram_reduce256<<<dim3(ntiles, 16, batch), 256, 0, st>>>(part, nt, used, n, G, strideG);
    check_launch();
}

void launch_gram(const float* A, int trans, int n, int nk, int lda, float* G, void* work,
                 hipStream_t st) {
    const int nt = (n + kGT - 1) / kGT;
    const int ntiles = nt * (nt + 1) / 2;
    const int kslab = gram_kslab(n, nk);
    const int nslab = (nk + kslab - 1) / kslab;
    float* part = (float*)work;
    int2* tiles = (int2*)((char*)work + (size_t)ntiles * nslab * kGT * kGT * sizeof(float));
    std::vector<int2> h;
    for (int i = 0; i < nt; ++i)
        for (int j = i; j < nt; ++j) h.push_back(make_int2(i, j));
    MPF_HIP(hipMemcpyAsync(tiles, h.data(), h.size() * sizeof(int2), hipMemcpyHostToDevice, st));
    MPF_HIP(hipStreamSynchronize(st));  // h goes out of scope
    dim3 grid(ntiles, 1, nslab);
    k_gram_mfma<<<grid, 256, 0, st>>>(A, trans, n, nk, lda, tiles, kslab, part);
    k_gram_reduce<<<ntiles, 256, 0, st>>>(part, tiles, ntiles, nslab, n, G);
    check_launch();
}

void launch_gemv_t(int rows, int cols, const float* A, int lda, size_t strideA, const double* x,
                   size_t stridex, double* y, size_t stridey, int batch, hipStream_t st) {
    dim3 grid((cols + 3) / 4, batch);
    if (rows % 4 == 0 && lda % 4 == 0 && strideA % 4 == 0 && ((uintptr_t)A & 15) == 0)
        k_gemv_t4<<<grid, 256, 0, st>>>(rows, cols, A, lda, strideA, x, stridex, y, stridey);
    else
        k_gemv_t<<<grid, 256, 0, st>>>(rows, cols, A, lda, strideA, x, stridex, y, stridey, 0.0);
    check_launch();
}

bool gemv_tn_applies(int rows, int lda, size_t strideA, const float* A) {
    return rows > 0 && rows <= 1024 && rows % 4 == 0 && lda % 4 == 0 && strideA % 4 == 0 && ((uintptr_t)A & 15) == 0;
}

void launch_gemv_tn(int rows, int cols, const float* A, int lda, size_t strideA, const double* x,
                    size_t stridex, double shift, const double* rhs, size_t striderhs, double* out,
                    size_t strideout, int batch, double* work, hipStream_t st) {
    MPF_REQUIRE(gemv_tn_applies(rows, lda, strideA, A), "gemv_tn: rows <= 1024, multiples of 4, aligned A");
    const int nslab = (cols + kNSlab - 1) / kNSlab;
    k_gemv_tn_part<<<dim3(nslab, batch), 256, 0, st>>>(rows, cols, A, lda, strideA, x, stridex, work);
    dim3 g2((rows + 255) / 256, batch);
    k_gemv_n_reduce<<<g2, 256, 0, st>>>(rows, nslab, work, shift, x, stridex, rhs, striderhs, out, strideout);
    check_launch();
}

size_t gemv_n_workspace(int rows, int cols, int batch) {
    return (size_t)batch * ((cols + kNSlab - 1) / kNSlab) * rows * sizeof(double);
}

void launch_gemv_n(int rows, int cols, const float* A, int lda, size_t strideA, const double* x,
                   size_t stridex, double shift, const double* z, size_t stridez, const double* rhs,
                   size_t striderhs, double* out, size_t strideout, int batch, double* work,
                   hipStream_t st) {
    const int nslab = (cols + kNSlab - 1) / kNSlab;
    dim3 g1((rows + 255) / 256, nslab, batch);
    k_gemv_n_part<<<g1, 256, 0, st>>>(rows, cols, A, lda, strideA, x, stridex, work);
    dim3 g2((rows + 255) / 256, batch);
    k_gemv_n_reduce<<<g2, 256, 0, st>>>(rows, nslab, work, shift, z, stridez, rhs, striderhs, out, strideout);
    check_launch();
}

// the (ib, jb) pairs of the trailing updates, elimination steps in order: step kb lists
// jb = kb + 1 .. nblk - 1 and, for each, ib = jb .. nblk - 1, and starts at
// chol_pair_start(nblk, kb) = sum over k < kb of T(nblk - k - 1), T(m) = m (m + 1) / 2
__host__ __device__ inline int chol_pair_start(int nblk, int kb) {
    int s = 0;
    for (int k = 0; k < kb; ++k) {
        const int m = nblk - k - 1;
        s += m * (m + 1) / 2;
    }
    return s;
}

__global__ void __launch_bounds__(256) k_chol_pairs(int nblk, int2* __restrict__ pairs) {
    const int kb = blockIdx.x * 256 + threadIdx.x;
    if (kb >= nblk) return;
    int i = chol_pair_start(nblk, kb);
    for (int jb = kb + 1; jb < nblk; ++jb)
        for (int ib = jb; ib < nblk; ++ib) pairs[i++] = make_int2(ib, jb);
}

void launch_cholesky(int n, const float* G, size_t strideG, double shift, double* A, size_t strideA,
                     int batch, int* d_info, int2* d_pairs, hipStream_t st) {
    dim3 g0((unsigned)(((size_t)n * n + 255) / 256), batch);
    k_to64_shift<<<g0, 256, 0, st>>>(n, G, strideG, shift, A, strideA);
    const int nblk = (n + kCB - 1) / kCB;
    k_chol_pairs<<<(nblk + 255) / 256, 256, 0, st>>>(nblk, d_pairs);
    for (int kb = 0; kb < nblk; ++kb) {
        k_potrf_diag<double><<<dim3(1, 1, batch), 256, 0, st>>>(n, A, strideA, kb, d_info);
        if (kb + 1 < nblk) {
            k_potrf_panel<double><<<dim3(nblk - kb - 1, 1, batch), 256, 0, st>>>(n, A, strideA, kb);
            const int p0 = chol_pair_start(nblk, kb), np = chol_pair_start(nblk, kb + 1) - p0;
            k_potrf_update<double><<<dim3(np, 1, batch), 256, 0, st>>>(n, A, strideA, kb, d_pairs + p0);
        }
    }
    check_launch();
}

size_t cholesky_pairs(int n) {
    const int nblk = (n + kCB - 1) / kCB;
    return (size_t)nblk * nblk * nblk / 2 + 16;
}

void launch_potrs(int n, const double* A, size_t strideA, const double* b, size_t strideb, double* y,
                  size_t stridey, int batch, hipStream_t st) {
    k_potrs<double><<<batch, kTrsvThreads, 0, st>>>(n, A, strideA, b, strideb, y, stridey);
    check_launch();
}

// complex symmetric A = L L^T in place (A: shifted FP64 matrices, both triangles)
void launch_cholesky_z(int n, zd* A, size_t strideA, int batch, int* d_info, int2* d_pairs, hipStream_t st) {
    const int nblk = (n + kCB - 1) / kCB;
    k_chol_pairs<<<(nblk + 255) / 256, 256, 0, st>>>(nblk, d_pairs);
    for (int kb = 0; kb < nblk; ++kb) {
        k_potrf_diag<zd><<<dim3(1, 1, batch), 256, 0, st>>>(n, A, strideA, kb, d_info);
        if (kb + 1 < nblk) {
            k_potrf_panel<zd><<<dim3(nblk - kb - 1, 1, batch), 256, 0, st>>>(n, A, strideA, kb);
            const int p0 = chol_pair_start(nblk, kb), np = chol_pair_start(nblk, kb + 1) - p0;
            k_potrf_update<zd><<<dim3(np, 1, batch), 256, 0, st>>>(n, A, strideA, kb, d_pairs + p0);
        }
    }
    check_launch();
}

void launch_potrs_z(int n, const zd* A, size_t strideA, const zd* b, size_t strideb, zd* y, size_t stridey,
                    int batch, hipStream_t st) {
    k_potrs<zd><<<batch, kTrsvThreads, 0, st>>>(n, A, strideA, b, strideb, y, stridey);
    check_launch();
}

void launch_sub(int n, const double* a, const double* b, double* out, hipStream_t st) {
    k_sub<<<(n + 255) / 256, 256, 0, st>>>(n, a, b, out);
    check_launch();
}

void launch_add_to(int n, double* y, const double* dy, hipStream_t st) {
    k_add_to<<<(n + 255) / 256, 256, 0, st>>>(n, y, dy);
    check_launch();
}

void launch_add_to_ratio(int n, int batch, double* y, const double* dy, double* ratio, hipStream_t st) {
    k_add_to_ratio<<<batch, 256, 0, st>>>(n, y, dy, ratio);
    check_launch();
}

void launch_gamma_update(int n, float* g, const double* d, hipStream_t st) {
    k_gamma_update<<<(n + 255) / 256, 256, 0, st>>>(n, g, d);
    check_launch();
}

void launch_sub_bcast(int n, int batch, const double* a, const double* u, double* out, hipStream_t st) {
    k_sub_bcast<<<(unsigned)(((size_t)n * batch + 255) / 256), 256, 0, st>>>(n, batch, a, u, out);
    check_launch();
}

void launch_col_sq(int rows, int cols, const float* A, int lda, double* d, hipStream_t st) {
    k_col_sq<<<(cols + 3) / 4, 256, 0, st>>>(rows, cols, A, lda, d);
    check_launch();
}

void launch_scale_cols(int rows, int cols, float* A, int lda, size_t strideA, int batch, const float* w,
                       hipStream_t st) {
    dim3 grid((unsigned)(((size_t)rows * cols + 255) / 256), batch);
    k_scale_cols<<<grid, 256, 0, st>>>(rows, cols, A, lda, strideA, w);
    check_launch();
}

void launch_lap_combine(int n, int batch, const double* u, double c, const double* x, const int32_t* nb,
                        const double* val, double* out, hipStream_t st) {
    k_lap_combine<<<dim3((n + 255) / 256, batch), 256, 0, st>>>(n, u, c, x, nb, val, out);
    check_launch();
}

void launch_axmy(size_t n, double a, const double* r, const double* u, double* z, hipStream_t st) {
    k_axmy<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, a, r, u, z);
    check_launch();
}

void launch_dots(int n, int batch, const double* a, const double* c, double* out, int ostride, hipStream_t st) {
    k_dots<<<batch, 256, 0, st>>>(n, a, c, out, ostride);
    check_launch();
}

void launch_cg_xr(int n, int batch, const double* alpha, double* x, const double* p, double* r,
                  const double* q, hipStream_t st) {
    k_cg_xr<<<dim3((n + 255) / 256, batch), 256, 0, st>>>(n, alpha, x, p, r, q);
    check_launch();
}

void launch_cg_p(int n, int batch, const double* beta, double* p, const double* z, hipStream_t st) {
    k_cg_p<<<dim3((n + 255) / 256, batch), 256, 0, st>>>(n, beta, p, z);
    check_launch();
}

void launch_mul(size_t n, float* a, const float* b, hipStream_t st) {
    if (n == 0) return;
    k_mul<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, a, b);
    check_launch();
}

void launch_scale_vec(int n, int batch, double* v, const float* w, hipStream_t st) {
    k_scale_vec<<<(unsigned)(((size_t)n * batch + 255) / 256), 256, 0, st>>>(n, batch, v, w);
    check_launch();
}

void launch_gamma_from(int n, int batch, const float* g0, const double* d, float* g, hipStream_t st) {
    k_gamma_from<<<(unsigned)(((size_t)n * batch + 255) / 256), 256, 0, st>>>(n, batch, g0, d, g);
    check_launch();
}

}  // namespace mpf
