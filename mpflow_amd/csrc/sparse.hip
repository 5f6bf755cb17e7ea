// Sparse kernels of the forward solve on CDNA4 (gfx950):
//   sigma-weighted gather assembly of S and R, the Jacobi diagonal, and the
//   multi-job / multi-harmonic / multi-RHS Jacobi-PCG building blocks.
//
// Work items are (job, slice, 256-row tile).
//  * SpMV: a 256-thread block = 4 wavefronts, each owning 64 consecutive rows.  The
//    wavefront loads the column-major (slot-major) ELL of its 64 rows with coalesced
//    loads, one lane per row, and then walks the rows, broadcasting each row's column
//    ids and values to all lanes (v_readlane for 64 RHS lanes, ds_bpermute for fewer);
//    every gathered neighbour is then one contiguous Rp*4-byte vector load.
//  * Streaming kernels (updates, init, residual): 1024-thread blocks, lane = RHS,
//    16 (Rp=64) row groups.
// Dot products are accumulated in FP64 per lane, combined across row groups and
// wavefronts in a fixed order into one partial per tile, and summed per (job, slice)
// by a one-block reduction in tile order: every scalar is deterministic run to run.
#include "krylov_dev.h"

namespace mpf {

__device__ __forceinline__ double sigma_of(const JobDev& J, float g) {
    // matParam (src/fem/equation_kernel.cu:35-43)
    return J.logarithmic ? J.ref * exp(2.302585092994046 * (double)g / 10.0) : J.ref * (double)g;
}

// --------------------------------------------------------------------------------
// assembly: S = sum_e sigma_e K_e, R = sum_e sigma_e M_e on the ELL pattern
// (updateMatrixKernel, src/fem/equation_kernel.cu:60-82), FP64 sums split hi/lo
// --------------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_assemble(const JobDev* __restrict__ jobs,
                                                     const int4* __restrict__ items) {
    const int4 it = items[blockIdx.x];
    const JobDev& J = jobs[it.x];
    const size_t ent = (size_t)it.y * kBlock + threadIdx.x;
    if (ent >= (size_t)J.width * J.np) return;
    const int b = J.ctb_off[ent], e = J.ctb_off[ent + 1];
    double S = 0.0, R = 0.0;
    for (int i = b; i < e; ++i) {
        // element -1: a sigma-independent contribution (complete-electrode-model blocks)
        const int el = J.ctb_elem[i];
        const double sg = el < 0 ? 1.0 : sigma_of(J, J.gamma[el]);
        S += sg * J.ctb_k[i];
        R += sg * J.ctb_m[i];
    }
    const float sh = (float)S, rh = (float)R;
    J.s[ent] = sh; J.s_lo[ent] = (float)(S - (double)sh);
    J.r[ent] = rh; J.r_lo[ent] = (float)(R - (double)rh);
}

// Jacobi diagonal 1/(S_ii + alpha_c R_ii) in FP32 (diagonalSparseKernel,
// src/numeric/preconditioner_kernel.cu:63-83); padding rows 0.
__global__ void __launch_bounds__(kBlock) k_dinv(const JobDev* __restrict__ jobs,
                                                 const int4* __restrict__ tiles) {
    const int4 t = tiles[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int row = t.z * kTileRows + threadIdx.x;
    const int s = t.y;
    float v = 0.0f;
    // a row without a diagonal entry (diag_slot -1, caller matrices only) gets 0
    if (row < J.n && J.diag_slot[row] >= 0) {
        const size_t e = (size_t)J.diag_slot[row] * J.np + row;
        const float a = J.s[e] + J.alpha[s] * J.r[e];
        v = a != 0.0f ? 1.0f / a : 0.0f;
    }
    J.dinv[(size_t)s * J.np + row] = v;
}

// --------------------------------------------------------------------------------
// PCG (ConjugateGradient::solve, src/numeric/conjugate_gradient.cpp:61-161)
// --------------------------------------------------------------------------------
// init: x = 0, r = b, z = D^-1 r, p = z; partials r.z and b.b
__global__ void __launch_bounds__(kEwBlock) k_pcg_init(const JobDev* __restrict__ jobs,
                                                       const int4* __restrict__ tiles) {
    const int4 t = tiles[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    const Ew L = ew_of(J);
    const float* dinv = J.dinv + (size_t)s * J.np;
    const float xs = J.xscale[s];
    const int tile0 = t.z * kTileRows;
    double rz = 0.0, bb = 0.0, bs[kAggPerTile] = {0.0, 0.0, 0.0, 0.0};
    for (int row = tile0 + L.g; row < tile0 + kTileRows; row += L.G) {
        const size_t i = vec_index(J, s, row, L.lane);
        float b = 0.0f;
        if (L.lane < J.R) b = J.use_rhs ? J.rhs[i] : (float)(xs * J.F[s * J.fstride + (size_t)row * J.Rp + L.lane]);
        const float z = dinv[row] * b;
        J.x[i] = 0.0f;
        J.rv[i] = b;
        J.p[i] = z;
        rz += (double)b * (double)z;
        bb += (double)b * (double)b;
        const int q = row_agg(J, row);
#pragma unroll
        for (int a = 0; a < kAggPerTile; ++a) bs[a] += a == q ? (double)b : 0.0;
    }
    // channels agg_channel(q): W^T r of the tile's aggregates (two-level preconditioner)
    const int ch[6] = {0, 1, agg_channel(0), agg_channel(1), agg_channel(2), agg_channel(3)};
    const double v[6] = {rz, bb, bs[0], bs[1], bs[2], bs[3]};
    block_partials<6, kEwBlock>(J, s, t.z, ch, v);
}

// two-level start: p = z = D^-1 b + W y (y from the coarse kernel, real rows only)
__global__ void __launch_bounds__(kEwBlock) k_pcg_init_p(const JobDev* __restrict__ jobs,
                                                         const int4* __restrict__ tiles) {
    const int4 t = tiles[blockIdx.x];
    const JobDev& J = jobs[t.x];
    if (!J.cinv) return;
    const int s = t.y;
    const int LPR = J.Rp / 4, tile0 = t.z * kTileRows;
    // 128-bit accesses: the tile's rows of p as [row][LPR] float4, y as [aggregate][LPR]
    const MPF_GLOBAL float4* Y =
        (const MPF_GLOBAL float4*)gp((const float*)J.cy) + ((size_t)s * kAggPerTile * J.ntiles + kAggPerTile * t.z) * LPR;
    MPF_GLOBAL float4* P = (MPF_GLOBAL float4*)gp(J.p) + ((size_t)s * J.np + tile0) * LPR;
    for (int i = threadIdx.x; i < kTileRows * LPR; i += kEwBlock) {
        const int row = tile0 + i / LPR;
        if (row >= J.n) continue;
        const float4 y = Y[row_agg(J, row) * LPR + i % LPR];
        float4 p = P[i];
        p.x += y.x; p.y += y.y; p.z += y.z; p.w += y.w;
        P[i] = p;
    }
}

__global__ void __launch_bounds__(kEwBlock) k_pcg_init_reduce(const JobDev* __restrict__ jobs,
                                                              const int2* __restrict__ js,
                                                              int* __restrict__ active) {
    const int2 t = js[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    const int ch[2] = {0, 1};
    double out[2];
    reduce_tiles<2>(J, s, ch, out);
    if (J.cinv) {
        // rho = b.D^-1 b + (W^T b).y, summed per coarse chunk by the coarse kernel
        const int chc[1] = {5};
        double oc[1];
        reduce_tiles<1>(J, s, chc, oc, coarse_chunks(J));
        out[0] = oc[0];
    }
    if ((int)threadIdx.x < J.Rp) {
        sc_ptr(J, SC_RHO, s)[threadIdx.x] = out[0];
        // a correction solve started from a previous solution stops relative to the
        // full right-hand side ||beta_c F||^2 left in SC_X2 by the FP64 residual
        sc_ptr(J, SC_BB, s)[threadIdx.x] = J.ref_b ? sc_ptr(J, SC_X2, s)[threadIdx.x] : out[1];
        sc_ptr(J, SC_RS, s)[threadIdx.x] = out[0];
    }
    if (threadIdx.x == 0) {
        J.st[s * kStInts + 0] = 1;
        J.st[s * kStInts + 1] = 0;
        J.st[s * kStInts + 2] = 0;
        atomicAdd(active, 1);
    }
}

// q = A_s p (+ 1 1^T p when dcFree, applied in the update); tile partials p.q and 1^T p.
// WMAX = padded ELL width (compile time).
template <int LG, int WMAX>
__global__ void __launch_bounds__(kEwBlock) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_pcg_spmv4(const JobDev* __restrict__ jobs,
                                                        const int4* __restrict__ tiles, int k4) {
    constexpr int Rp = 1 << LG;
    constexpr int LPR = Rp / 4;
    const int4 t = tiles[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    if (!slice_active(J, s)) return;
    const int wave = threadIdx.x >> 6, l64 = threadIdx.x & 63;
    const int row0 = t.z * kTileRows + wave * kSpmvRowsPerWave;
    const size_t slice_off = (size_t)s * J.np * Rp;
    MPF_GLOBAL float4* Q4 = (MPF_GLOBAL float4*)(gp(J.q) + slice_off);
    double v[2][4] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
    // LDS: the wavefronts' ELL entries (later reused for the partial sums) and the tile's
    // own rows of p, from which the in-tile gathers and p itself are read
    __shared__ union {
        EllStage<WMAX> stage;
        double red[2][kEwBlock / 64][64];
    } sh;
    __shared__ XTile<LG> xs;
    const int tile0 = t.z * kTileRows;
    // k4: the iteration index modulo kRing, which selects this iteration's p
    const float* P = pbuf(J, k4) + slice_off;
    const __amdgpu_buffer_rsrc_t prs = uniform_rsrc(P, (size_t)J.np * Rp * sizeof(float));
    ell_gather_wave<LG, WMAX, true>(J, s, row0, l64, P, sh.stage, [&](int row, int c4, float4 acc) {
        const float4 own = xt_own<LG>(xs, prs, row, tile0, c4);
        Q4[(size_t)row * LPR + c4] = acc;
        v[0][0] += (double)own.x * (double)acc.x; v[0][1] += (double)own.y * (double)acc.y;
        v[0][2] += (double)own.z * (double)acc.z; v[0][3] += (double)own.w * (double)acc.w;
        v[1][0] += (double)own.x; v[1][1] += (double)own.y; v[1][2] += (double)own.z; v[1][3] += (double)own.w;
    }, xs);
    const int ch[2] = {0, 1};
    block_partials4_buf<LG, 2>(J, s, t.z, ch, v, sh.red);
}

// k_pcg_spmv4 for every live harmonic of one (job, tile) work item (tiles[b].y unused):
// the wavefront's ELL entries (id, S, R) are loaded once, kept in registers and restaged
// in LDS with each harmonic's alpha_s (the same S + alpha_s R value as k_pcg_spmv4), so
// the tile's matrix is read once for all harmonics instead of once per harmonic.  Per
// harmonic the tile's rows of p go to LDS, then q, p.q and 1^T p follow k_pcg_spmv4.
template <int LG, int WMAX>
__global__ void __launch_bounds__(kEwBlock) k_pcg_spmv4h(const JobDev* __restrict__ jobs,
                                                         const int4* __restrict__ tiles, int k4) {
    constexpr int Rp = 1 << LG;
    constexpr int LPR = Rp / 4;
    constexpr int NX = (kTileRows * LPR + kEwBlock - 1) / kEwBlock;
    constexpr int NE = (WMAX * kSpmvRowsPerWave + 63) / 64;
    const int4 t = tiles[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int wave = threadIdx.x >> 6, l64 = threadIdx.x & 63;
    const int tile0 = t.z * kTileRows;
    const int row0 = tile0 + wave * kSpmvRowsPerWave;
    const int np = J.np;
    __shared__ union {
        EllStage<WMAX> stage;
        double red[2][kEwBlock / 64][64];
    } sh;
    __shared__ XTile<LG> xs;
    const int wl = __builtin_amdgcn_readfirstlane((int)J.gw[row0 / kSpmvRowsPerWave]);
    const int wb = ell_bin<WMAX>(wl);
    int idv[NE];
    float sv[NE], rv[NE];
    {
        const MPF_GLOBAL int32_t* ids = gp(J.ids);
        const MPF_GLOBAL float* Sv = gp((const float*)J.s);
        const MPF_GLOBAL float* Rv = gp((const float*)J.r);
#pragma unroll
        for (int k = 0; k < NE; ++k) {
            const int e = l64 + 64 * k, q = e / kSpmvRowsPerWave, r = e % kSpmvRowsPerWave;
            idv[k] = 0; sv[k] = 0.0f; rv[k] = 0.0f;
            if (q < wl) {
                const size_t g = (size_t)q * np + row0 + r;
                idv[k] = ids[g]; sv[k] = Sv[g]; rv[k] = Rv[g];
            }
        }
#pragma unroll
        for (int k = 0; k < NE; ++k) {
            const int q = (l64 + 64 * k) / kSpmvRowsPerWave;
            idv[k] = xt_key<LG>(idv[k], tile0, q < wl);
        }
    }
    const int ch[2] = {0, 1};
    for (int s = 0; s < J.C; ++s) {
        if (!slice_active(J, s)) continue;
        const size_t slice_off = (size_t)s * np * Rp;
        const float* X = pbuf(J, k4) + slice_off;
        const MPF_GLOBAL float4* X4 = (const MPF_GLOBAL float4*)gp(X) + (size_t)tile0 * LPR;
        float4 xv[NX];
#pragma unroll
        for (int k = 0; k < NX; ++k) {
            const int i = threadIdx.x + k * kEwBlock;
            if (i < kTileRows * LPR) xv[k] = X4[i];
        }
        // the previous harmonic's readers of xs and of the stage / partial-sum room are done
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NX; ++k) {
            const int i = threadIdx.x + k * kEwBlock;
            if (i < kTileRows * LPR) xs[i] = xt_value<LG>(i, xv[k]);
        }
        const float al = J.alpha[s];
#pragma unroll
        for (int k = 0; k < NE; ++k) {
            const int e = l64 + 64 * k, q = e / kSpmvRowsPerWave, r = e % kSpmvRowsPerWave;
            if (q < wb) sh.stage[wave][q][r] = make_int2(idv[k], q < wl ? __float_as_int(sv[k] + al * rv[k]) : 0);
        }
        __syncthreads();
        MPF_GLOBAL float4* Q4 = (MPF_GLOBAL float4*)(gp(J.q) + slice_off);
        double v[2][4] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
        const __amdgpu_buffer_rsrc_t xrs = uniform_rsrc(X, (size_t)np * Rp * sizeof(float));
        ell_gather_bin<LG, WMAX, true>(wb, np, row0, l64, X, sh.stage[wave], xs, tile0,
                                       [&](int row, int c4, float4 acc) {
            const float4 own = xt_own<LG>(xs, xrs, row, tile0, c4);
            Q4[(size_t)row * LPR + c4] = acc;
            v[0][0] += (double)own.x * (double)acc.x; v[0][1] += (double)own.y * (double)acc.y;
            v[0][2] += (double)own.z * (double)acc.z; v[0][3] += (double)own.w * (double)acc.w;
            v[1][0] += (double)own.x; v[1][1] += (double)own.y; v[1][2] += (double)own.z; v[1][3] += (double)own.w;
        });
        block_partials4_buf<LG, 2>(J, s, t.z, ch, v, sh.red);
    }
}

template <int LG, int WMAX, int U>
__global__ void __launch_bounds__(kEwBlock) k_pcg_spmv(const JobDev* __restrict__ jobs,
                                                       const int4* __restrict__ tiles, int k4) {
    constexpr int Rp = 1 << LG;
    constexpr int G = 64 / Rp;                       // rows per wavefront step
    constexpr int STEPS = kSpmvRowsPerWave / G;      // steps per wavefront
    static_assert(STEPS % U == 0, "unroll must divide the steps");
    const int4 t = tiles[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    if (!slice_active(J, s)) return;
    const int wave = threadIdx.x >> 6, l64 = threadIdx.x & 63;
    const int lane = l64 & (Rp - 1), sub = l64 >> LG;
    const int np = J.np;
    const int row0 = t.z * kTileRows + wave * kSpmvRowsPerWave;
    const float al = J.alpha[s];
    const MPF_GLOBAL int32_t* ids = gp(J.ids);
    const MPF_GLOBAL float* Sv = gp((const float*)J.s);
    const MPF_GLOBAL float* Rv = gp((const float*)J.r);
    int idr[WMAX];
    float vr[WMAX];
#pragma unroll
    for (int q = 0; q < WMAX; ++q) { idr[q] = 0; vr[q] = 0.0f; }
    if (l64 < kSpmvRowsPerWave) {
#pragma unroll
        for (int q = 0; q < WMAX; ++q) {
            const size_t e = (size_t)q * np + row0 + l64;
            idr[q] = ids[e];
            vr[q] = Sv[e] + al * Rv[e];
        }
    }
    const float* Pbase = pbuf(J, k4);         // p of this iteration (k4 = k % kRing)
    const MPF_GLOBAL float* P = gp(Pbase) + (size_t)s * np * Rp;
    MPF_GLOBAL float* Q = gp(J.q) + (size_t)s * np * Rp;
    // gathers go through a buffer resource: 32-bit offsets, reads past the slice return 0
    // without a memory access (used for the zero-valued padding slots)
    const uint32_t range = (uint32_t)((size_t)np * Rp * sizeof(float));
    const __amdgpu_buffer_rsrc_t rs = uniform_rsrc(Pbase + (size_t)s * np * Rp, (size_t)range);
    double pq = 0.0, ps = 0.0;
    for (int i = 0; i < STEPS; i += U) {
        int src[U];
        float own[U], acc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            src[u] = (i + u) * G + sub;
            own[u] = P[(size_t)(row0 + src[u]) * Rp + lane];
            acc[u] = 0.0f;
        }
#pragma unroll
        for (int q = 0; q < WMAX; ++q) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int id = bcast_i<LG>(idr[q], src[u]);
                const float v = bcast_f<LG>(vr[q], src[u]);
                float x;
                if constexpr (LG == 6)   // id is wave-uniform: scalar offset, shared lane offset
                    // (the range check applies to the vector offset, so padding moves that)
                    x = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
                        rs, v != 0.0f ? (uint32_t)lane * 4u : range, v != 0.0f ? id * (Rp * 4) : 0, 0));
                else
                    x = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(
                        rs, v != 0.0f ? (uint32_t)(id * Rp + lane) * 4u : range, 0, 0));
                acc[u] += v * x;
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            Q[(size_t)(row0 + src[u]) * Rp + lane] = acc[u];
            pq += (double)own[u] * (double)acc[u];
            ps += (double)own[u];
        }
    }
    const int ch[2] = {0, 1};
    const double v[2] = {pq, ps};
    block_partials<2, kEwBlock>(J, s, t.z, ch, v);
}

// Iteration i runs three launches: SpMV (tile partials of p.q and 1^T p), update and
// pupdate.  Each update / pupdate block sums the partials it needs over all tiles of
// its (job, slice) itself, in the same order as every other block (bit-identical
// scalars, no cross-block synchronisation); tile 0's block also keeps the solver state.
//
// update: alpha = rho / p.q (with dcFree q = A p + 1 1^T p, so p.q += (1^T p)^2);
// r -= alpha q (a zero p.q zeroes r, and x in pupdate, updateVectorKernel
// conjugate_gradient_kernel.cu:77-89); tile partials of r.z (z = D^-1 r) and r.r.
// x += alpha p is left to pupdate, which reads p anyway.  Tile 0's block records
// whether the slice is live in this iteration (state [3]): pupdate's blocks test that
// instead of [0], which its stop rule may clear while they run, so that every tile of a
// slice that stops in this iteration still takes its last x update.
template <int LG>
__global__ void __launch_bounds__(kUpdBlock, 6) k_pcg_update4(const JobDev* __restrict__ jobs,
                                                           const int4* __restrict__ tiles, int iter) {
    constexpr int Rp = 1 << LG, LPR = Rp / 4, ROWS = kUpdBlock / LPR;
    const int par = iter & 1, k4 = iter & (kRing - 1);
    const int4 t = tiles[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    const bool live = slice_active(J, s);
    if (t.z == 0 && threadIdx.x == 0) J.st[s * kStInts + 3] = live ? 1 : 0;
    if (!live) return;
    __shared__ double sh_al[Rp], sh_dc[Rp];
    __shared__ int sh_zero[Rp];
    {
        // p.q and 1^T p arrive as one partial per kChunkTiles-tile chunk (k_pcg_chunks)
        double out[2] = {0.0, 0.0};
        const int cnt = coarse_chunks(J);
        if (s == J.dc) {
            const int ch[2] = {5, 6};
            reduce_tiles<2, kUpdBlock>(J, s, ch, out, cnt);
        } else {
            const int ch[1] = {5};
            double o[1];
            reduce_tiles<1, kUpdBlock>(J, s, ch, o, cnt);
            out[0] = o[0];
        }
        if ((int)threadIdx.x < Rp) {
            const int lane = threadIdx.x;
            const double ps = s == J.dc ? out[1] : 0.0;
            const double pq = out[0] + ps * ps;
            const double al = pq != 0.0 ? sc_ptr(J, rho_slot(par), s)[lane] / pq : 0.0;
            sh_al[lane] = al;
            sh_dc[lane] = ps;
            sh_zero[lane] = pq == 0.0;
            if (t.z == 0) {
                // alpha and p.q per ring slot: pupdate applies kRing iterations' x
                // updates at once
                sc_ptr(J, SC_PQ0 + k4, s)[lane] = pq;
                sc_ptr(J, SC_PSUM, s)[lane] = ps;
                sc_ptr(J, SC_PA0 + k4, s)[lane] = al;
            }
        }
        __syncthreads();
    }
    const int c4 = threadIdx.x % LPR, r0 = threadIdx.x / LPR;
    float a[4], dcv[4];
    bool zero[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int col = 4 * c4 + j;
        zero[j] = sh_zero[col] != 0;
        a[j] = (float)sh_al[col];
        dcv[j] = (float)sh_dc[col];
    }
    const size_t base = (size_t)s * J.np * LPR;
    const MPF_GLOBAL float* dinv = gp((const float*)J.dinv) + (size_t)s * J.np;
    const MPF_GLOBAL float4* Q = (const MPF_GLOBAL float4*)gp((const float*)J.q) + base;
    MPF_GLOBAL float4* Rr = (MPF_GLOBAL float4*)gp(J.rv) + base;
    double v[2][4] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
    // W^T r per aggregate (read in FP32 by the coarse kernel): summed in FP32 within a
    // wavefront, in FP64 across the block
    float g[kAggPerTile][4];
#pragma unroll
    for (int q = 0; q < kAggPerTile; ++q)
#pragma unroll
        for (int j = 0; j < 4; ++j) g[q][j] = 0.0f;
    for (int row = t.z * kTileRows + r0; row < (t.z + 1) * kTileRows; row += ROWS) {
        const size_t i = (size_t)row * LPR + c4;
        const float4 q4 = Q[i];
        const float4 r4 = Rr[i];
        const float d = dinv[row];
        const float dc = row < J.n ? 1.0f : 0.0f;
        const int qa = row_agg(J, row);
        float qv[4] = {q4.x, q4.y, q4.z, q4.w}, rv[4] = {r4.x, r4.y, r4.z, r4.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            rv[j] = zero[j] ? 0.0f : rv[j] - a[j] * (qv[j] + dc * dcv[j]);
            const float z = d * rv[j];
            v[0][j] += (double)rv[j] * (double)z;
            v[1][j] += (double)rv[j] * (double)rv[j];
#pragma unroll
            for (int q = 0; q < kAggPerTile; ++q) g[q][j] += q == qa ? rv[j] : 0.0f;
        }
        Rr[i] = make_float4(rv[0], rv[1], rv[2], rv[3]);
    }
    // channels agg_channel(q): W^T r of the tile's aggregates (two-level preconditioner),
    // with a coarse level also c_a (W^T r)_a in FP32 for the coarse kernel's staging
    const int ch[2] = {2, 3};
    const int chf[kAggPerTile] = {agg_channel(0), agg_channel(1), agg_channel(2), agg_channel(3)};
    MPF_GLOBAL float* fo = nullptr;
    float fs[kAggPerTile] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (J.cinv) {
        fo = gp(J.cg) + ((size_t)s * kAggPerTile * J.ntiles + kAggPerTile * t.z) * Rp;
#pragma unroll
        for (int q = 0; q < kAggPerTile; ++q) fs[q] = J.csc[kAggPerTile * t.z + q];
    }
    block_partials4f<LG, 2, kAggPerTile, kUpdBlock>(J, s, t.z, ch, v, chf, g, fo, fs);
}

template <int LG>
constexpr int coarse_threads() { return (1 << LG) * kCoarseRows > 1024 ? 1024 : (1 << LG) * kCoarseRows; }

// After the SpMV: the tile partials p.q (channel 0) and 1^T p (channel 1) summed per
// kChunkTiles-tile chunk in a fixed order into channels 5 and 6, so every update block
// reduces a handful of chunk partials instead of every tile's.  One block per (chunk,
// job, slice).
template <int LG>
__global__ void __launch_bounds__(coarse_threads<LG>()) k_pcg_chunks(const JobDev* __restrict__ jobs,
                                                                     const int2* __restrict__ js) {
    constexpr int Rp = 1 << LG, NT = coarse_threads<LG>(), G = NT / Rp;
    const int2 pr = js[blockIdx.y];
    const JobDev& J = jobs[pr.x];
    const int s = pr.y;
    if (J.Rp != Rp || !slice_active(J, s)) return;
    const int nt = J.ntiles, t0 = blockIdx.x * kChunkTiles;
    if (t0 >= nt) return;
    const int te = min(t0 + kChunkTiles, nt);
    const int lane = threadIdx.x & (Rp - 1), g = threadIdx.x >> LG;
    double a = 0.0, b = 0.0;
    for (int t = t0 + g; t < te; t += G) {
        a += part_ptr(J, 0, s, t)[lane];
        b += part_ptr(J, 1, s, t)[lane];
    }
    __shared__ double red[2][NT];
    red[0][threadIdx.x] = a;
    red[1][threadIdx.x] = b;
    __syncthreads();
    if (g == 0) {
        double x = 0.0, y = 0.0;
        for (int gg = 0; gg < G; ++gg) { x += red[0][gg * Rp + lane]; y += red[1][gg * Rp + lane]; }
        part_ptr(J, 5, s, blockIdx.x)[lane] = x;
        part_ptr(J, 6, s, blockIdx.x)[lane] = y;
    }
}

// Coarse correction of the two-level preconditioner, one block per (chunk of kCoarseRows
// coarse rows = kChunkTiles tiles, job, slice): y_a = c_a sum_u cinv_s[a][u] c_u g_u with
// g = W^T r (aggregate kAggPerTile t + q: channel agg_channel(q) of tile t), written to
// cy; then per chunk the partial sums of
//   rho = sum_t (r.D^-1 r)_t + sum_a g_a y_a   (from channel ch_rz)  -> channel 5
//   r.r = sum_t (r.r)_t                        (from channel ch_rr)  -> channel 6
// so the next kernel reduces one partial per chunk instead of one per tile.  The product
// Y[64 x Rp] = cinv_s[chunk rows][:] G[: x Rp] runs on the matrix cores
// (v_mfma_f32_16x16x4_f32: FP32 operands, FP32 accumulation).  Wavefront w
// owns coarse rows 16w..16w+15 of the chunk and all Rp columns (Rp / 16 accumulator
// tiles).  G is staged through LDS kCoarseK aggregates at a time; each lane reads four
// consecutive cinv entries of its row with one 128-bit load and feeds them to four MFMAs
// (step j takes k = 4 (lane / 16) + j), with the matching G rows read from LDS.
typedef float v4f __attribute__((ext_vector_type(4)));
constexpr int kCoarseK = 64;
template <int LG>
__global__ void __launch_bounds__(256) k_pcg_coarse_mfma(const JobDev* __restrict__ jobs,
                                                         const int2* __restrict__ js, int ch_rz, int ch_rr,
                                                         int in_iteration) {
    constexpr int Rp = 1 << LG, CT = Rp / 16, GLD = Rp + 4;
    static_assert(kCoarseRows == 64, "four wavefronts of 16 coarse rows");
    const int2 pr = js[blockIdx.y];
    const JobDev& J = jobs[pr.x];
    const int s = pr.y;
    if (J.Rp != Rp || (in_iteration && !slice_active(J, s)) || (!in_iteration && !J.cinv)) return;
    const int nt = J.ntiles, na = kAggPerTile * nt, ld = J.cld;
    const int a0 = blockIdx.x * kCoarseRows;
    if (a0 >= na) return;
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63, r16 = l & 15, kb = l >> 4;
    __shared__ float gs[kCoarseK][GLD];
    __shared__ double red[2][16][Rp];
    const MPF_GLOBAL float* csc = gp((const float*)J.csc);
    v4f acc[CT];
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) acc[ct] = v4f{0.f, 0.f, 0.f, 0.f};
    // rows past na compute a copy of the last row, discarded
    const int arow = min(a0 + 16 * wave + r16, na - 1);
    const MPF_GLOBAL float* crow = J.cinv ? gp(J.cinv) + ((size_t)s * na + arow) * ld : nullptr;
    // software pipeline: the next step's cinv fragments and G values are loaded while the
    // current step's are staged and multiplied, so that each wavefront waits for the
    // inverse (read from L2 or the MALL) once per step of kCoarseK aggregates, not four times
    constexpr int NK = kCoarseK / 16, NS = kCoarseK * Rp / 256;
    static_assert(kCoarseK * Rp % 256 == 0, "G staging splits evenly over the block");
    float4 cn[NK];
    float gn[NS];
    auto load_step = [&](int u0) {
#pragma unroll
        for (int q = 0; q < NK; ++q) {
            const int u = u0 + 16 * q + 4 * kb;   // cinv rows are zero-padded to ld, a multiple of 4
            cn[q] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (u < ld) cn[q] = *(const MPF_GLOBAL float4*)(crow + u);
        }
        const int un = min(kCoarseK, na - u0);
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            const int i = threadIdx.x + 256 * q, u = i / Rp, c = i % Rp, a = u0 + u;
            // in the iteration the update kernel left c_a g_a in FP32; at the start only
            // the FP64 partials exist
            gn[q] = u >= un ? 0.0f
                    : in_iteration ? gp((const float*)J.cg)[((size_t)s * na + a) * Rp + c]
                                   : (float)((double)csc[a] * part_ptr(J, agg_channel(a % kAggPerTile), s, a / kAggPerTile)[c]);
        }
    };
    if (J.cinv) load_step(0);
    for (int u0 = 0; J.cinv && u0 < na; u0 += kCoarseK) {
        float4 cc[NK];
#pragma unroll
        for (int q = 0; q < NK; ++q) cc[q] = cn[q];
        __syncthreads();   // the previous step's G has been read
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            const int i = threadIdx.x + 256 * q;
            gs[i / Rp][i % Rp] = gn[q];
        }
        if (u0 + kCoarseK < na) load_step(u0 + kCoarseK);
        __syncthreads();
#pragma unroll
        for (int k0 = 0; k0 < kCoarseK; k0 += 16) {
            const float4 c4 = cc[k0 / 16];
            const float cv[4] = {c4.x, c4.y, c4.z, c4.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float* grow = gs[k0 + 4 * kb + j];
#pragma unroll
                for (int ct = 0; ct < CT; ++ct)
                    acc[ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(cv[j], grow[16 * ct + r16], acc[ct], 0, 0, 0);
            }
        }
    }
    // lane (r16, kb) holds Y[16 w + 4 kb + v][16 ct + r16], v = 0..3
    double rz[CT], rr[CT];
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
        rz[ct] = 0.0;
        rr[ct] = 0.0;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int a = a0 + 16 * wave + 4 * kb + v, t = a / kAggPerTile, col = 16 * ct + r16;
            if (a < na) {
                if (a % kAggPerTile == 0) {
                    rz[ct] += part_ptr(J, ch_rz, s, t)[col];
                    rr[ct] += part_ptr(J, ch_rr, s, t)[col];
                }
                if (J.cinv) {
                    const float y = csc[a] * acc[ct][v];
                    gp(J.cy)[((size_t)s * na + a) * Rp + col] = y;
                    rz[ct] += part_ptr(J, agg_channel(a % kAggPerTile), s, t)[col] * (double)y;
                }
            }
        }
    }
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
        red[0][4 * wave + kb][16 * ct + r16] = rz[ct];
        red[1][4 * wave + kb][16 * ct + r16] = rr[ct];
    }
    __syncthreads();
    if ((int)threadIdx.x < 2 * Rp) {
        const int c = threadIdx.x / Rp, col = threadIdx.x % Rp;
        double v = 0.0;
        for (int g = 0; g < 16; ++g) v += red[c][g][col];
        part_ptr(J, 5 + c, s, blockIdx.x)[col] = v;
    }
}

// Sc^-1/2 per aggregate: Sc_a = sum of S_ii(gamma) over the aggregate's real rows / the
// same sum at sigma = 1 (harmonic independent; one block per (job, tile) of slice 0, the
// tile's kAggPerTile aggregates)
__global__ void __launch_bounds__(kTileRows) k_coarse_scale(const JobDev* __restrict__ jobs,
                                                           const int4* __restrict__ tiles,
                                                           const double* __restrict__ d1) {
    const int4 t = tiles[blockIdx.x];
    const JobDev& J = jobs[t.x];
    if (t.y != 0 || !J.csc) return;
    const int row = t.z * kTileRows + threadIdx.x;
    double v = 0.0;
    if (row < J.n) {
        const size_t e = (size_t)J.diag_slot[row] * J.np + row;
        v = (double)J.s[e] + (double)J.s_lo[e];
    }
    // per aggregate of the tile (the rows with aggq = q): tree sums over the tile
    const int qa = row_agg(J, row);
    __shared__ double red[kAggPerTile][kTileRows];
#pragma unroll
    for (int q = 0; q < kAggPerTile; ++q) red[q][threadIdx.x] = q == qa ? v : 0.0;
    __syncthreads();
    for (int h = kTileRows / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
#pragma unroll
            for (int q = 0; q < kAggPerTile; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + h];
        __syncthreads();
    }
    if ((int)threadIdx.x < kAggPerTile) {
        const int a = kAggPerTile * t.z + threadIdx.x;
        const double sc = d1[a] > 0.0 ? red[threadIdx.x][0] / d1[a] : 0.0;
        J.csc[a] = sc > 0.0 ? (float)(1.0 / sqrt(sc)) : 0.0f;
    }
}

// pupdate: rho' = r.z, beta = rho' / rho (rho == 0 zeroes p), p = z + beta p, written to
// the next ring buffer (p of iteration k in buffer k % kRing).  x moves every kRing
// iterations: after iteration k = kRing - 1 (mod kRing), x takes the updates
// alpha_i p_i of the last kRing iterations in order (alpha and p.q as the update kernels
// left them; a zero p.q zeroes x, as updateVectorKernel does at every iteration); a slice
// that stops elsewhere in the cycle gets the updates it is owed from k_pcg_flush.
// Tile 0's
// block also sums r.r, applies the stop rule (ConjugateGradient::solve,
// src/numeric/conjugate_gradient.cpp:118-150) and counts the step.  The blocks test the
// slice state the update kernel recorded for this iteration, so a slice that stops here
// still has every tile's x updated (its p is updated too and not read again).
template <int LG, bool XUP>
__global__ void __launch_bounds__(kUpdBlock, 6) k_pcg_pupdate4(const JobDev* __restrict__ jobs,
                                                            const int4* __restrict__ tiles, int iter,
                                                            int* __restrict__ active) {
    constexpr int Rp = 1 << LG, LPR = Rp / 4, ROWS = kUpdBlock / LPR;
    const int par = iter & 1, k4 = iter & (kRing - 1);
    const int4 t = tiles[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    if (gp(J.st)[s * kStInts + 3] == 0) return;
    const bool book = t.z == 0;
    __shared__ double sh_b[Rp];
    __shared__ int sh_bz[Rp];
    __shared__ int not_done;
    // two-level: the coarse corrections y_a of the tile's kAggPerTile aggregates (zero without)
    __shared__ float4 sh_y[kAggPerTile][LPR];
    if ((int)threadIdx.x < kAggPerTile * LPR) {
        float4 y = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (J.cinv)
            y = ((const MPF_GLOBAL float4*)gp((const float*)J.cy))
                [((size_t)s * kAggPerTile * J.ntiles + kAggPerTile * t.z) * LPR + threadIdx.x];
        sh_y[threadIdx.x / LPR][threadIdx.x % LPR] = y;
    }
    {
        double out[2] = {0.0, 0.0};
        // rho and r.r come as one partial per kChunkTiles-tile chunk (k_pcg_coarse_mfma,
        // channels 5, 6)
        const int cnt = coarse_chunks(J);
        const int c_rz = 5, c_rr = 6;
        if (book) {
            const int ch[2] = {c_rz, c_rr};
            reduce_tiles<2, kUpdBlock>(J, s, ch, out, cnt);
        } else {
            const int ch[1] = {c_rz};
            double o[1];
            reduce_tiles<1, kUpdBlock>(J, s, ch, o, cnt);
            out[0] = o[0];
        }
        if (threadIdx.x == 0) not_done = 0;
        __syncthreads();
        if ((int)threadIdx.x < Rp) {
            const int lane = threadIdx.x;
            const double rho = out[0];
            const double rho_old = sc_ptr(J, rho_slot(par), s)[lane];
            sh_b[lane] = rho_old != 0.0 ? rho / rho_old : 0.0;
            sh_bz[lane] = rho_old == 0.0;
            if (book) {
                const double rr = out[1];
                const double bb = sc_ptr(J, SC_BB, s)[lane];
                sc_ptr(J, SC_RR, s)[lane] = rr;
                sc_ptr(J, rho_slot(par ^ 1), s)[lane] = rho;
                sc_ptr(J, SC_BETA, s)[lane] = sh_b[lane];
                if (J.rule == MPF_STOP_REFERENCE) {
                    if (lane == J.col0) {
                        // column 0 of the pattern only: sqrt(rho / ref) >= eps keeps
                        // iterating; NaN (ref == 0) stops
                        const double crit = fabs(sqrt(rho / bb));
                        if (crit >= (double)FLT_EPSILON) not_done = 1;
                    }
                } else if (lane < J.R) {
                    if (bb > 0.0 && rr > J.tol * J.tol * bb) atomicOr(&not_done, 1);
                }
                // a NaN residual with a nonzero right-hand side is a breakdown
                if (lane < J.R && bb > 0.0 && rr != rr) atomicOr(&not_done, 2);
            }
        }
        __syncthreads();
        if (book && threadIdx.x == 0) {
            const int steps = J.st[s * kStInts + 1] + 1;
            J.st[s * kStInts + 1] = steps;
            if (not_done & 2) {
                J.st[s * kStInts + 0] = 0;
                J.st[s * kStInts + 2] = 1;
                atomicSub(active, 1);
            } else if (!not_done) {
                J.st[s * kStInts + 0] = 0;
                atomicSub(active, 1);
            } else if (steps >= J.maxit) {
                J.st[s * kStInts + 0] = 0;
                J.st[s * kStInts + 2] = 1;
                atomicSub(active, 1);
            }
        }
    }
    const int c4 = threadIdx.x % LPR, r0 = threadIdx.x / LPR;
    float b[4];
    bool zero[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        b[j] = (float)sh_b[4 * c4 + j];
        zero[j] = sh_bz[4 * c4 + j] != 0;
    }
    const size_t base = (size_t)s * J.np * LPR;
    const MPF_GLOBAL float* dinv = gp((const float*)J.dinv) + (size_t)s * J.np;
    const MPF_GLOBAL float4* Rr = (const MPF_GLOBAL float4*)gp((const float*)J.rv) + base;
    // p_k is read from this iteration's buffer, p_{k+1} written to the next one
    const MPF_GLOBAL float4* Pk = (const MPF_GLOBAL float4*)gp((const float*)pbuf(J, k4)) + base;
    MPF_GLOBAL float4* Pn = (MPF_GLOBAL float4*)gp(pbuf(J, (k4 + 1) & (kRing - 1))) + base;
    // two-level: z = D^-1 r + y_a on the real rows of aggregate a (a = kAggPerTile t + q
    // for the rows of tile t with aggq = q)
    if constexpr (XUP) {
        // k4 == kRing - 1: the cycle's x updates, oldest first; p of the cycle's first iteration sits where
        // p_{k+1} goes and is read before the write
        float a[kRing][4];
        bool xz[kRing][4];
#pragma unroll
        for (int q = 0; q < kRing; ++q)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                a[q][j] = (float)sc_ptr(J, SC_PA0 + q, s)[4 * c4 + j];
                xz[q][j] = sc_ptr(J, SC_PQ0 + q, s)[4 * c4 + j] == 0.0;
            }
        MPF_GLOBAL float4* X = (MPF_GLOBAL float4*)gp(J.x) + base;
        const MPF_GLOBAL float4* Pq[kRing - 1];
#pragma unroll
        for (int q = 0; q < kRing - 1; ++q) Pq[q] = (const MPF_GLOBAL float4*)gp((const float*)pbuf(J, q)) + base;
        for (int row = t.z * kTileRows + r0; row < (t.z + 1) * kTileRows; row += ROWS) {
            const size_t i = (size_t)row * LPR + c4;
            const float4 r4 = Rr[i], p4 = Pk[i], x4 = X[i];
            float4 pq4[kRing];
#pragma unroll
            for (int q = 0; q < kRing - 1; ++q) pq4[q] = Pq[q][i];
            pq4[kRing - 1] = p4;
            float xv[4] = {x4.x, x4.y, x4.z, x4.w};
#pragma unroll
            for (int q = 0; q < kRing; ++q) {
                const float pv[4] = {pq4[q].x, pq4[q].y, pq4[q].z, pq4[q].w};
#pragma unroll
                for (int j = 0; j < 4; ++j) xv[j] = xz[q][j] ? 0.0f : xv[j] + a[q][j] * pv[j];
            }
            X[i] = make_float4(xv[0], xv[1], xv[2], xv[3]);
            const float d = dinv[row];
            const float w = row < J.n ? 1.0f : 0.0f;
            const float4 y4 = sh_y[row_agg(J, row)][c4];
            Pn[i] = make_float4(zero[0] ? 0.0f : d * r4.x + w * y4.x + b[0] * p4.x,
                                zero[1] ? 0.0f : d * r4.y + w * y4.y + b[1] * p4.y,
                                zero[2] ? 0.0f : d * r4.z + w * y4.z + b[2] * p4.z,
                                zero[3] ? 0.0f : d * r4.w + w * y4.w + b[3] * p4.w);
        }
        return;
    }
    for (int row = t.z * kTileRows + r0; row < (t.z + 1) * kTileRows; row += ROWS) {
        const size_t i = (size_t)row * LPR + c4;
        const float4 r4 = Rr[i], p4 = Pk[i];
        const float d = dinv[row];
        const float w = row < J.n ? 1.0f : 0.0f;
        const float4 y4 = sh_y[row_agg(J, row)][c4];
        Pn[i] = make_float4(zero[0] ? 0.0f : d * r4.x + w * y4.x + b[0] * p4.x,
                            zero[1] ? 0.0f : d * r4.y + w * y4.y + b[1] * p4.y,
                            zero[2] ? 0.0f : d * r4.z + w * y4.z + b[2] * p4.z,
                            zero[3] ? 0.0f : d * r4.w + w * y4.w + b[3] * p4.w);
    }
}

// The x updates owed by a slice whose step count n is not a multiple of kRing: those of
// its last n % kRing iterations (ring slots 0 ..), in order, as pupdate would apply them.
__global__ void __launch_bounds__(kUpdBlock) k_pcg_flush(const JobDev* __restrict__ jobs,
                                                         const int4* __restrict__ tiles) {
    const int4 t = tiles[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    const int m = gp(J.st)[s * kStInts + 1] & (kRing - 1);
    if (m == 0) return;
    // 128-bit accesses; a thread keeps the same 4 lanes (kUpdBlock is a multiple of Rp / 4),
    // so their alpha and p.q of the owed iterations are read once
    const int LPR = J.Rp / 4, c4 = threadIdx.x % LPR;
    float a[kRing - 1][4];
    bool zq[kRing - 1][4];
#pragma unroll
    for (int q = 0; q < kRing - 1; ++q)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            a[q][j] = q < m ? (float)sc_ptr(J, SC_PA0 + q, s)[4 * c4 + j] : 0.0f;
            zq[q][j] = q < m && sc_ptr(J, SC_PQ0 + q, s)[4 * c4 + j] == 0.0;
        }
    const size_t base = ((size_t)s * J.np + (size_t)t.z * kTileRows) * LPR;
    MPF_GLOBAL float4* X = (MPF_GLOBAL float4*)gp(J.x) + base;
    for (int i = threadIdx.x; i < kTileRows * LPR; i += kUpdBlock) {
        float4 x = X[i];
        float xv[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
        for (int q = 0; q < kRing - 1; ++q) {
            if (q >= m) break;
            const float4 p = ((const MPF_GLOBAL float4*)gp((const float*)pbuf(J, q)) + base)[i];
            const float pv[4] = {p.x, p.y, p.z, p.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) xv[j] = zq[q][j] ? 0.0f : xv[j] + a[q][j] * pv[j];
        }
        X[i] = make_float4(xv[0], xv[1], xv[2], xv[3]);
    }
}



// --------------------------------------------------------------------------------
// FP64 refinement: residual r64 = beta F - (A64 + dc) x64 written as the FP32 RHS
// --------------------------------------------------------------------------------
__global__ void __launch_bounds__(kEwBlock) k_colsum64(const JobDev* __restrict__ jobs,
                                                       const int4* __restrict__ tiles) {
    const int4 t = tiles[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    const Ew L = ew_of(J);
    double a = 0.0;
    if (s == J.dc)
        for (int row = t.z * kTileRows + L.g; row < (t.z + 1) * kTileRows; row += L.G)
            a += J.x64[vec_index(J, s, row, L.lane)];
    const int ch[1] = {0};
    const double v[1] = {a};
    block_partials<1, kEwBlock>(J, s, t.z, ch, v);
}

__global__ void __launch_bounds__(kEwBlock) k_colsum64_reduce(const JobDev* __restrict__ jobs,
                                                              const int2* __restrict__ js) {
    const int2 t = js[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    const int ch[1] = {0};
    double out[1];
    reduce_tiles<1>(J, s, ch, out);
    if ((int)threadIdx.x < J.Rp)
        sc_ptr(J, SC_X0, s)[threadIdx.x] = (s == J.dc && J.dc_rank1) ? out[0] : 0.0;
}

// r = beta_c F - A_c x64 - dcs in FP64 for every lane of one (job, slice, tile) item:
// a thread owns two lanes of one row (128-bit gathers of x64), the block takes PR rows
// at a time, and the ELL entries of those rows are staged through LDS kResSlots slots at
// a time with coalesced loads (id and the FP64 value S + alpha R, each formed once per
// row and slot instead of once per lane).  The sums run over the slots in ELL order.
// XT: the tile's 256 rows of x64 and a row of zeros sit in LDS (64 KB at 32 lanes) and
// the stage holds byte keys as in the SpMV (xt_key): every slot adds an LDS read (zero for
// a neighbour outside the tile) and a buffer load (out of range, 0, for one inside), so
// only the neighbours outside the tile are gathered from memory; the sums are the same.
constexpr int kResSlots = 8;
template <int LG, bool XT>
__global__ void __launch_bounds__(kEwBlock) k_residual64(const JobDev* __restrict__ jobs,
                                                         const int4* __restrict__ tiles) {
    constexpr int Rp = 1 << LG, T = Rp / 2, PR = kEwBlock / T;
    constexpr int RB = Rp * 8;                              // bytes per row of x64
    constexpr int NXT = XT ? (kTileRows + 1) * T : 1;       // double2 elements of the LDS tile
    const int4 t4 = tiles[blockIdx.x];
    const JobDev& J = jobs[t4.x];
    if (J.Rp != Rp) return;   // the launcher checks that every job has 2^LG lanes
    const int s = t4.y;
    const double al = J.alpha64[s];
    const double xs = J.xscale64[s];
    const int t = threadIdx.x & (T - 1), rl = threadIdx.x / T;
    const int l0 = 2 * t;
    const double dcs0 = sc_ptr(J, SC_X0, s)[l0], dcs1 = sc_ptr(J, SC_X0, s)[l0 + 1];
    const MPF_GLOBAL double2* X2 = (const MPF_GLOBAL double2*)(gp(J.x64) + (size_t)s * J.np * Rp);
    float* out = J.rhs;
    __shared__ struct {
        int id[kResSlots][PR];      // XT: byte keys
        double v[kResSlots][PR];
    } se;
    __shared__ union {
        double2 xt[NXT];
        double red[2][PR][Rp];
    } sb;
    const int tile0 = t4.z * kTileRows;
    __amdgpu_buffer_rsrc_t rs;
    if constexpr (XT) {
        rs = uniform_rsrc(X2, (size_t)J.np * RB);
        constexpr int NL = kTileRows * T / kEwBlock;
        static_assert(kTileRows * T % kEwBlock == 0, "the tile's rows split evenly over the block");
        double2 xv[NL];
#pragma unroll
        for (int k = 0; k < NL; ++k) xv[k] = X2[(size_t)tile0 * T + threadIdx.x + k * kEwBlock];
#pragma unroll
        for (int k = 0; k < NL; ++k) sb.xt[threadIdx.x + k * kEwBlock] = xv[k];
        if ((int)threadIdx.x < T) sb.xt[kTileRows * T + threadIdx.x] = make_double2(0.0, 0.0);
    }
    double rr0 = 0.0, rr1 = 0.0, bb0 = 0.0, bb1 = 0.0;
    for (int base = tile0; base < tile0 + kTileRows; base += PR) {
        const int row = base + rl;
        // slots past a 16-row group's widest row are padding (value 0)
        int wp = J.width;
        if (J.gw) {
            wp = 0;
            for (int g = base >> 4; g < (base + PR) >> 4; ++g) wp = max(wp, (int)J.gw[g]);
        }
        const int wr = J.gw ? J.gw[row >> 4] : J.width;
        double a0 = 0.0, a1 = 0.0;
        for (int q0 = 0; q0 < wp; q0 += kResSlots) {
            __syncthreads();   // the previous step's entries have been read (and the tile staged)
            for (int i = threadIdx.x; i < kResSlots * PR; i += kEwBlock) {
                const int qq = i / PR, r2 = i % PR, q = q0 + qq;
                int id = XT ? -(kTileRows + 1) * RB : 0;
                double v = 0.0;
                if (q < wp) {
                    const size_t e = (size_t)q * J.np + base + r2;
                    id = J.ids[e];
                    if constexpr (XT) {
                        const int local = id - tile0;
                        id = (unsigned)local < (unsigned)kTileRows ? -(local + 1) * RB : id * RB;
                    }
                    v = ((double)J.s[e] + (double)J.s_lo[e]) + al * ((double)J.r[e] + (double)J.r_lo[e]);
                }
                se.id[qq][r2] = id;
                se.v[qq][r2] = v;
            }
            __syncthreads();
            const int qn = wr - q0;
#pragma unroll
            for (int qq = 0; qq < kResSlots; ++qq) {
                if (qq < qn) {
                    const double v = se.v[qq][rl];
                    double2 xv;
                    if constexpr (XT) {
                        const int key = se.id[qq][rl];
                        const v4u g = __builtin_amdgcn_raw_buffer_load_b128(rs, (uint32_t)(key + 16 * t), 0, 0);
                        double2 gd;
                        __builtin_memcpy(&gd, &g, sizeof(gd));
                        const uint32_t lo = min((uint32_t)(-key) - (uint32_t)RB, (uint32_t)(kTileRows * RB));
                        const double2 l = *reinterpret_cast<const double2*>(
                            reinterpret_cast<const char*>(sb.xt) + lo + 16 * t);
                        xv = make_double2(l.x + gd.x, l.y + gd.y);
                    } else {
                        xv = X2[(size_t)se.id[qq][rl] * T + t];
                    }
                    a0 += v * xv.x;
                    a1 += v * xv.y;
                }
            }
        }
        double b0 = 0.0, b1 = 0.0, r0 = 0.0, r1 = 0.0;
        if (row < J.n) {
            const double* F = J.F + s * J.fstride + (size_t)row * Rp;
            if (l0 < J.R) { b0 = xs * F[l0]; r0 = b0 - a0 - dcs0; }
            if (l0 + 1 < J.R) { b1 = xs * F[l0 + 1]; r1 = b1 - a1 - dcs1; }
        }
        *(float2*)(out + vec_index(J, s, row, l0)) = make_float2((float)r0, (float)r1);
        rr0 += r0 * r0; rr1 += r1 * r1;
        bb0 += b0 * b0; bb1 += b1 * b1;
    }
    // per-lane tile partials of r.r and b.b, summed over the rows in a fixed order (the
    // buffer shares LDS with the x64 tile: every gather of the block is done first)
    __syncthreads();
    sb.red[0][rl][l0] = rr0; sb.red[0][rl][l0 + 1] = rr1;
    sb.red[1][rl][l0] = bb0; sb.red[1][rl][l0 + 1] = bb1;
    __syncthreads();
    if ((int)threadIdx.x < 2 * Rp) {
        const int c = threadIdx.x / Rp, lane = threadIdx.x % Rp;
        double a = 0.0;
        for (int g = 0; g < PR; ++g) a += sb.red[c][g][lane];
        part_ptr(J, c, s, t4.z)[lane] = a;
    }
}

__global__ void __launch_bounds__(kEwBlock) k_residual64_reduce(const JobDev* __restrict__ jobs,
                                                                const int2* __restrict__ js) {
    const int2 t = js[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    const int ch[2] = {0, 1};
    double out[2];
    reduce_tiles<2>(J, s, ch, out);
    if ((int)threadIdx.x < J.Rp) {
        sc_ptr(J, SC_X1, s)[threadIdx.x] = out[0];
        sc_ptr(J, SC_X2, s)[threadIdx.x] = out[1];
    }
}

// x64 = x (first) or x64 += x
__global__ void __launch_bounds__(kEwBlock) k_x64_update(const JobDev* __restrict__ jobs,
                                                         const int4* __restrict__ tiles, int first) {
    const int4 t = tiles[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    // 128-bit FP32 reads of x, 2 x 128-bit FP64 accesses of x64 per 4 lanes
    const int LPR = J.Rp / 4;
    const size_t base = ((size_t)s * J.np + (size_t)t.z * kTileRows) * LPR;
    const MPF_GLOBAL float4* X = (const MPF_GLOBAL float4*)gp((const float*)J.x) + base;
    MPF_GLOBAL double2* D = (MPF_GLOBAL double2*)gp(J.x64) + 2 * base;
    for (int i = threadIdx.x; i < kTileRows * LPR; i += kEwBlock) {
        const float4 v = X[i];
        double2 lo = make_double2((double)v.x, (double)v.y), hi = make_double2((double)v.z, (double)v.w);
        if (!first) {
            const double2 a = D[2 * i], b = D[2 * i + 1];
            lo.x += a.x; lo.y += a.y; hi.x += b.x; hi.y += b.y;
        }
        D[2 * i] = lo;
        D[2 * i + 1] = hi;
    }
}

// --------------------------------------------------------------------------------
// launch wrappers
// --------------------------------------------------------------------------------
static inline void check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error(MPF_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
}

void launch_assemble(const JobDev* d_jobs, const int4* d_items, int nitems, hipStream_t st) {
    if (!nitems) return;
    k_assemble<<<nitems, kBlock, 0, st>>>(d_jobs, d_items);
    check_launch();
}

void launch_dinv(const JobDev* d_jobs, const int4* d_tiles, int ntiles, hipStream_t st) {
    k_dinv<<<ntiles, kBlock, 0, st>>>(d_jobs, d_tiles);
    check_launch();
}

template <int LG>
static void coarse_launch(const JobDev* d_jobs, const int2* d_js, int njs, int nchunk, int ch_rz, int ch_rr,
                          hipStream_t st) {
    k_pcg_coarse_mfma<LG><<<dim3(nchunk, njs), 256, 0, st>>>(d_jobs, d_js, ch_rz, ch_rr, ch_rz != 0);
}

// every lane count present in the list runs its own instantiation; blocks whose job has
// another lane count exit at once
template <int LG>
static void chunks_launch(const JobDev* d_jobs, const int2* d_js, int njs, int nchunk, hipStream_t st) {
    k_pcg_chunks<LG><<<dim3(nchunk, njs), coarse_threads<LG>(), 0, st>>>(d_jobs, d_js);
}

static void launch_coarse_all(const JobDev* d_jobs, const int2* d_js, int njs, int nchunk, int lg, int ch_rz,
                              int ch_rr, hipStream_t st) {
    if (lg == 6 || lg < 0) coarse_launch<6>(d_jobs, d_js, njs, nchunk, ch_rz, ch_rr, st);
    if (lg == 5 || lg < 0) coarse_launch<5>(d_jobs, d_js, njs, nchunk, ch_rz, ch_rr, st);
    if (lg == 4 || lg < 0) coarse_launch<4>(d_jobs, d_js, njs, nchunk, ch_rz, ch_rr, st);
}

void launch_pcg_init(const JobDev* d_jobs, const int4* d_tiles, int ntiles, const int2* d_js,
                     int njs, int* d_active, hipStream_t st, int nchunk) {
    k_pcg_init<<<ntiles, kEwBlock, 0, st>>>(d_jobs, d_tiles);
    if (nchunk > 0) {
        // the init list mixes lane counts: all three instantiations run
        launch_coarse_all(d_jobs, d_js, njs, nchunk, -1, 0, 1, st);
        k_pcg_init_p<<<ntiles, kEwBlock, 0, st>>>(d_jobs, d_tiles);
    }
    k_pcg_init_reduce<<<njs, kEwBlock, 0, st>>>(d_jobs, d_js, d_active);
    check_launch();
}

void launch_coarse_scale(const JobDev* d_jobs, const int4* d_tiles, int ntiles, const double* d1,
                         hipStream_t st) {
    k_coarse_scale<<<ntiles, kTileRows, 0, st>>>(d_jobs, d_tiles, d1);
    check_launch();
}

// SpMV variant 0: 128-bit gathers (k_pcg_spmv4); variant 1: 32-bit gathers with
// v_readlane broadcast (k_pcg_spmv, kept for comparison by the tuner); variant 2: 128-bit
// gathers, all harmonics of a (job, tile) per block (k_pcg_spmv4h; d_tiles lists
// (job, -, tile) items).
template <int LG>
static void spmv_dispatch(int wpad, int variant, const JobDev* d_jobs, const int4* d_tiles, int ntiles,
                          hipStream_t st, int k4) {
    constexpr int STEPS = kSpmvRowsPerWave / (64 >> LG);
    constexpr int U2 = STEPS >= 2 ? 2 : 1;
    if (variant == 2) {
        switch (wpad) {
            case 8:  k_pcg_spmv4h<LG, 8><<<ntiles, kEwBlock, 0, st>>>(d_jobs, d_tiles, k4); return;
            case 16: k_pcg_spmv4h<LG, 16><<<ntiles, kEwBlock, 0, st>>>(d_jobs, d_tiles, k4); return;
            case 24: k_pcg_spmv4h<LG, 24><<<ntiles, kEwBlock, 0, st>>>(d_jobs, d_tiles, k4); return;
            default: throw Error(MPF_EINVAL, "spmv: the harmonic-fused variant takes padded widths up to 24");
        }
    }
    if (variant == 0) {
        switch (wpad) {
            case 8:  k_pcg_spmv4<LG, 8><<<ntiles, kEwBlock, 0, st>>>(d_jobs, d_tiles, k4); return;
            case 16: k_pcg_spmv4<LG, 16><<<ntiles, kEwBlock, 0, st>>>(d_jobs, d_tiles, k4); return;
            case 24: k_pcg_spmv4<LG, 24><<<ntiles, kEwBlock, 0, st>>>(d_jobs, d_tiles, k4); return;
            default: break;   // wider rows use the 32-bit variant (no register spills)
        }
    }
    switch (wpad) {
        case 8:  k_pcg_spmv<LG, 8, U2><<<ntiles, kEwBlock, 0, st>>>(d_jobs, d_tiles, k4); break;
        case 16: k_pcg_spmv<LG, 16, U2><<<ntiles, kEwBlock, 0, st>>>(d_jobs, d_tiles, k4); break;
        case 24: k_pcg_spmv<LG, 24, U2><<<ntiles, kEwBlock, 0, st>>>(d_jobs, d_tiles, k4); break;
        case 32: k_pcg_spmv<LG, 32, 1><<<ntiles, kEwBlock, 0, st>>>(d_jobs, d_tiles, k4); break;
        case 40: k_pcg_spmv<LG, 40, 1><<<ntiles, kEwBlock, 0, st>>>(d_jobs, d_tiles, k4); break;
        case 48: k_pcg_spmv<LG, 48, 1><<<ntiles, kEwBlock, 0, st>>>(d_jobs, d_tiles, k4); break;
        case 56: k_pcg_spmv<LG, 56, 1><<<ntiles, kEwBlock, 0, st>>>(d_jobs, d_tiles, k4); break;
        case 64: k_pcg_spmv<LG, 64, 1><<<ntiles, kEwBlock, 0, st>>>(d_jobs, d_tiles, k4); break;
        default: throw Error(MPF_EINVAL, "spmv: unsupported padded ELL width");
    }
}

void launch_pcg_flush(const JobDev* d_jobs, const int4* d_tiles, int ntiles, hipStream_t st) {
    if (ntiles == 0) return;
    k_pcg_flush<<<ntiles, kUpdBlock, 0, st>>>(d_jobs, d_tiles);
    check_launch();
}

void launch_pcg_spmv(const JobDev* d_jobs, const int4* d_tiles, int ntiles, int lg, int wpad, int variant,
                     hipStream_t st, int k4) {
    if (lg == 6) spmv_dispatch<6>(wpad, variant, d_jobs, d_tiles, ntiles, st, k4);
    else if (lg == 5) spmv_dispatch<5>(wpad, variant, d_jobs, d_tiles, ntiles, st, k4);
    else if (lg == 4) spmv_dispatch<4>(wpad, variant, d_jobs, d_tiles, ntiles, st, k4);
    else throw Error(MPF_EINVAL, "spmv: unsupported lane count");
}

void launch_pcg_iter(const JobDev* d_jobs, const int4* d_tiles, int ntiles, int* d_active, int iter,
                     int lg, int wpad, int variant, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1,
                     const int2* d_js, int njs, int nchunk, const int4* d_ftiles, int nftiles) {
    if (ev0) MPF_HIP(hipEventRecord(ev0, st));
    if (variant == 2) {
        MPF_REQUIRE(d_ftiles, "pcg iteration: the harmonic-fused SpMV needs a (job, tile) list");
        if (nftiles) launch_pcg_spmv(d_jobs, d_ftiles, nftiles, lg, wpad, variant, st, iter & (kRing - 1));
    } else {
        launch_pcg_spmv(d_jobs, d_tiles, ntiles, lg, wpad, variant, st, iter & (kRing - 1));
    }
    if (ev1) MPF_HIP(hipEventRecord(ev1, st));
    MPF_REQUIRE(d_js && njs > 0 && nchunk > 0, "pcg iteration: no (job, slice) list");
    if (lg == 6) chunks_launch<6>(d_jobs, d_js, njs, nchunk, st);
    else if (lg == 5) chunks_launch<5>(d_jobs, d_js, njs, nchunk, st);
    else chunks_launch<4>(d_jobs, d_js, njs, nchunk, st);
    if (lg == 6) k_pcg_update4<6><<<ntiles, kUpdBlock, 0, st>>>(d_jobs, d_tiles, iter);
    else if (lg == 5) k_pcg_update4<5><<<ntiles, kUpdBlock, 0, st>>>(d_jobs, d_tiles, iter);
    else k_pcg_update4<4><<<ntiles, kUpdBlock, 0, st>>>(d_jobs, d_tiles, iter);
    // coarse correction where the job has one; chunk sums of r.z and r.r for every job
    launch_coarse_all(d_jobs, d_js, njs, nchunk, lg, 2, 3, st);
    // the last iteration of each direction-ring cycle also moves x
    if ((iter & (kRing - 1)) == kRing - 1) {
        if (lg == 6) k_pcg_pupdate4<6, true><<<ntiles, kUpdBlock, 0, st>>>(d_jobs, d_tiles, iter, d_active);
        else if (lg == 5) k_pcg_pupdate4<5, true><<<ntiles, kUpdBlock, 0, st>>>(d_jobs, d_tiles, iter, d_active);
        else k_pcg_pupdate4<4, true><<<ntiles, kUpdBlock, 0, st>>>(d_jobs, d_tiles, iter, d_active);
    } else {
        if (lg == 6) k_pcg_pupdate4<6, false><<<ntiles, kUpdBlock, 0, st>>>(d_jobs, d_tiles, iter, d_active);
        else if (lg == 5) k_pcg_pupdate4<5, false><<<ntiles, kUpdBlock, 0, st>>>(d_jobs, d_tiles, iter, d_active);
        else k_pcg_pupdate4<4, false><<<ntiles, kUpdBlock, 0, st>>>(d_jobs, d_tiles, iter, d_active);
    }
    check_launch();
}

void launch_residual64(const JobDev* d_jobs, const int4* d_tiles, int ntiles, const int2* d_js,
                       int njs, int lg, size_t slice_bytes, hipStream_t st) {
    k_colsum64<<<ntiles, kEwBlock, 0, st>>>(d_jobs, d_tiles);
    k_colsum64_reduce<<<njs, kEwBlock, 0, st>>>(d_jobs, d_js);
    // the LDS tile of x64 needs 31-bit byte offsets within a slice (and fits 32 lanes)
    const bool xt = slice_bytes < (size_t)(1u << 31) - (1u << 16);
    if (lg == 4 && xt) k_residual64<4, true><<<ntiles, kEwBlock, 0, st>>>(d_jobs, d_tiles);
    else if (lg == 4) k_residual64<4, false><<<ntiles, kEwBlock, 0, st>>>(d_jobs, d_tiles);
    else if (lg == 5 && xt) k_residual64<5, true><<<ntiles, kEwBlock, 0, st>>>(d_jobs, d_tiles);
    else if (lg == 5) k_residual64<5, false><<<ntiles, kEwBlock, 0, st>>>(d_jobs, d_tiles);
    else if (lg == 6) k_residual64<6, false><<<ntiles, kEwBlock, 0, st>>>(d_jobs, d_tiles);
    else throw Error(MPF_EINVAL, "residual: unsupported lane count");
    k_residual64_reduce<<<njs, kEwBlock, 0, st>>>(d_jobs, d_js);
    check_launch();
}

void launch_x64_update(const JobDev* d_jobs, const int4* d_tiles, int ntiles, int first,
                       hipStream_t st) {
    k_x64_update<<<ntiles, kEwBlock, 0, st>>>(d_jobs, d_tiles, first);
    check_launch();
}


}  // namespace mpf
