// Device helpers shared by the Krylov kernels (PCG in sparse.hip, BiCGSTAB in
// bicgstab.hip): tile partials and their fixed-order reductions, buffer resources and
// the 128-bit multi-RHS ELL gather core of the SpMV.
#pragma once

#include <cfloat>
#include <cmath>
#include <type_traits>

#include "jobs.h"

namespace mpf {

__device__ __forceinline__ MPF_GLOBAL double* part_ptr(const JobDev& J, int ch, int s, int tile) {
    return gp(J.part) + (((size_t)ch * J.C + s) * J.ntiles + tile) * J.Rp;
}

__device__ __forceinline__ MPF_GLOBAL double* sc_ptr(const JobDev& J, int f, int s) {
    return gp(J.sc) + ((size_t)f * J.C + s) * J.Rp;
}

__device__ __forceinline__ bool slice_active(const JobDev& J, int s) {
    return gp(J.st)[s * kStInts] != 0;
}

// Combine per-thread values of NV channels across the row groups of an NT-thread block
// (thread = group * Rp + lane) in a fixed order and write them as this tile's partials.
template <int NV, int NT>
__device__ __forceinline__ void block_partials(const JobDev& J, int s, int tile, const int* ch,
                                               const double (&v)[NV]) {
    __shared__ double red[NV][NT];
    const int Rp = J.Rp, lg = J.lg;
    const int lane = threadIdx.x & (Rp - 1), g = threadIdx.x >> lg, G = NT >> lg;
#pragma unroll
    for (int c = 0; c < NV; ++c) red[c][threadIdx.x] = v[c];
    __syncthreads();
    if (g == 0) {
#pragma unroll
        for (int c = 0; c < NV; ++c) {
            double a = red[c][lane];
            for (int gg = 1; gg < G; ++gg) a += red[c][gg * Rp + lane];
            part_ptr(J, ch[c], s, tile)[lane] = a;
        }
    }
}

// Sum NV partial channels over all tiles of (job, slice) in tile order with an
// NT-thread block; each row group sums a contiguous range of tiles (loads unrolled so
// many are in flight), then the groups are combined in order.  The result is valid in
// threads with g == 0 (lane = column).
// `count` (default: the job's tile count) partials are summed.
template <int NV, int NT = kEwBlock>
__device__ __forceinline__ void reduce_tiles(const JobDev& J, int s, const int* ch, double (&out)[NV],
                                             int count = -1) {
    __shared__ double red[NV][NT];
    const int Rp = J.Rp, lg = J.lg;
    const int lane = threadIdx.x & (Rp - 1), g = threadIdx.x >> lg, G = NT >> lg;
    const int nt = count < 0 ? J.ntiles : count;
    double a[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) a[c] = 0.0;
    const int per = (nt + G - 1) / G;
    const int t0 = g * per, t1 = min(nt, t0 + per);
    int t = t0;
    for (; t + 8 <= t1; t += 8) {
#pragma unroll
        for (int c = 0; c < NV; ++c) {
            const MPF_GLOBAL double* pp = part_ptr(J, ch[c], s, t) + lane;
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = pp[(size_t)u * Rp];
#pragma unroll
            for (int u = 0; u < 8; ++u) a[c] += v[u];
        }
    }
    for (; t < t1; ++t) {
#pragma unroll
        for (int c = 0; c < NV; ++c) a[c] += part_ptr(J, ch[c], s, t)[lane];
    }
#pragma unroll
    for (int c = 0; c < NV; ++c) red[c][threadIdx.x] = a[c];
    __syncthreads();
    if (g == 0) {
#pragma unroll
        for (int c = 0; c < NV; ++c) {
            double r = red[c][lane];
            for (int gg = 1; gg < G; ++gg) r += red[c][gg * Rp + lane];
            out[c] = r;
        }
    }
}


// streaming-kernel lane mapping (kEwBlock threads per 256-row tile)
struct Ew {
    int lane, g, G;
};
__device__ __forceinline__ Ew ew_of(const JobDev& J) {
    Ew e;
    e.lane = threadIdx.x & (J.Rp - 1);
    e.g = threadIdx.x >> J.lg;
    e.G = kEwBlock >> J.lg;
    return e;
}

// 4th dword of a raw buffer resource on gfx9-family GPUs (32-bit data, no swizzle)
constexpr int kBufferWord3 = 0x00020000;
typedef unsigned int v4u __attribute__((ext_vector_type(4)));

// Buffer resource built from wave-uniform values (readfirstlane keeps it in SGPRs, so
// buffer loads need no per-lane waterfall).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t uniform_rsrc(const void* base, size_t bytes) {
    const uint64_t b = (uint64_t)base;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)b);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(b >> 32));
    const int n = __builtin_amdgcn_readfirstlane((int)(bytes > 0x7fffffffu ? 0x7fffffffu : bytes));
    return __builtin_amdgcn_make_buffer_rsrc((void*)(((uint64_t)hi << 32) | lo), 0, n, kBufferWord3);
}

// broadcast a per-row register value held by lane `src` to the lanes of a row group
template <int LG>
__device__ __forceinline__ int bcast_i(int v, int src) {
    if constexpr (LG == 6) return __builtin_amdgcn_readlane(v, src);
    else return __builtin_amdgcn_ds_bpermute(src << 2, v);
}
template <int LG>
__device__ __forceinline__ float bcast_f(float v, int src) {
    return __int_as_float(bcast_i<LG>(__float_as_int(v), src));
}

// rows per wavefront in the SpMV: a 1024-thread block = 16 wavefronts x 16 rows = one
// 256-row tile
constexpr int kSpmvRowsPerWave = 16;

// Multi-RHS ELL gather of one wavefront's 16 rows: y_row = sum_q (S + alpha_s R)[row, q]
// X[id(row, q)] for the slice's Rp columns.  A lane owns 4 consecutive columns and Rp/4
// lanes cover one row, so one buffer_load_dwordx4 fetches 64/(Rp/4) neighbour rows.  The
// wavefront's ELL entries (id, value) sit in LDS as [slot][row]; every lane of a row reads
// the same entry (broadcast).  W (compile time, >= the group's widest row) fixes the
// unrolled slot loop.  Zero-valued slots (padding) gather through an offset past the
// buffer's range, which the buffer unit answers with 0 without a memory access.
// epi(row, c4, acc) consumes each row's 4-column result.
// With an LDS copy `xs` of the block's first kXtRows = 255 rows of X (rows tile0 ...)
// followed by one row of zeros (the tile's last row is read from memory like a neighbour
// outside the tile), neighbours inside the tile are read from LDS and only the others
// from memory.  The stage then holds a byte key per entry (xt_key): every slot issues one
// memory gather at offset key + 16 c4 (past the buffer's range, so answered with 0 and no
// memory access, when the key is negative) and one LDS read of row
// min(-key / (4 Rp) - 1, kXtRows) (the zero row for a neighbour read from memory), and
// adds the two: exactly one of them is the neighbour's value.
constexpr int kXtRows = kTileRows - 1;
template <int LG>
constexpr int xt_zero_offset() { return kXtRows * (1 << LG) * 4; }
template <int LG>
__device__ __forceinline__ int xt_key(int id, int tile0, bool real) {
    constexpr int RB = (1 << LG) * 4;
    if (!real) return -(kXtRows + 1) * RB;            // padding slot: the zero row, no gather
    const int local = id - tile0;
    return (unsigned)local < (unsigned)kXtRows ? -(local + 1) * RB : id * RB;
}
template <int LG, int W, bool XT, class Epi>
__device__ __forceinline__ void ell_gather_rows(int np, int row0, int l64, const float* X,
                                                const int2 (*ell)[kSpmvRowsPerWave], const float4* xs,
                                                int tile0, Epi&& epi) {
    constexpr int Rp = 1 << LG;
    constexpr int LPR = Rp / 4;                       // lanes per row
    constexpr int RPS = 64 / LPR;                     // rows per wavefront step
    constexpr int STEPS = kSpmvRowsPerWave / RPS;
    const int rsub = l64 / LPR, c4 = l64 % LPR;
    const uint32_t range = (uint32_t)((size_t)np * Rp * sizeof(float));
    const __amdgpu_buffer_rsrc_t rs = uniform_rsrc(X, (size_t)range);
    // at most 4 (LDS tile) or 8 gathers (16 or 32 VGPRs) are in flight per lane, which
    // keeps the kernels at 8 wavefronts per SIMD
#pragma unroll 1
    for (int st = 0; st < STEPS; ++st) {
        const int src = st * RPS + rsub;              // < 16
        float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if constexpr (XT) {
            // groups of G slots: the G memory gathers of a group are all issued before
            // any is used; every slot adds its memory gather and its LDS row (one of
            // them zero), so that no branch holds a gather back
            static_assert(W % 2 == 0, "slot bins are even");
            constexpr uint32_t RB = Rp * 4, ZERO = xt_zero_offset<LG>();
            const char* xsb = reinterpret_cast<const char*>(xs) + 16 * c4;
            auto group = [&](auto gsize, int q0) {
                constexpr int G = decltype(gsize)::value;
                v4u raw[G];
                int2 e[G];
#pragma unroll
                for (int u = 0; u < G; ++u) {
                    e[u] = ell[q0 + u][src];
                    raw[u] = __builtin_amdgcn_raw_buffer_load_b128(rs, (uint32_t)(e[u].x + 16 * c4), 0, 0);
                }
#pragma unroll
                for (int u = 0; u < G; ++u) {
                    const float v = __int_as_float(e[u].y);
                    const uint32_t lo = min((uint32_t)(-e[u].x) - RB, ZERO);
                    const float4 l = *reinterpret_cast<const float4*>(xsb + lo);
                    acc.x = fmaf(v, l.x + __uint_as_float(raw[u][0]), acc.x);
                    acc.y = fmaf(v, l.y + __uint_as_float(raw[u][1]), acc.y);
                    acc.z = fmaf(v, l.z + __uint_as_float(raw[u][2]), acc.z);
                    acc.w = fmaf(v, l.w + __uint_as_float(raw[u][3]), acc.w);
                }
            };
#pragma unroll 1
            for (int q0 = 0; q0 + 4 <= W; q0 += 4) group(std::integral_constant<int, 4>(), q0);
            if constexpr (W % 4 == 2) group(std::integral_constant<int, 2>(), W - 2);
            epi(row0 + src, c4, acc);
            continue;
        }
        // groups of 4 slots (a last group of 2 in the 10-slot bin): the group's gathers
        // are all issued before any is used
        auto group = [&](auto gsize, int q0) {
            constexpr int GM = decltype(gsize)::value;
            v4u raw[GM];
            float v[GM];
#pragma unroll
            for (int u = 0; u < GM; ++u) {
                const int2 e = ell[q0 + u][src];
                v[u] = __int_as_float(e.y);
                const uint32_t off = v[u] != 0.0f ? (uint32_t)(e.x * Rp + 4 * c4) * 4u : range;
                raw[u] = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
            }
#pragma unroll
            for (int u = 0; u < GM; ++u) {
                acc.x += v[u] * __uint_as_float(raw[u][0]);
                acc.y += v[u] * __uint_as_float(raw[u][1]);
                acc.z += v[u] * __uint_as_float(raw[u][2]);
                acc.w += v[u] * __uint_as_float(raw[u][3]);
            }
        };
        static_assert(W % 2 == 0, "slot bins are even");
#pragma unroll 1
        for (int q0 = 0; q0 + 4 <= W; q0 += 4) group(std::integral_constant<int, 4>(), q0);
        if constexpr (W % 4 == 2) group(std::integral_constant<int, 2>(), W - 2);
        epi(row0 + src, c4, acc);
    }
}

// LDS copy of a tile's first kXtRows rows of X for ell_gather_wave, then a row of zeros
// (kTileRows x Rp floats in all): element i of the tile's rows is staged as xt_value(i, x)
template <int LG>
using XTile = float4[kTileRows * (1 << LG) / 4];
template <int LG>
__device__ __forceinline__ float4 xt_value(int i, float4 x) {
    return i < kXtRows * ((1 << LG) / 4) ? x : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}
// the tile's own row `row` of X (buffer resource rs over the slice's X): the LDS row
// (zeros for the tile's last row) plus a buffer load that is out of range (0, no memory
// access) except for the last row -- no branch, so the LDS read stays an LDS instruction
template <int LG>
__device__ __forceinline__ float4 xt_own(const float4* xs, __amdgpu_buffer_rsrc_t rs, int row, int tile0, int c4) {
    constexpr int LPR = (1 << LG) / 4;
    const int local = row - tile0;
    const float4 l = xs[min(local, kXtRows) * LPR + c4];
    const uint32_t off = local == kXtRows ? (uint32_t)(row * LPR + c4) * 16u : 0xFFFFFFF0u;
    const v4u g = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
    return make_float4(l.x + __uint_as_float(g[0]), l.y + __uint_as_float(g[1]), l.z + __uint_as_float(g[2]),
                       l.w + __uint_as_float(g[3]));
}

// LDS for ell_gather_wave: one [WMAX][16] (id, value) table per wavefront of the block
template <int WMAX>
using EllStage = int2[kEwBlock / 64][WMAX][kSpmvRowsPerWave];

// Slot-count bin of a wavefront whose widest row has wl entries: bins 10, 12, 16, 20 and
// WMAX (P2 meshes: edge-midpoint rows 9 wide, vertex rows 19).
template <int WMAX>
__device__ __forceinline__ int ell_bin(int wl) {
    constexpr int WS = WMAX > 12 ? 12 : WMAX;
    constexpr int W10 = WMAX > 10 ? 10 : WMAX;
    return (wl <= W10 && W10 < WS) ? W10
         : (wl <= WS ? WS : (wl <= 16 && WMAX > 16 ? 16 : (wl <= 20 && WMAX > 20 ? 20 : WMAX)));
}

// ell_gather_rows with the slot loop unrolled for bin wb (= ell_bin<WMAX>)
template <int LG, int WMAX, bool XT, class Epi>
__device__ __forceinline__ void ell_gather_bin(int wb, int np, int row0, int l64, const float* X,
                                               const int2 (*ell)[kSpmvRowsPerWave], const float4* xs, int tile0,
                                               Epi&& epi) {
    constexpr int WS = WMAX > 12 ? 12 : WMAX;
    constexpr int W10 = WMAX > 10 ? 10 : WMAX;
    if (wb == W10 && W10 < WS) ell_gather_rows<LG, W10, XT>(np, row0, l64, X, ell, xs, tile0, epi);
    else if (wb == WS && WS < WMAX) ell_gather_rows<LG, WS, XT>(np, row0, l64, X, ell, xs, tile0, epi);
    else if (wb == 16 && WMAX > 16) ell_gather_rows<LG, (WMAX > 16 ? 16 : WMAX), XT>(np, row0, l64, X, ell, xs, tile0, epi);
    else if (wb == 20 && WMAX > 20) ell_gather_rows<LG, (WMAX > 20 ? 20 : WMAX), XT>(np, row0, l64, X, ell, xs, tile0, epi);
    else ell_gather_rows<LG, WMAX, XT>(np, row0, l64, X, ell, xs, tile0, epi);
}

// One wavefront's 16 rows of y = A_s X.  The slot count is binned by the wavefront's group
// width (ell_bin): a group whose rows are all at most 12 entries wide (P2 midpoint rows: 9)
// runs the 12-slot body.  The wavefront first stages its (id, S + alpha_s R) entries in LDS
// with coalesced loads spread over all 64 lanes; slots between the group width and the bin
// are zero.  Contains a block barrier: every wavefront of the block must call it.
// XT: xs is LDS room for the tile's rows of X, loaded here by the whole block.
template <int LG, int WMAX, bool XT = false, class Epi>
__device__ __forceinline__ void ell_gather_wave(const JobDev& J, int s, int row0, int l64, const float* X,
                                                EllStage<WMAX>& stage, Epi&& epi, float4* xs = nullptr) {
    constexpr int LPR = (1 << LG) / 4;
    const int tile0 = row0 / kTileRows * kTileRows;
    const int wave = threadIdx.x >> 6;
    const int wl = __builtin_amdgcn_readfirstlane((int)J.gw[row0 / kSpmvRowsPerWave]);
    const int wb = ell_bin<WMAX>(wl);
    const int np = J.np;
    const float al = J.alpha[s];
    const MPF_GLOBAL int32_t* ids = gp(J.ids);
    const MPF_GLOBAL float* Sv = gp((const float*)J.s);
    const MPF_GLOBAL float* Rv = gp((const float*)J.r);
    // every load of the prologue (the tile's rows of X, the wavefront's ELL entries) is
    // issued before any of them is written to LDS
    constexpr int NX = XT ? (kTileRows * LPR + kEwBlock - 1) / kEwBlock : 1;
    constexpr int NE = (WMAX * kSpmvRowsPerWave + 63) / 64;
    float4 xv[NX];
    int idv[NE];
    float sv[NE], rv[NE];
    const MPF_GLOBAL float4* X4 = (const MPF_GLOBAL float4*)gp(X) + (size_t)tile0 * LPR;
    if constexpr (XT) {
#pragma unroll
        for (int k = 0; k < NX; ++k) {
            const int i = threadIdx.x + k * kEwBlock;
            if (i < kTileRows * LPR) xv[k] = X4[i];
        }
    }
#pragma unroll
    for (int k = 0; k < NE; ++k) {
        const int e = l64 + 64 * k, q = e / kSpmvRowsPerWave, r = e % kSpmvRowsPerWave;
        idv[k] = 0; sv[k] = 0.0f; rv[k] = 0.0f;
        if (q < wl) {
            const size_t g = (size_t)q * np + row0 + r;
            idv[k] = ids[g]; sv[k] = Sv[g]; rv[k] = Rv[g];
        }
    }
    if constexpr (XT) {
#pragma unroll
        for (int k = 0; k < NX; ++k) {
            const int i = threadIdx.x + k * kEwBlock;
            if (i < kTileRows * LPR) xs[i] = xt_value<LG>(i, xv[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < NE; ++k) {
        const int e = l64 + 64 * k, q = e / kSpmvRowsPerWave, r = e % kSpmvRowsPerWave;
        if (q < wb) {
            int2 v = make_int2(0, 0);
            if constexpr (XT) v.x = xt_key<LG>(idv[k], tile0, q < wl);
            if (q < wl) {
                if constexpr (!XT) v.x = idv[k];
                v.y = __float_as_int(sv[k] + al * rv[k]);
            }
            stage[wave][q][r] = v;
        }
    }
    __syncthreads();
    ell_gather_bin<LG, WMAX, XT>(wb, np, row0, l64, X, stage[wave], xs, tile0, epi);
}

// ---- 128-bit streaming variants: a lane owns 4 consecutive RHS columns -------------
// Combine per-lane FP64 values of 4 columns across the row sub-groups of each
// wavefront (fixed xor tree) and across the NT / 64 wavefronts (fixed order), then
// write the tile partials of NV channels.
template <int LG, int NV, int NT = kEwBlock>
__device__ __forceinline__ void block_partials4(const JobDev& J, int s, int tile, const int* ch,
                                                double (&v)[NV][4]) {
    constexpr int LPR = (1 << LG) / 4;
    const int wave = threadIdx.x >> 6, l64 = threadIdx.x & 63;
    const int rsub = l64 / LPR, c4 = l64 % LPR;
#pragma unroll
    for (int off = LPR; off < 64; off <<= 1)
#pragma unroll
        for (int c = 0; c < NV; ++c)
#pragma unroll
            for (int j = 0; j < 4; ++j) v[c][j] += __shfl_xor(v[c][j], off, 64);
    __shared__ double red[NV][NT / 64][64];
    if (rsub == 0)
#pragma unroll
        for (int c = 0; c < NV; ++c)
#pragma unroll
            for (int j = 0; j < 4; ++j) red[c][wave][4 * c4 + j] = v[c][j];
    __syncthreads();
    if ((int)threadIdx.x < (1 << LG)) {
#pragma unroll
        for (int c = 0; c < NV; ++c) {
            double a = 0.0;
            for (int w = 0; w < NT / 64; ++w) a += red[c][w][threadIdx.x];
            part_ptr(J, ch[c], s, tile)[threadIdx.x] = a;
        }
    }
}

// block_partials4 with NF more channels carried in FP32 through the wavefront tree
// (short sums that their consumer reads in FP32) and in FP64 across the wavefronts; with
// fo set, channel c's sums are also written scaled by fs[c] in FP32 to fo[c * Rp + lane]
template <int LG, int NV, int NF, int NT>
__device__ __forceinline__ void block_partials4f(const JobDev& J, int s, int tile, const int* ch,
                                                 double (&v)[NV][4], const int* chf, float (&f)[NF][4],
                                                 MPF_GLOBAL float* fo = nullptr, const float* fs = nullptr) {
    constexpr int LPR = (1 << LG) / 4;
    const int wave = threadIdx.x >> 6, l64 = threadIdx.x & 63;
    const int rsub = l64 / LPR, c4 = l64 % LPR;
#pragma unroll
    for (int off = LPR; off < 64; off <<= 1) {
#pragma unroll
        for (int c = 0; c < NV; ++c)
#pragma unroll
            for (int j = 0; j < 4; ++j) v[c][j] += __shfl_xor(v[c][j], off, 64);
#pragma unroll
        for (int c = 0; c < NF; ++c)
#pragma unroll
            for (int j = 0; j < 4; ++j) f[c][j] += __shfl_xor(f[c][j], off, 64);
    }
    __shared__ double red[NV + NF][NT / 64][64];
    if (rsub == 0) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
#pragma unroll
            for (int c = 0; c < NV; ++c) red[c][wave][4 * c4 + j] = v[c][j];
#pragma unroll
            for (int c = 0; c < NF; ++c) red[NV + c][wave][4 * c4 + j] = (double)f[c][j];
        }
    }
    __syncthreads();
    if ((int)threadIdx.x < (1 << LG)) {
#pragma unroll
        for (int c = 0; c < NV + NF; ++c) {
            double a = 0.0;
            for (int w = 0; w < NT / 64; ++w) a += red[c][w][threadIdx.x];
            part_ptr(J, c < NV ? ch[c] : chf[c - NV], s, tile)[threadIdx.x] = a;
            if (c >= NV && fo) fo[(c - NV) * (1 << LG) + threadIdx.x] = (float)((double)fs[c - NV] * a);
        }
    }
}

// partial-sum chunks of a job: kChunkTiles tiles (kCoarseRows coarse rows) each, one
// partial per chunk in channels 5 and 6
__device__ __forceinline__ int coarse_chunks(const JobDev& J) {
    return (J.ntiles + kChunkTiles - 1) / kChunkTiles;
}

// Coarse aggregates: the rows of tile t with aggq = q form aggregate kAggPerTile t + q
// (the device numbering bisects every tile into kAggRows-node leaves; a job without the
// table takes local rows [kAggRows q, kAggRows (q + 1))).  The FP64 tile partials of
// W^T r for quarter q sit in channel agg_channel(q).
__device__ __forceinline__ constexpr int agg_channel(int q) { return q == 0 ? 4 : 6 + q; }
__device__ __forceinline__ int row_agg(const JobDev& J, int row) {
    return J.aggq ? (int)J.aggq[row] : (row % kTileRows) / kAggRows;
}
static_assert(kAggPerTile == 4 && kAggPerTile * kAggRows == kTileRows,
              "aggregate channels 4, 7, 8, 9: four quarters per tile");

// block_partials4 with the reduction buffer supplied by the caller (it may share LDS with
// data the wavefronts used before: the block synchronises before writing it)
template <int LG, int NV>
__device__ __forceinline__ void block_partials4_buf(const JobDev& J, int s, int tile, const int* ch,
                                                    double (&v)[NV][4], double (*red)[kEwBlock / 64][64]) {
    constexpr int LPR = (1 << LG) / 4;
    const int wave = threadIdx.x >> 6, l64 = threadIdx.x & 63;
    const int rsub = l64 / LPR, c4 = l64 % LPR;
#pragma unroll
    for (int off = LPR; off < 64; off <<= 1)
#pragma unroll
        for (int c = 0; c < NV; ++c)
#pragma unroll
            for (int j = 0; j < 4; ++j) v[c][j] += __shfl_xor(v[c][j], off, 64);
    __syncthreads();
    if (rsub == 0)
#pragma unroll
        for (int c = 0; c < NV; ++c)
#pragma unroll
            for (int j = 0; j < 4; ++j) red[c][wave][4 * c4 + j] = v[c][j];
    __syncthreads();
    if ((int)threadIdx.x < (1 << LG)) {
#pragma unroll
        for (int c = 0; c < NV; ++c) {
            double a = 0.0;
            for (int w = 0; w < kEwBlock / 64; ++w) a += red[c][w][threadIdx.x];
            part_ptr(J, ch[c], s, tile)[threadIdx.x] = a;
        }
    }
}

// rho of iteration `par` parity: r.z of the current residual lives in one of two slots so
// that the kernel computing the next rho never overwrites the one its other blocks read
__device__ __forceinline__ int rho_slot(int par) { return par ? SC_RHO1 : SC_RHO; }

}  // namespace mpf
