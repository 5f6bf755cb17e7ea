// Standalone Krylov solves on a caller's sparse matrix: ConjugateGradient::solve and
// BiCGSTAB::solve with a SparseMatrix operator and the diagonal preconditioner
// (src/numeric/conjugate_gradient.cpp:61-161, src/numeric/bicgstab.cpp:83-193,
// src/numeric/preconditioner.cpp:59-76), all right-hand sides at once.
//
// The matrix arrives in the reference's SparseMatrix layout (sparse_matrix.cpp:48-114,
// 396-423): 32 slots per row, entry (row, slot) at row*32 + slot, empty slots with column
// id 0xFFFFFFFF.  It is imported into the product's column-major ELL (one harmonic, S =
// the values, R = 0) and solved by the same Krylov driver as the EIT models.
#include <algorithm>
#include <cstring>
#include <memory>

#include "blockcg.h"
#include "krylov_dev.h"
#include "model.h"
#include "numeric.h"

namespace mpf {

constexpr uint32_t kInvalidIndex = 0xFFFFFFFFu;
constexpr int kRefSlots = 32;

// one thread per (padded) row: slot-major ELL, diagonal slot (-1: the row has no
// diagonal entry, its Jacobi inverse is 0 as in diagonalSparseKernel,
// preconditioner_kernel.cu:63-83), last used slot + 1.  A column id that is neither
// 0xFFFFFFFF nor below n sets *bad (the solve is refused before any kernel reads it).
__global__ void k_ell_import(int n, int np, int density, const uint32_t* __restrict__ cid,
                             const float* __restrict__ val, int wpad, int32_t* __restrict__ ids,
                             float* __restrict__ s, int32_t* __restrict__ diag, int* __restrict__ used,
                             int* __restrict__ bad) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= np) return;
    int last = 0, dg = -1;
    for (int q = 0; q < wpad; ++q) {
        int id = row;
        float v = 0.0f;
        if (row < n && q < density) {
            const uint32_t c = cid[(size_t)row * kRefSlots + q];
            if (c != kInvalidIndex) {
                if (c >= (uint32_t)n) {
                    atomicOr(bad, 1);
                } else {
                    id = (int)c;
                    v = val[(size_t)row * kRefSlots + q];
                    last = q + 1;
                    if (id == row && dg < 0) dg = q;
                }
            }
        }
        ids[(size_t)q * np + row] = id;
        s[(size_t)q * np + row] = v;
    }
    diag[row] = dg;
    used[row] = last;
}

__global__ void k_group_width(int ngroups, const int* __restrict__ used, uint8_t* __restrict__ gw) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroups) return;
    int w = 0;
    for (int r = 0; r < 16; ++r) w = max(w, used[g * 16 + r]);
    gw[g] = (uint8_t)w;
}

// column sums of X0 (column-major, ld) in FP64, fixed order: cs[lane]; with `sq` the
// sums of squares instead (b.b of the stop rule).  One 256-thread block per column.
__global__ void k_col_sums(int n, const float* __restrict__ X, int ldx, int sq, double* __restrict__ cs) {
    const int lane = blockIdx.x;
    __shared__ double red[256];
    double a = 0.0;
    for (int row = threadIdx.x; row < n; row += 256) {
        const double v = (double)X[row + (size_t)lane * ldx];
        a += sq ? v * v : v;
    }
    red[threadIdx.x] = a;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) cs[lane] = red[0];
}

// Initial residual of the caller's start X0 (conjugate_gradient.cpp:80-88):
// r = B - A X0, and with dcFree r += 1 (1^T X0) as the reference writes it (addScalar
// adds the column sums of X0, so with X0 != 0 the start residual is that of the operator
// A - 1 1^T; X0 = 0, EIT::solve's start, is unaffected).  FP64 per entry; written as
// rhs [row][lane] (FP32) and F [row][lane] (FP64), the system of the correction solved
// from 0.
__global__ void k_rhs_import(int n, int np, int R, int Rp, const float* __restrict__ B, int ldb,
                             const uint32_t* __restrict__ cid, const float* __restrict__ val, int density,
                             const float* __restrict__ X0, int ldx, const double* __restrict__ cs,
                             float* __restrict__ rhs, double* __restrict__ F) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)np * Rp) return;
    const int row = (int)(i / Rp), lane = (int)(i % Rp);
    double r = 0.0;
    if (row < n && lane < R) {
        r = (double)B[row + (size_t)lane * ldb];
        const float* x = X0 + (size_t)lane * ldx;
        for (int q = 0; q < density; ++q) {
            const uint32_t c = cid[(size_t)row * kRefSlots + q];
            if (c < (uint32_t)n) r -= (double)val[(size_t)row * kRefSlots + q] * (double)x[c];
        }
        if (cs) r += cs[lane];
    }
    rhs[i] = (float)r;
    F[i] = r;
}

// X = X0 + the correction (FP32 Krylov solution, or its FP64 refinement)
__global__ void k_x_export(int n, int R, int Rp, const float* __restrict__ x, const double* __restrict__ x64,
                           int use64, float* __restrict__ X, int ldx) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * R) return;
    const int lane = (int)(i / n), row = (int)(i % n);
    const size_t k = (size_t)row * Rp + lane;
    float* o = X + row + (size_t)lane * ldx;
    *o = use64 ? (float)((double)*o + x64[k]) : *o + x[k];
}

int krylov_solve(mpf_ctx* ctx, int n, int density, const uint32_t* d_cid, const float* d_val, int R,
                 const float* d_B, int ldb, float* d_X, int ldx, int dc_free, const mpf_solve_opts& o,
                 mpf_solve_stats* stats) {
    MPF_REQUIRE(n >= 1 && density >= 1 && density <= kRefSlots, "krylov solve: bad rows or density");
    MPF_REQUIRE(R >= 1 && R <= 64, "krylov solve: 1 <= R <= 64 right-hand sides");
    MPF_REQUIRE(ldb >= n && ldx >= n, "krylov solve: leading dimension < rows");
    const bool bicg = o.krylov == MPF_KRYLOV_BICGSTAB;
    const bool block = o.krylov == MPF_KRYLOV_BLOCK_CG;
    MPF_REQUIRE(o.krylov == MPF_KRYLOV_CG || bicg || block, "krylov solve: unknown Krylov method");
    hipStream_t st = ctx->stream;
    const int np = (n + kTileRows - 1) / kTileRows * kTileRows;
    const int wpad = (density + kSlotPad - 1) / kSlotPad * kSlotPad;
    MPF_REQUIRE(!bicg || wpad <= 24, "BiCGSTAB: padded ELL width above 24 is not supported");
    const int Rp = R <= 16 ? 16 : (R <= 32 ? 32 : 64);
    const int ntiles = np / kTileRows;
    const size_t ell = (size_t)wpad * np, V = (size_t)np * Rp;
    DBuf<int32_t> ids(ell), diag(np), used(np), bad(1);
    DBuf<uint8_t> gw(np / 16);
    DBuf<float> s(ell), zero(ell), dinv(np), x(V), rv(V), p(V), q(V), pring(V * (kRing - 1)), rhs(V);
    DBuf<double> F(V), x64(V), part((size_t)kParts * ntiles * Rp), sc((size_t)SC_NFIELDS * Rp), cs(R);
    DBuf<float> rh, yv, sv, zv, tv;
    if (bicg) { rh.alloc(V); yv.alloc(V); sv.alloc(V); zv.alloc(V); tv.alloc(V); }
    MPF_HIP(hipMemsetAsync(zero.p, 0, zero.bytes(), st));
    MPF_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), st));
    MPF_HIP(hipMemsetAsync(sc.p, 0, sc.bytes(), st));
    k_ell_import<<<(np + 255) / 256, 256, 0, st>>>(n, np, density, d_cid, d_val, wpad, ids.p, s.p, diag.p, used.p,
                                                  bad.p);
    int hbad = 0;
    MPF_HIP(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
    MPF_HIP(hipStreamSynchronize(st));
    MPF_REQUIRE(hbad == 0, "krylov solve: a column id is neither 0xFFFFFFFF nor below n");
    k_group_width<<<(np / 16 + 255) / 256, 256, 0, st>>>(np / 16, used.p, gw.p);
    // the stop rules compare with b.b of the caller's B (conjugate_gradient.cpp:104-105)
    k_col_sums<<<R, 256, 0, st>>>(n, d_B, ldb, 1, sc.p + (size_t)SC_X2 * Rp);
    // dcFree (CG): the reference adds the column sums of X0 to the start residual
    const bool dc_sum = dc_free && !bicg;
    if (dc_sum) k_col_sums<<<R, 256, 0, st>>>(n, d_X, ldx, 0, cs.p);
    k_rhs_import<<<(unsigned)((V + 255) / 256), 256, 0, st>>>(n, np, R, Rp, d_B, ldb, d_cid, d_val, density, d_X,
                                                              ldx, dc_sum ? cs.p : nullptr, rhs.p, F.p);
    MPF_HIP(hipGetLastError());
    std::vector<JobDev> hj(1);
    JobDev& J = hj[0];
    std::memset(&J, 0, sizeof(J));
    J.n = n; J.np = np; J.width = wpad; J.C = 1; J.R = R; J.Rp = Rp; J.lg = Rp == 16 ? 4 : (Rp == 32 ? 5 : 6);
    J.ntiles = ntiles; J.col0 = 0; J.nnz = 0;
    // ConjugateGradient applies dcFree as A + 1 1^T; BiCGSTAB ignores it (bicgstab.cpp:87)
    // unless the product's constant projection is asked for (bicgstab_ignore_dc == 0)
    J.dc = dc_free ? ((bicg && o.bicgstab_ignore_dc) ? -1 : 0) : -1;
    J.dc_rank1 = bicg ? 0 : 1;
    J.ids = ids.p; J.s = s.p; J.r = zero.p; J.s_lo = zero.p; J.r_lo = zero.p; J.dinv = dinv.p;
    J.alpha[0] = 0.0f; J.xscale[0] = 1.0f; J.alpha64[0] = 0.0; J.xscale64[0] = 1.0;
    J.F = F.p; J.rhs = rhs.p; J.use_rhs = 1;
    J.x = x.p; J.rv = rv.p; J.p = p.p; J.q = q.p;
    for (int b = 0; b < kRing - 1; ++b) J.pr[b] = pring.p + V * b;
    J.rh = rh.p; J.yv = yv.p; J.sv = sv.p; J.zv = zv.p; J.tv = tv.p;
    J.x64 = x64.p; J.part = part.p; J.sc = sc.p;
    J.diag_slot = diag.p; J.gw = gw.p;
    DBuf<JobDev> jobs(1);
    jobs.upload(hj.data(), 1, st);
    std::vector<int4> ht;
    for (int t = 0; t < ntiles; ++t) ht.push_back(make_int4(0, 0, t, 0));
    DBuf<int4> tiles(ntiles);
    tiles.upload(ht.data(), ht.size(), st);
    const int2 hjs = make_int2(0, 0);
    DBuf<int2> js(1);
    js.upload(&hjs, 1, st);
    launch_dinv(jobs.p, tiles.p, ntiles, st);
    bool capped = false;
    const bool mixed = o.precision == MPF_PREC_MIXED;
    // the correction from the caller's X0 starts at 0 (EIT::solve's fields start at 0,
    // eit.cpp:165; then the correction is the solution itself) and stops relative to b.b
    std::unique_ptr<KrylovDriver> drv;
    BcgArgs ba;
    if (block) {
        ba.n = n; ba.np = np; ba.R = R; ba.Rp = Rp; ba.wpad = wpad; ba.dc = dc_free ? 1 : 0;
        ba.ids = ids.p; ba.s = s.p; ba.zero = zero.p; ba.dinv = dinv.p; ba.rhs = rhs.p;
        ba.x = x.p; ba.p = p.p; ba.q = q.p; ba.bad = bad.p;
    } else {
        drv.reset(new KrylovDriver(ctx, {&hj}));
    }
    const int max_iter = o.max_iter > 0 ? o.max_iter : 10 * n;
    // one solve of the current rhs: from the caller's b.b (ref_b) or relative to the rhs
    auto solve = [&](int rule, double tol, bool ref_b) {
        if (!block) return drv->run(o, rule, tol, true, ref_b, stats);
        ba.bb_ref = ref_b ? sc.p + (size_t)SC_X2 * Rp : nullptr;
        const BcgResult res = block_pcg(ba, rule, tol, max_iter, o.check_every, st);
        if (stats) {
            stats->steps += res.steps;
            stats->max_steps = std::max<int64_t>(stats->max_steps, res.steps);
            stats->not_converged += (res.capped || res.broke) ? 1 : 0;
        }
        return res.capped || res.broke;
    };
    capped |= solve(o.stop_rule, o.tol, true);
    if (mixed) {
        launch_x64_update(jobs.p, tiles.p, ntiles, 1, st);
        for (int k = 0; k < o.refine_steps; ++k) {
            launch_residual64(jobs.p, tiles.p, ntiles, js.p, 1, hj[0].lg, (size_t)hj[0].np * hj[0].Rp * sizeof(double), st);
            capped |= solve(MPF_STOP_TOLERANCE, o.refine_tol, false);
            launch_x64_update(jobs.p, tiles.p, ntiles, 0, st);
        }
    }
    k_x_export<<<(unsigned)(((size_t)n * R + 255) / 256), 256, 0, st>>>(n, R, Rp, x.p, x64.p, mixed ? 1 : 0, d_X, ldx);
    MPF_HIP(hipGetLastError());
    MPF_HIP(hipStreamSynchronize(st));
    return capped ? MPF_ENOCONV : MPF_OK;
}

// ---------------------------------------------------------------------------
// SparseMatrix::multiply (sparse_matrix.cpp:262-288; multiplyKernel,
// sparse_matrix_kernel.cu:157-196) and the product-layout multi-harmonic SpMV, both on
// the 128-bit gather core of the Krylov SpMV (ell_gather_rows): a 1024-thread block is
// one 256-row tile, each wavefront stages its 16 rows' (column id, value) pairs in LDS
// and gathers whole neighbour rows of a row-interleaved copy of X.
// ---------------------------------------------------------------------------

// X (column-major, leading dimension ldx) -> Xt [chunk][np][RP]: chunk c holds columns
// c*RP .. c*RP + RP - 1; rows >= rows and columns >= cols are zero.  64-row tiles are
// transposed through LDS so that both sides are coalesced.
template <int RP>
__global__ void __launch_bounds__(256) k_cols_to_rows(int rows, int np, int cols, const float* __restrict__ X,
                                                      int ldx, float* __restrict__ Xt) {
    __shared__ float tile[RP][65];
    const int row0 = blockIdx.x * 64, chunk = blockIdx.y;
    for (int i = threadIdx.x; i < RP * 64; i += 256) {
        const int j = i / 64, r = i % 64;
        const int row = row0 + r, col = chunk * RP + j;
        tile[j][r] = (row < rows && col < cols) ? X[row + (size_t)col * ldx] : 0.0f;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < RP * 64; i += 256) {
        const int r = i / RP, j = i % RP;
        Xt[((size_t)chunk * np + row0 + r) * RP + j] = tile[j][r];
    }
}

struct SeamAlpha {
    float a[kMaxSlices];
};

// REF: the reference layout -- entry (row, slot) at row*32 + slot, `density` slots,
// 0xFFFFFFFF empty -- with Y column-major nrows x ncols (leading dimension ldy); a
// column id that is neither 0xFFFFFFFF nor below xrows sets *bad and contributes 0.
// Otherwise the product layout: slot-major [W][np] ids / S / R (padding slots id = row,
// value 0), value S + alpha_c R for harmonic c = blockIdx.y, Y = [c][np][RP]; an id
// outside [0, np) sets *bad and contributes 0.
template <int LG, int W, bool REF>
__global__ void __launch_bounds__(kEwBlock) k_seam_spmv(int nrows, int xrows, int np, int ncols, int density,
                                                        const uint32_t* __restrict__ cid,
                                                        const float* __restrict__ val,
                                                        const int32_t* __restrict__ ids, const float* __restrict__ S,
                                                        const float* __restrict__ Rv, SeamAlpha al,
                                                        const float* __restrict__ Xt, float* __restrict__ Y, int ldy,
                                                        int* __restrict__ bad) {
    constexpr int Rp = 1 << LG, LPR = Rp / 4;
    __shared__ int2 stage[kEwBlock / 64][W][kSpmvRowsPerWave];
    __shared__ float yt[REF ? Rp : 1][REF ? kTileRows + 1 : 1];
    const int wave = threadIdx.x >> 6, l64 = threadIdx.x & 63;
    const int tile0 = blockIdx.x * kTileRows;
    const int row0 = tile0 + wave * kSpmvRowsPerWave;
    const int chunk = blockIdx.y;
    for (int e = l64; e < W * kSpmvRowsPerWave; e += 64) {
        int2 v = make_int2(0, 0);
        int q, r;
        if constexpr (REF) {
            // 16 rows x 32 slots are contiguous in the reference layout
            r = e / W;
            q = e % W;
            const int row = row0 + r;
            if (row < nrows && q < density) {
                const uint32_t c = cid[(size_t)row * kRefSlots + q];
                if (c != kInvalidIndex) {
                    if (c < (uint32_t)xrows) v = make_int2((int)c, __float_as_int(val[(size_t)row * kRefSlots + q]));
                    else atomicOr(bad, 1);
                }
            }
        } else {
            q = e / kSpmvRowsPerWave;
            r = e % kSpmvRowsPerWave;
            const size_t g = (size_t)q * np + row0 + r;
            const int id = ids[g];
            const float a = S[g] + al.a[chunk] * Rv[g];
            if (id >= 0 && id < np) v = make_int2(id, __float_as_int(a));
            else atomicOr(bad, 1);
        }
        stage[wave][q][r] = v;
    }
    __syncthreads();
    const float* X = Xt + (size_t)chunk * np * Rp;
    ell_gather_rows<LG, W, false>(np, row0, l64, X, stage[wave], nullptr, tile0, [&](int row, int c4, float4 acc) {
        if constexpr (REF) {
            yt[4 * c4 + 0][row - tile0] = acc.x;
            yt[4 * c4 + 1][row - tile0] = acc.y;
            yt[4 * c4 + 2][row - tile0] = acc.z;
            yt[4 * c4 + 3][row - tile0] = acc.w;
        } else {
            ((float4*)(Y + ((size_t)chunk * np + row) * Rp))[c4] = acc;
        }
    });
    if constexpr (REF) {
        __syncthreads();
        for (int i = threadIdx.x; i < Rp * kTileRows; i += kEwBlock) {
            const int j = i / kTileRows, r = i % kTileRows;
            const int row = tile0 + r, col = chunk * Rp + j;
            if (row < nrows && col < ncols) Y[row + (size_t)col * ldy] = yt[j][r];
        }
    }
}

template <int LG, bool REF>
static void seam_launch(int W, dim3 grid, int nrows, int xrows, int np, int ncols, int density, const uint32_t* cid,
                        const float* val, const int32_t* ids, const float* S, const float* R, const SeamAlpha& al,
                        const float* Xt, float* Y, int ldy, int* bad, hipStream_t st) {
#define MPF_SEAM(w) k_seam_spmv<LG, w, REF><<<grid, kEwBlock, 0, st>>>(nrows, xrows, np, ncols, density, cid, val, \
                                                                     ids, S, R, al, Xt, Y, ldy, bad)
    switch (W) {
        case 8: MPF_SEAM(8); break;
        case 16: MPF_SEAM(16); break;
        case 24: MPF_SEAM(24); break;
        case 32: MPF_SEAM(32); break;
        case 40: if constexpr (!REF) { MPF_SEAM(40); break; } [[fallthrough]];
        case 48: if constexpr (!REF) { MPF_SEAM(48); break; } [[fallthrough]];
        case 56: if constexpr (!REF) { MPF_SEAM(56); break; } [[fallthrough]];
        case 64: if constexpr (!REF) { MPF_SEAM(64); break; } [[fallthrough]];
        default: throw Error(MPF_EINVAL, "spmv: unsupported ELL width");
    }
#undef MPF_SEAM
    MPF_HIP(hipGetLastError());
}

void sparse_multiply(mpf_ctx* ctx, const float* d_values, const uint32_t* d_cid, const float* d_X,
                     int result_rows, int matrix_rows, int columns, int density, float* d_Y) {
    MPF_REQUIRE(result_rows >= 1 && matrix_rows >= 1 && columns >= 1, "sparse multiply: empty shape");
    MPF_REQUIRE(density >= 0 && density <= kRefSlots, "sparse multiply: density above 32 slots");
    hipStream_t st = ctx->stream;
    const int np = round_up(std::max(result_rows, matrix_rows), kTileRows);
    const int Rp = columns <= 16 ? 16 : (columns <= 32 ? 32 : 64);
    const int nchunk = (columns + Rp - 1) / Rp;
    const int W = std::max(8, round_up(density, 8));
    DBuf<float> Xt((size_t)nchunk * np * Rp);
    DBuf<int> bad(1);
    MPF_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), st));
    const dim3 tgrid(np / 64, nchunk), grid(np / kTileRows, nchunk);
    const SeamAlpha al{};
    if (Rp == 16) {
        k_cols_to_rows<16><<<tgrid, 256, 0, st>>>(matrix_rows, np, columns, d_X, matrix_rows, Xt.p);
        seam_launch<4, true>(W, grid, result_rows, matrix_rows, np, columns, density, d_cid, d_values, nullptr,
                             nullptr, nullptr, al, Xt.p, d_Y, result_rows, bad.p, st);
    } else if (Rp == 32) {
        k_cols_to_rows<32><<<tgrid, 256, 0, st>>>(matrix_rows, np, columns, d_X, matrix_rows, Xt.p);
        seam_launch<5, true>(W, grid, result_rows, matrix_rows, np, columns, density, d_cid, d_values, nullptr,
                             nullptr, nullptr, al, Xt.p, d_Y, result_rows, bad.p, st);
    } else {
        k_cols_to_rows<64><<<tgrid, 256, 0, st>>>(matrix_rows, np, columns, d_X, matrix_rows, Xt.p);
        seam_launch<6, true>(W, grid, result_rows, matrix_rows, np, columns, density, d_cid, d_values, nullptr,
                             nullptr, nullptr, al, Xt.p, d_Y, result_rows, bad.p, st);
    }
    int h = 0;
    MPF_HIP(hipMemcpyAsync(&h, bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
    MPF_HIP(hipStreamSynchronize(st));
    MPF_REQUIRE(h == 0, "sparse multiply: a column id is neither 0xFFFFFFFF nor below matrix_rows");
}

void ell_spmv_product(const int32_t* ids, const float* S, const float* zero, int np, int width, int Rp,
                      const float* X, float* Y, int* bad, hipStream_t st) {
    const SeamAlpha al{};
    const dim3 grid(np / kTileRows, 1);
    if (Rp == 16)
        seam_launch<4, false>(width, grid, np, np, np, Rp, width, nullptr, nullptr, ids, S, zero, al, X, Y, Rp, bad, st);
    else if (Rp == 32)
        seam_launch<5, false>(width, grid, np, np, np, Rp, width, nullptr, nullptr, ids, S, zero, al, X, Y, Rp, bad, st);
    else
        seam_launch<6, false>(width, grid, np, np, np, Rp, width, nullptr, nullptr, ids, S, zero, al, X, Y, Rp, bad, st);
}

void ell_spmv(mpf_ctx* ctx, int np, int width, const int32_t* d_ids, const float* d_S, const float* d_R, int C,
              const float* alpha, const float* d_X, float* d_Y, int ld) {
    hipStream_t st = ctx->stream;
    SeamAlpha al{};
    for (int c = 0; c < C; ++c) al.a[c] = alpha[c];
    DBuf<int> bad(1);
    MPF_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), st));
    const dim3 grid(np / kTileRows, C);
    if (ld == 16)
        seam_launch<4, false>(width, grid, np, np, np, ld, width, nullptr, nullptr, d_ids, d_S, d_R, al, d_X, d_Y,
                              ld, bad.p, st);
    else if (ld == 32)
        seam_launch<5, false>(width, grid, np, np, np, ld, width, nullptr, nullptr, d_ids, d_S, d_R, al, d_X, d_Y,
                              ld, bad.p, st);
    else
        seam_launch<6, false>(width, grid, np, np, np, ld, width, nullptr, nullptr, d_ids, d_S, d_R, al, d_X, d_Y,
                              ld, bad.p, st);
    int h = 0;
    MPF_HIP(hipMemcpyAsync(&h, bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
    MPF_HIP(hipStreamSynchronize(st));
    MPF_REQUIRE(h == 0, "mpf_ell_spmv: a column id outside [0, np)");
}

}  // namespace mpf
