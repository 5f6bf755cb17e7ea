// Standalone Krylov solves on a caller's sparse matrix: ConjugateGradient::solve and
// BiCGSTAB::solve with a SparseMatrix operator and the diagonal preconditioner
// (src/numeric/conjugate_gradient.cpp:61-161, src/numeric/bicgstab.cpp:83-193,
// src/numeric/preconditioner.cpp:59-76), all right-hand sides at once.
//
// The matrix arrives in the reference's SparseMatrix layout (sparse_matrix.cpp:48-114,
// 396-423): 32 slots per row, entry (row, slot) at row*32 + slot, empty slots with column
// id 0xFFFFFFFF.  It is imported into the product's column-major ELL (one harmonic, S =
// the values, R = 0) and solved by the same Krylov driver as the EIT models.
#include <cstring>

#include "model.h"
#include "numeric.h"

namespace mpf {

constexpr uint32_t kInvalidIndex = 0xFFFFFFFFu;
constexpr int kRefSlots = 32;

// one thread per (padded) row: slot-major ELL, diagonal slot, last used slot + 1
__global__ void k_ell_import(int n, int np, int density, const uint32_t* __restrict__ cid,
                             const float* __restrict__ val, int wpad, int32_t* __restrict__ ids,
                             float* __restrict__ s, int32_t* __restrict__ diag, int* __restrict__ used) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= np) return;
    int last = 0, dg = 0;
    for (int q = 0; q < wpad; ++q) {
        int id = row;
        float v = 0.0f;
        if (row < n && q < density) {
            const uint32_t c = cid[(size_t)row * kRefSlots + q];
            if (c != kInvalidIndex) {
                id = (int)c;
                v = val[(size_t)row * kRefSlots + q];
                last = q + 1;
                if (id == row) dg = q;
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

// B (column-major, ld) -> rhs [row][lane] (FP32) and F [row][lane] (FP64)
__global__ void k_rhs_import(int n, int np, int R, int Rp, const float* __restrict__ B, int ldb,
                             float* __restrict__ rhs, double* __restrict__ F) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)np * Rp) return;
    const int row = (int)(i / Rp), lane = (int)(i % Rp);
    const float b = (row < n && lane < R) ? B[row + (size_t)lane * ldb] : 0.0f;
    rhs[i] = b;
    F[i] = (double)b;
}

__global__ void k_x_export(int n, int R, int Rp, const float* __restrict__ x, const double* __restrict__ x64,
                           int use64, float* __restrict__ X, int ldx) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * R) return;
    const int lane = (int)(i / n), row = (int)(i % n);
    const size_t k = (size_t)row * Rp + lane;
    X[row + (size_t)lane * ldx] = use64 ? (float)x64[k] : x[k];
}

int krylov_solve(mpf_ctx* ctx, int n, int density, const uint32_t* d_cid, const float* d_val, int R,
                 const float* d_B, int ldb, float* d_X, int ldx, int dc_free, const mpf_solve_opts& o,
                 mpf_solve_stats* stats) {
    MPF_REQUIRE(n >= 1 && density >= 1 && density <= kRefSlots, "krylov solve: bad rows or density");
    MPF_REQUIRE(R >= 1 && R <= 64, "krylov solve: 1 <= R <= 64 right-hand sides");
    MPF_REQUIRE(ldb >= n && ldx >= n, "krylov solve: leading dimension < rows");
    const bool bicg = o.krylov == MPF_KRYLOV_BICGSTAB;
    MPF_REQUIRE(o.krylov == MPF_KRYLOV_CG || bicg, "krylov solve: unknown Krylov method");
    hipStream_t st = ctx->stream;
    const int np = (n + kTileRows - 1) / kTileRows * kTileRows;
    const int wpad = (density + kSlotPad - 1) / kSlotPad * kSlotPad;
    MPF_REQUIRE(!bicg || wpad <= 24, "BiCGSTAB: padded ELL width above 24 is not supported");
    const int Rp = R <= 16 ? 16 : (R <= 32 ? 32 : 64);
    const int ntiles = np / kTileRows;
    const size_t ell = (size_t)wpad * np, V = (size_t)np * Rp;
    DBuf<int32_t> ids(ell), diag(np), used(np);
    DBuf<uint8_t> gw(np / 16);
    DBuf<float> s(ell), zero(ell), dinv(np), x(V), rv(V), p(V), q(V), pring(V * (kRing - 1)), rhs(V);
    DBuf<double> F(V), x64(V), part((size_t)kParts * ntiles * Rp), sc((size_t)SC_NFIELDS * Rp);
    DBuf<float> rh, yv, sv, zv, tv;
    if (bicg) { rh.alloc(V); yv.alloc(V); sv.alloc(V); zv.alloc(V); tv.alloc(V); }
    MPF_HIP(hipMemsetAsync(zero.p, 0, zero.bytes(), st));
    k_ell_import<<<(np + 255) / 256, 256, 0, st>>>(n, np, density, d_cid, d_val, wpad, ids.p, s.p, diag.p, used.p);
    k_group_width<<<(np / 16 + 255) / 256, 256, 0, st>>>(np / 16, used.p, gw.p);
    k_rhs_import<<<(unsigned)((V + 255) / 256), 256, 0, st>>>(n, np, R, Rp, d_B, ldb, rhs.p, F.p);
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
    KrylovDriver drv(ctx, {&hj});
    bool capped = false;
    const bool mixed = o.precision == MPF_PREC_MIXED;
    // the first solve starts from x = 0 (EIT::solve fills the fields with 0, eit.cpp:165)
    capped |= drv.run(o, o.stop_rule, o.tol, true, false, stats);
    if (mixed) {
        launch_x64_update(jobs.p, tiles.p, ntiles, 1, st);
        for (int k = 0; k < o.refine_steps; ++k) {
            launch_residual64(jobs.p, tiles.p, ntiles, js.p, 1, st);
            capped |= drv.run(o, MPF_STOP_TOLERANCE, o.refine_tol, true, false, stats);
            launch_x64_update(jobs.p, tiles.p, ntiles, 0, st);
        }
    }
    k_x_export<<<(unsigned)(((size_t)n * R + 255) / 256), 256, 0, st>>>(n, R, Rp, x.p, x64.p, mixed ? 1 : 0, d_X, ldx);
    MPF_HIP(hipGetLastError());
    MPF_HIP(hipStreamSynchronize(st));
    return capped ? MPF_ENOCONV : MPF_OK;
}

}  // namespace mpf
