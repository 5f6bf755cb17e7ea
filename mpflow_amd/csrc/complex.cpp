// Host model of the complex-conductivity EIT (complex.h).
//
// models::EIT over Equation<thrust::complex<float>, ...> (src/models/eit.cpp:139-245,
// 289-304): the set-up (mesh order, ELL pattern, gather plan, excitation and measurement
// maps) is the real model's; S and R are assembled from complex sigma, every (frame,
// harmonic) pair runs complex CG (conjugate orthogonal by default, the reference's
// conjugated dot products on request) in the same launches, and the voltages and
// Jacobian are complex.  Solver::solveAbsolute (src/solver/solver.cpp:195-225) with
// Inverse's plain transposes (inverse.cpp:148-152, 196-204) becomes the data-space step
// (J J^T + lam^2 I) y = r, delta = J^T y, whose matrix is complex symmetric: FP64 sums
// of the FP32 Jacobian and an FP64 complex L L^T.
#include "complex.h"

#include <algorithm>
#include <cstring>

#include "krylov.h"
#include "model.h"

namespace mpf {

struct ComplexModel::Impl {
    mpf_ctx* ctx = nullptr;
    MeshSetup S;
    int frames = 1, quirks = 1, exact = 0;
    int Rq = 2, ntile_job = 0, ntiles = 0, nasm = 0, njs = 0;
    DBuf<int32_t> ids, diag_slot, ctb_off, ctb_elem, elements, w_off, w_node, colmap;
    DBuf<double> ctb_k, ctb_m, F, w_val, colscale;
    DBuf<float> ek, em;
    DBuf<float2> sv, rv, dinv, x, r, p, q, gamma;
    DBuf<zd> part, sc;
    DBuf<int> st, live;
    std::vector<CJob> hjobs;
    DBuf<CJob> jobs;
    DBuf<int4> tiles, asm_items;
    DBuf<int2> js;
    int* h_live = nullptr;
    int* h_st = nullptr;
    // Gauss-Newton
    DBuf<float2> Jm;
    DBuf<zd> U64, G, y, rgn, vE;
    DBuf<int2> pairs;
    DBuf<int> info;

    ~Impl() {
        if (h_live) (void)hipHostFree(h_live);
        if (h_st) (void)hipHostFree(h_st);
    }

    void ensure_gn() {
        const int m = S.mrows;
        if (Jm.p) return;
        Jm.alloc((size_t)frames * S.E * m);
        U64.alloc((size_t)frames * m);
        G.alloc((size_t)frames * m * m);
        y.alloc((size_t)frames * m);
        rgn.alloc((size_t)frames * m);
        vE.alloc((size_t)frames * S.E);
        pairs.alloc(cholesky_pairs(m));
        info.alloc(1);
    }
};

ComplexModel::ComplexModel(mpf_ctx* c, const mpf_model_desc& d) : impl(new Impl) {
    Impl& M = *impl;
    MPF_REQUIRE(d.frames >= 1, "model: frames < 1");
    MPF_REQUIRE(d.contact_impedance <= 0.0 && d.source_type == MPF_SOURCE_CURRENT,
                "complex conductivity: current sources and the gap model only");
    M.ctx = c;
    M.frames = d.frames;
    M.quirks = d.quirks;
    M.exact = d.jacobian_exact;
    build_setup(d, M.S);
    const MeshSetup& S = M.S;
    MPF_REQUIRE(S.Ru <= 64, "complex conductivity: at most 64 unique pattern columns");
    hipStream_t st0 = c->stream;
    const int np = S.np, W = S.wpad;
    M.ids.alloc((size_t)W * np); M.ids.upload(S.ids.data(), S.ids.size(), st0);
    M.diag_slot.alloc(np); M.diag_slot.upload(S.diag_slot.data(), np, st0);
    M.ctb_off.alloc(S.ctb_off.size()); M.ctb_off.upload(S.ctb_off.data(), S.ctb_off.size(), st0);
    M.ctb_elem.alloc(S.ctb_elem.size()); M.ctb_elem.upload(S.ctb_elem.data(), S.ctb_elem.size(), st0);
    M.ctb_k.alloc(S.ctb_k.size()); M.ctb_k.upload(S.ctb_k.data(), S.ctb_k.size(), st0);
    M.ctb_m.alloc(S.ctb_m.size()); M.ctb_m.upload(S.ctb_m.data(), S.ctb_m.size(), st0);
    M.elements.alloc(S.elements.size()); M.elements.upload(S.elements.data(), S.elements.size(), st0);
    M.ek.alloc(S.elem_k.size()); M.ek.upload(S.elem_k.data(), S.elem_k.size(), st0);
    M.em.alloc(S.elem_m.size()); M.em.upload(S.elem_m.data(), S.elem_m.size(), st0);
    M.F.alloc(S.F.size()); M.F.upload(S.F.data(), S.F.size(), st0);
    M.w_off.alloc(S.w_off.size()); M.w_off.upload(S.w_off.data(), S.w_off.size(), st0);
    M.w_node.alloc(std::max<size_t>(1, S.w_node.size()));
    if (!S.w_node.empty()) M.w_node.upload(S.w_node.data(), S.w_node.size(), st0);
    M.w_val.alloc(std::max<size_t>(1, S.w_val.size()));
    if (!S.w_val.empty()) M.w_val.upload(S.w_val.data(), S.w_val.size(), st0);
    M.colmap.alloc(S.R); M.colmap.upload(S.colmap.data(), S.R, st0);
    M.colscale.alloc(S.R); M.colscale.upload(S.colscale.data(), S.R, st0);

    // lane stride: a power of two >= the unique columns, two lanes per thread
    M.Rq = 2;
    while (M.Rq < S.Ru) M.Rq *= 2;
    M.ntile_job = np / kTileRows;
    const size_t ell = (size_t)W * np;
    const size_t V = (size_t)S.C * np * M.Rq;
    M.sv.alloc(ell * M.frames); M.rv.alloc(ell * M.frames);
    M.dinv.alloc((size_t)S.C * np * M.frames);
    M.x.alloc(V * M.frames); M.r.alloc(V * M.frames); M.p.alloc(V * M.frames); M.q.alloc(V * M.frames);
    M.part.alloc((size_t)kCParts * S.C * M.ntile_job * M.Rq * M.frames);
    M.sc.alloc((size_t)ZF_N * S.C * M.Rq * M.frames);
    M.st.alloc((size_t)M.frames * kMaxSlices * 4);
    M.live.alloc(1);
    M.gamma.alloc((size_t)S.E * M.frames);
    MPF_HIP(hipMemsetAsync(M.st.p, 0, M.st.bytes(), st0));
    MPF_HIP(hipHostMalloc((void**)&M.h_live, sizeof(int), hipHostMallocDefault));
    MPF_HIP(hipHostMalloc((void**)&M.h_st, M.st.bytes(), hipHostMallocDefault));

    M.hjobs.resize(M.frames);
    for (int f = 0; f < M.frames; ++f) {
        CJob& J = M.hjobs[f];
        std::memset(&J, 0, sizeof(J));
        J.n = S.N; J.np = np; J.width = W; J.C = S.C; J.Ru = S.Ru; J.Rq = M.Rq;
        J.ntiles = M.ntile_job;
        J.maxit = 10 * S.N;
        J.rule = MPF_STOP_REFERENCE;
        J.col0 = S.colmap[0];
        J.E = S.E; J.logarithmic = S.logarithmic;
        J.tol = 1e-6; J.ref = S.ref;
        J.ids = M.ids.p; J.diag_slot = M.diag_slot.p;
        J.sv = M.sv.p + ell * f; J.rv = M.rv.p + ell * f;
        J.dinv = M.dinv.p + (size_t)S.C * np * f;
        for (int k = 0; k < S.C; ++k) { J.alpha[k] = (float)S.alpha[k]; J.beta[k] = (float)S.beta[k]; }
        J.F = M.F.p; J.Fp = S.Rp;
        J.x = M.x.p + V * f; J.r = M.r.p + V * f; J.p = M.p.p + V * f; J.q = M.q.p + V * f;
        J.part = M.part.p + (size_t)kCParts * S.C * M.ntile_job * M.Rq * f;
        J.sc = M.sc.p + (size_t)ZF_N * S.C * M.Rq * f;
        J.st = M.st.p + (size_t)f * kMaxSlices * 4;
        J.gamma = M.gamma.p + (size_t)S.E * f;
        J.ctb_off = M.ctb_off.p; J.ctb_elem = M.ctb_elem.p; J.ctb_k = M.ctb_k.p; J.ctb_m = M.ctb_m.p;
    }
    M.jobs.alloc(M.frames);
    M.jobs.upload(M.hjobs.data(), M.frames, st0);

    std::vector<int4> ht, ha;
    std::vector<int2> hj;
    const int nasm_job = (int)((ell + 255) / 256);
    // (frame, tile, harmonic) order: a tile's harmonics share its matrix rows in L2
    for (int f = 0; f < M.frames; ++f) {
        for (int k = 0; k < S.C; ++k) hj.push_back(make_int2(f, k));
        for (int t = 0; t < M.ntile_job; ++t)
            for (int k = 0; k < S.C; ++k) ht.push_back(make_int4(f, k, t, 0));
        for (int b = 0; b < nasm_job; ++b) ha.push_back(make_int4(f, b, 0, 0));
    }
    M.ntiles = (int)ht.size(); M.njs = (int)hj.size(); M.nasm = (int)ha.size();
    std::vector<int4> hto(ht.size());
    xcd_order(ht.data(), M.ntiles, hto.data());
    M.tiles.alloc(M.ntiles); M.tiles.upload(hto.data(), hto.size(), st0);
    M.js.alloc(M.njs); M.js.upload(hj.data(), hj.size(), st0);
    M.asm_items.alloc(M.nasm); M.asm_items.upload(ha.data(), ha.size(), st0);
    MPF_HIP(hipStreamSynchronize(st0));
}

ComplexModel::~ComplexModel() = default;

const MeshSetup& ComplexModel::setup() const { return impl->S; }
int ComplexModel::frames() const { return impl->frames; }

int ComplexModel::forward(const mpf_solve_opts& o, const float2* d_gamma, float2* d_U, zd* d_U64, float2* d_J,
                          mpf_solve_stats* stats) {
    Impl& M = *impl;
    const MeshSetup& S = M.S;
    hipStream_t st0 = M.ctx->stream;
    MPF_REQUIRE(d_gamma, "solve: null gamma");
    MPF_REQUIRE(o.krylov == MPF_KRYLOV_CG, "complex conductivity: the CG solver only");
    MPF_REQUIRE(o.precision == MPF_PREC_FP32, "complex conductivity: FP32 CG only");
    MPF_REQUIRE(o.preconditioner == MPF_PRECOND_JACOBI, "complex conductivity: the diagonal preconditioner only");
    MPF_REQUIRE(o.stop_rule == MPF_STOP_REFERENCE || o.stop_rule == MPF_STOP_TOLERANCE, "solve: unknown stop rule");
    MPF_HIP(hipMemcpyAsync(M.gamma.p, d_gamma, sizeof(float2) * (size_t)S.E * M.frames, hipMemcpyDeviceToDevice,
                           st0));
    for (auto& J : M.hjobs) {
        J.herm = o.complex_hermitian ? 1 : 0;
        J.rule = o.stop_rule;
        J.tol = o.tol;
        J.maxit = o.max_iter > 0 ? o.max_iter : 10 * S.N;
    }
    M.jobs.upload(M.hjobs.data(), M.frames, st0);
    launch_c_assemble(M.jobs.p, M.asm_items.p, M.nasm, st0);
    launch_c_dinv(M.jobs.p, M.tiles.p, M.ntiles, st0);
    launch_c_init(M.jobs.p, M.tiles.p, M.ntiles, M.js.p, M.njs, st0);
    *M.h_live = M.njs;
    MPF_HIP(hipMemcpyAsync(M.live.p, M.h_live, sizeof(int), hipMemcpyHostToDevice, st0));
    const int every = o.check_every > 0 ? o.check_every : 16;
    const int maxit = M.hjobs[0].maxit;
    int it = 0;
    while (true) {
        for (int i = 0; i < every; ++i) launch_c_iter(M.jobs.p, M.tiles.p, M.ntiles, M.js.p, M.njs, M.live.p, st0);
        it += every;
        MPF_HIP(hipMemcpyAsync(M.h_live, M.live.p, sizeof(int), hipMemcpyDeviceToHost, st0));
        MPF_HIP(hipStreamSynchronize(st0));
        if (*M.h_live <= 0 || it >= maxit + every) break;
    }
    MPF_HIP(hipMemcpyAsync(M.h_st, M.st.p, M.st.bytes(), hipMemcpyDeviceToHost, st0));
    MPF_HIP(hipStreamSynchronize(st0));
    bool capped = false;
    for (int f = 0; f < M.frames; ++f)
        for (int k = 0; k < S.C; ++k) {
            const int* sp = M.h_st + ((size_t)f * kMaxSlices + k) * 4;
            if (sp[2] || sp[0]) capped = true;
            if (stats) {
                stats->steps += sp[1];
                stats->max_steps = std::max<int64_t>(stats->max_steps, sp[1]);
                stats->not_converged += (sp[2] || sp[0]) ? 1 : 0;
            }
        }
    CMeasDev md{M.w_off.p, M.w_node.p, M.w_val.p, M.colmap.p, M.colscale.p, S.D, S.M, S.Dp, S.Mp};
    if (d_U || d_U64) launch_c_voltages(M.jobs.p, M.frames, md, d_U, d_U64, st0);
    if (d_J) {
        CJacDev jd;
        std::memset(&jd, 0, sizeof(jd));
        jd.elements = M.elements.p; jd.ek = M.ek.p; jd.em = M.em.p;
        jd.colmap = M.colmap.p; jd.colscale = M.colscale.p;
        jd.E = S.E; jd.k = S.k; jd.D = S.D; jd.M = S.M; jd.Dp = S.Dp; jd.Mp = S.Mp; jd.exact = M.exact;
        for (int k = 0; k < S.C; ++k) {
            jd.alpha[k] = (float)S.alpha[k];
            // reference formula: height; exact 2.5D derivative: 1 / beta_c
            jd.weight[k] = M.exact ? (float)(1.0 / S.beta[k]) : (float)S.height;
        }
        launch_c_jacobian(M.jobs.p, M.frames, jd, d_J, st0);
    }
    MPF_HIP(hipStreamSynchronize(st0));
    return capped ? MPF_ENOCONV : MPF_OK;
}

int ComplexModel::gn_step(float2* d_gamma, const zd* d_Umeas, double lambda, const mpf_solve_opts& o, zd* d_Ucalc,
                          mpf_solve_stats* stats) {
    Impl& M = *impl;
    const MeshSetup& S = M.S;
    MPF_REQUIRE(d_gamma && d_Umeas, "gn_step: null argument");
    M.ensure_gn();
    hipStream_t st0 = M.ctx->stream;
    const int m = S.mrows, E = S.E;
    const int rc = forward(o, d_gamma, nullptr, M.U64.p, M.Jm.p, stats);
    if (rc != MPF_OK && rc != MPF_ENOCONV) return rc;
    if (d_Ucalc) MPF_HIP(hipMemcpyAsync(d_Ucalc, M.U64.p, sizeof(zd) * (size_t)m * M.frames, hipMemcpyDeviceToDevice, st0));
    // r = U_meas - U(gamma); (J J^T + lam^2 I) y = r; gamma += J^T y
    launch_c_sub(m * M.frames, d_Umeas, M.U64.p, M.rgn.p, st0);
    const double lam2 = (S.ref * lambda) * (S.ref * lambda);
    launch_c_gram(m, E, M.Jm.p, (size_t)m * E, lam2, M.G.p, (size_t)m * m, M.frames, st0);
    MPF_HIP(hipMemsetAsync(M.info.p, 0, sizeof(int), st0));
    launch_cholesky_z(m, M.G.p, (size_t)m * m, M.frames, M.info.p, M.pairs.p, st0);
    launch_potrs_z(m, M.G.p, (size_t)m * m, M.rgn.p, m, M.y.p, m, M.frames, st0);
    launch_c_gemv_t(m, E, M.Jm.p, (size_t)m * E, M.y.p, m, M.vE.p, E, M.frames, d_gamma, st0);
    int hinfo = 0;
    MPF_HIP(hipMemcpyAsync(&hinfo, M.info.p, sizeof(int), hipMemcpyDeviceToHost, st0));
    MPF_HIP(hipStreamSynchronize(st0));
    if (hinfo) throw Error(MPF_EINVAL, "gn_step: complex normal-equation matrix has a zero pivot");
    return rc;
}

}  // namespace mpf
