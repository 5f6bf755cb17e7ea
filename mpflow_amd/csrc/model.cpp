// Host orchestration of the EIT forward solve and the Tikhonov Gauss-Newton step.
//
// models::EIT::solve (src/models/eit.cpp:139-213) runs the C harmonics one after the
// other, each with ~13 launches and a device->host copy per CG iteration.  Here every
// harmonic of every frame iterates in the same launches; the host only polls a device
// counter of unconverged (frame, harmonic) pairs every `check_every` iterations.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "cem.h"
#include "model.h"
#include "fem.h"

std::pair<hipEvent_t, hipEvent_t> mpf_ctx::next_events() {
    if (ev_used == ev.size()) {
        hipEvent_t a, b;
        MPF_HIP(hipEventCreate(&a));
        MPF_HIP(hipEventCreate(&b));
        ev.emplace_back(a, b);
    }
    return ev[ev_used++];
}

void mpf_ctx::collect_events() {
    for (size_t i = 0; i < ev_used; ++i) {
        float ms = 0.0f;
        MPF_HIP(hipEventElapsedTime(&ms, ev[i].first, ev[i].second));
        prof_ms += ms;
        ++prof_launches;
    }
    ev_used = 0;
}

namespace mpf {

Model::Model(mpf_ctx* c, const mpf_model_desc& d) : ctx(c) {
    MPF_REQUIRE(d.frames >= 1, "model: frames < 1");
    frames = d.frames;
    quirks = d.quirks;
    build_setup(d, S);
    hipStream_t st0 = ctx->stream;
    const int np = S.np, W = S.wpad;
    ids.alloc((size_t)W * np); ids.upload(S.ids.data(), S.ids.size(), st0);
    diag_slot.alloc(np); diag_slot.upload(S.diag_slot.data(), np, st0);
    gw.alloc(np / 16); gw.upload(S.group_width.data(), np / 16, st0);
    aggq.alloc(np); aggq.upload(S.agg_q.data(), np, st0);
    ctb_off.alloc(S.ctb_off.size()); ctb_off.upload(S.ctb_off.data(), S.ctb_off.size(), st0);
    ctb_elem.alloc(S.ctb_elem.size()); ctb_elem.upload(S.ctb_elem.data(), S.ctb_elem.size(), st0);
    ctb_k.alloc(S.ctb_k.size()); ctb_k.upload(S.ctb_k.data(), S.ctb_k.size(), st0);
    ctb_m.alloc(S.ctb_m.size()); ctb_m.upload(S.ctb_m.data(), S.ctb_m.size(), st0);
    elements.alloc(S.elements.size()); elements.upload(S.elements.data(), S.elements.size(), st0);
    ek.alloc(S.elem_k.size()); ek.upload(S.elem_k.data(), S.elem_k.size(), st0);
    em.alloc(S.elem_m.size()); em.upload(S.elem_m.data(), S.elem_m.size(), st0);
    {
        // the Jacobian visits elements in the order of their lowest device node: the
        // device numbering is a recursive bisection, so consecutive elements share nodes
        std::vector<int64_t> key(S.E);
        for (int e = 0; e < S.E; ++e) {
            int lo = S.elements[(size_t)e * S.k];
            for (int a = 1; a < S.k; ++a) lo = std::min(lo, S.elements[(size_t)e * S.k + a]);
            key[e] = ((int64_t)lo << 32) | e;
        }
        std::sort(key.begin(), key.end());
        std::vector<int32_t> ord(S.E);
        for (int e = 0; e < S.E; ++e) ord[e] = (int32_t)(key[e] & 0xFFFFFFFF);
        jac_order.alloc(S.E);
        jac_order.upload(ord.data(), S.E, st0);
    }
    F.alloc(S.F.size()); F.upload(S.F.data(), S.F.size(), st0);
    if (S.fixed) {
        auto up_i = [&](DBuf<int32_t>& b, const std::vector<int32_t>& v) { b.alloc(std::max<size_t>(v.size(), 1)); if (!v.empty()) b.upload(v.data(), v.size(), st0); };
        auto up_d = [&](DBuf<double>& b, const std::vector<double>& v) { b.alloc(std::max<size_t>(v.size(), 1)); if (!v.empty()) b.upload(v.data(), v.size(), st0); };
        up_i(cpl_off, S.cpl_off); up_i(cpl_node, S.cpl_node); up_i(cpl_elem, S.cpl_elem);
        up_d(cpl_k, S.cpl_k); up_d(cpl_m, S.cpl_m);
        up_i(drow_off, S.drow_off); up_i(drow_node, S.drow_node); up_i(drow_elem, S.drow_elem);
        up_d(drow_k, S.drow_k); up_d(drow_m, S.drow_m);
    }
    w_off.alloc(S.w_off.size()); w_off.upload(S.w_off.data(), S.w_off.size(), st0);
    w_node.alloc(std::max<size_t>(1, S.w_node.size())); if (!S.w_node.empty()) w_node.upload(S.w_node.data(), S.w_node.size(), st0);
    w_val.alloc(std::max<size_t>(1, S.w_val.size())); if (!S.w_val.empty()) w_val.upload(S.w_val.data(), S.w_val.size(), st0);
    colmap.alloc(S.R); colmap.upload(S.colmap.data(), S.R, st0);
    colscale.alloc(S.R); colscale.upload(S.colscale.data(), S.R, st0);
    {
        std::vector<int32_t> inv(np);
        for (int r = 0; r < np; ++r) inv[r] = r;
        for (size_t n = 0; n < S.perm.size(); ++n) inv[S.perm[n]] = (int32_t)n;
        node_of_row.alloc(np); node_of_row.upload(inv.data(), np, st0);
    }

    const size_t ell = (size_t)W * np;
    const size_t V = vec_elems();
    const int ntile_job = np / kTileRows;
    s.alloc(ell * frames); r.alloc(ell * frames); s_lo.alloc(ell * frames); r_lo.alloc(ell * frames);
    dinv.alloc((size_t)S.C * np * frames);
    x.alloc(V * frames); rv.alloc(V * frames); p.alloc(V * frames); q.alloc(V * frames);
    pring.alloc(V * frames * (kRing - 1));
    rhs.alloc(V * frames);
    x64.alloc(V * frames);
    if (S.fixed) {
        fslice.alloc((size_t)S.C * np * S.Rp * frames);
        cur.alloc(V * frames);
        MPF_HIP(hipMemsetAsync(cur.p, 0, sizeof(double) * V * frames, st0));
    }
    part.alloc((size_t)kParts * S.C * ntile_job * S.Rp * frames);
    sc.alloc((size_t)SC_NFIELDS * S.C * S.Rp * frames);
    gamma.alloc((size_t)S.E * frames);
    const int nct = kAggPerTile * S.coarse_nt;   // coarse rows (aggregates)
    if (nct) {
        cinv.alloc(S.coarse_inv.size()); cinv.upload(S.coarse_inv.data(), S.coarse_inv.size(), st0);
        cd1.alloc(nct); cd1.upload(S.coarse_d1.data(), nct, st0);
        csc.alloc((size_t)nct * frames);
        cy.alloc((size_t)S.C * nct * S.Rp * frames);
        cg.alloc((size_t)S.C * nct * S.Rp * frames);
        MPF_HIP(hipMemsetAsync(cy.p, 0, cy.bytes(), st0));
    }
    MPF_HIP(hipMemsetAsync(x64.p, 0, x64.bytes(), st0));
    MPF_HIP(hipMemsetAsync(x.p, 0, x.bytes(), st0));

    hjobs.resize(frames);
    for (int f = 0; f < frames; ++f) {
        JobDev& J = hjobs[f];
        std::memset(&J, 0, sizeof(J));
        J.n = S.N; J.np = np; J.width = W; J.C = S.C; J.R = S.Ru; J.Rp = S.Rp;
        J.col0 = S.cem_coupled ? 0 : S.colmap[0];
        J.lg = S.Rp == 16 ? 4 : (S.Rp == 32 ? 5 : 6);
        J.ntiles = ntile_job;
        J.dc = 0;  // Open sources: harmonic 0 is solved dcFree (eit.cpp:175-177)
        J.maxit = 10 * S.N;
        J.rule = MPF_STOP_REFERENCE;
        J.tol = 1e-6;
        J.ids = ids.p;
        J.s = s.p + ell * f; J.r = r.p + ell * f; J.s_lo = s_lo.p + ell * f; J.r_lo = r_lo.p + ell * f;
        J.dinv = dinv.p + (size_t)S.C * np * f;
        for (int c = 0; c < S.C; ++c) {
            J.alpha[c] = (float)S.alpha[c]; J.xscale[c] = (float)S.beta[c];
            J.alpha64[c] = S.alpha[c]; J.xscale64[c] = S.beta[c];
        }
        J.F = F.p;
        if (S.cem_coupled)
            // the unit electrode sources Bc, the same for every harmonic
            for (int c = 0; c < S.C; ++c) { J.xscale[c] = 1.0f; J.xscale64[c] = 1.0; }
        if (S.fixed) {
            // voltage sources: one excitation per slice, no beta scaling
            J.F = fslice.p + (size_t)S.C * np * S.Rp * f;
            J.fstride = (size_t)np * S.Rp;
            J.cur = cur.p + V * f;
            for (int c = 0; c < S.C; ++c) { J.xscale[c] = 1.0f; J.xscale64[c] = 1.0; }
        }
        J.rhs = rhs.p + V * f;
        J.use_rhs = 0;
        J.x = x.p + V * f; J.rv = rv.p + V * f; J.p = p.p + V * f; J.q = q.p + V * f;
        for (int b = 0; b < kRing - 1; ++b) J.pr[b] = pring.p + V * frames * b + V * f;
        J.x64 = x64.p + V * f;
        J.part = part.p + (size_t)kParts * S.C * ntile_job * S.Rp * f;
        J.sc = sc.p + (size_t)SC_NFIELDS * S.C * S.Rp * f;
        J.st = nullptr;          // the Krylov driver points this at its own state
        J.active = nullptr;
        J.nnz = S.nnz;
        J.ctb_off = ctb_off.p; J.ctb_elem = ctb_elem.p; J.ctb_k = ctb_k.p; J.ctb_m = ctb_m.p;
        J.gamma = gamma.p + (size_t)S.E * f;
        J.ref = S.ref; J.logarithmic = S.logarithmic; J.E = S.E;
        J.diag_slot = diag_slot.p;
        J.gw = gw.p;
        J.aggq = aggq.p;
        if (nct) {
            J.cinv = cinv.p;
            J.cld = S.coarse_ld;
            J.csc = csc.p + (size_t)nct * f;
            J.cy = cy.p + (size_t)S.C * nct * S.Rp * f;
            J.cg = cg.p + (size_t)S.C * nct * S.Rp * f;
        }
    }
    jobs.alloc(frames);
    upload_jobs();
    if (S.cem_coupled) {
        const int L = S.P;
        bc_off.alloc(S.bc_off.size()); bc_off.upload(S.bc_off.data(), S.bc_off.size(), st0);
        bc_row.alloc(std::max<size_t>(1, S.bc_row.size())); bc_row.upload(S.bc_row.data(), S.bc_row.size(), st0);
        bc_val.alloc(std::max<size_t>(1, S.bc_val.size())); bc_val.upload(S.bc_val.data(), S.bc_val.size(), st0);
        elen.alloc(L); elen.upload(S.elen.data(), L, st0);
        pi.alloc(S.pi.size()); pi.upload(S.pi.data(), S.pi.size(), st0);
        prep.alloc(S.pRu); prep.upload(S.prep.data(), S.pRu, st0);
        cem_T.alloc((size_t)frames * S.C * L * S.pRp);
        cem_phi.alloc((size_t)frames * S.C * np * S.pRp);
        hjobs_phi = hjobs;
        for (int f = 0; f < frames; ++f) {
            JobDev& J = hjobs_phi[f];
            J.Rp = S.pRp;
            J.R = S.pRu;
            J.lg = S.pRp == 16 ? 4 : (S.pRp == 32 ? 5 : 6);
            J.x64 = cem_phi.p + (size_t)S.C * np * S.pRp * f;
        }
        jobs_phi.alloc(frames);
        jobs_phi.upload(hjobs_phi.data(), frames, st0);
    }

    // work tables
    std::vector<int4> ht;
    std::vector<int2> hj;
    std::vector<int4> ha;
    const int nasm_job = (int)((ell + kBlock - 1) / kBlock);
    // (frame, tile, harmonic) order: the harmonics of one tile run back to back on one XCD
    // and read the tile's matrix rows (ids, S, R) from its L2 after the first
    for (int f = 0; f < frames; ++f) {
        for (int c = 0; c < S.C; ++c) hj.push_back(make_int2(f, c));
        for (int t = 0; t < ntile_job; ++t)
            for (int c = 0; c < S.C; ++c) ht.push_back(make_int4(f, c, t, 0));
        for (int b = 0; b < nasm_job; ++b) ha.push_back(make_int4(f, b, 0, 0));
    }
    ntiles = (int)ht.size(); njs = (int)hj.size(); nasm = (int)ha.size();
    std::vector<int4> hto(ht.size());
    xcd_order(ht.data(), ntiles, hto.data());
    tiles.alloc(ntiles); tiles.upload(hto.data(), hto.size(), st0);
    js.alloc(njs); js.upload(hj.data(), hj.size(), st0);
    asm_items.alloc(nasm); asm_items.upload(ha.data(), ha.size(), st0);
    MPF_HIP(hipStreamSynchronize(st0));
}

Model::~Model() {
}

void Model::upload_jobs() {
    jobs.upload(hjobs.data(), hjobs.size(), ctx->stream);
}

// Per-model set-up of a forward solve: Krylov method checks, the dc handling of the
// job table, then Equation::update (S and R once per sigma, not once per harmonic) and
// the Jacobi diagonal.
void Model::prepare(const mpf_solve_opts& o) {
    hipStream_t st0 = ctx->stream;
    MPF_REQUIRE(o.krylov == MPF_KRYLOV_CG || o.krylov == MPF_KRYLOV_BICGSTAB, "solve: unknown Krylov method");
    MPF_REQUIRE(o.preconditioner == MPF_PRECOND_JACOBI || o.preconditioner == MPF_PRECOND_TWO_LEVEL,
                "solve: unknown preconditioner");
    if (o.krylov == MPF_KRYLOV_BICGSTAB) {
        MPF_REQUIRE(S.wpad <= 24, "BiCGSTAB: padded ELL width above 24 is not supported");
        ensure_bicg();
    }
    // Open sources: CG solves harmonic 0 dcFree (eit.cpp:175-177); the reference's BiCGSTAB
    // ignores the flag (bicgstab.cpp:87), which bicgstab_ignore_dc reproduces
    for (auto& J : hjobs) {
        J.dc = (o.krylov == MPF_KRYLOV_BICGSTAB && o.bicgstab_ignore_dc) ? -1 : 0;
        J.dc_rank1 = o.krylov == MPF_KRYLOV_CG ? 1 : 0;
        // Dirichlet rows, or the contact term of every mode: no constant mode
        if (S.fixed || S.cem_coupled) { J.dc = -1; J.dc_rank1 = 0; }
    }
    upload_jobs();
    launch_assemble(jobs.p, asm_items.p, nasm, st0);
    launch_dinv(jobs.p, tiles.p, ntiles, st0);
    if (S.fixed) launch_fixed_rhs(jobs.p, frames, fixdev(), st0);
    if (S.coarse_nt && o.preconditioner == MPF_PRECOND_TWO_LEVEL && o.krylov == MPF_KRYLOV_CG)
        launch_coarse_scale(jobs.p, tiles.p, ntiles, cd1.p, st0);
}

// FP64 refinement report, voltages and (optionally) the Jacobian after the solves.
void Model::finish(const mpf_solve_opts& o, float* d_U, double* d_U64, float* d_J, mpf_solve_stats* stats) {
    hipStream_t st0 = ctx->stream;
    const bool mixed = o.precision == MPF_PREC_MIXED;
    if (mixed && stats) {
        std::vector<double> hsc((size_t)SC_NFIELDS * S.C * S.Rp * frames);
        MPF_HIP(hipMemcpyAsync(hsc.data(), sc.p, hsc.size() * sizeof(double), hipMemcpyDeviceToHost, st0));
        MPF_HIP(hipStreamSynchronize(st0));
        double worst = stats->refine_residual;
        for (int f = 0; f < frames; ++f)
            for (int c = 0; c < S.C; ++c)
                for (int l = 0; l < S.Ru; ++l) {
                    const double* b = &hsc[(size_t)SC_NFIELDS * S.C * S.Rp * f];
                    const double rr = b[((size_t)SC_X1 * S.C + c) * S.Rp + l];
                    const double bb = b[((size_t)SC_X2 * S.C + c) * S.Rp + l];
                    // NaN compares false: report it rather than skip it
                    if (bb > 0) {
                        const double v = std::sqrt(rr / bb);
                        if (!(v <= worst)) worst = v;
                    }
                }
        stats->refine_residual = worst;
    }
    if (S.fixed) launch_fixed_currents(jobs.p, frames, fixdev(), mixed ? 1 : 0, st0);
    if (S.cem_coupled) {
        // Schur complement on the electrode potentials: voltages, then the pattern fields
        CemDev cd;
        std::memset(&cd, 0, sizeof(cd));
        cd.bc_off = bc_off.p; cd.bc_row = bc_row.p; cd.bc_val = bc_val.p; cd.elen = elen.p;
        cd.pi = pi.p; cd.prep = prep.p; cd.zp = S.zp;
        cd.L = S.P; cd.R = S.R; cd.D = S.D; cd.M = S.M; cd.Dp = S.Dp; cd.Mp = S.Mp; cd.pRu = S.pRu; cd.pRp = S.pRp;
        for (int c = 0; c < S.C; ++c) {
            const double hb = S.height * S.beta[c];
            cd.coef[c] = S.n_c[c] * hb * hb / S.zp;
            cd.tsc[c] = hb / S.zp;
        }
        launch_cem(jobs.p, frames, S.C, S.np, cd, mixed ? 1 : 0, d_U, d_U64, cem_T.p, cem_phi.p, st0);
    } else {
        MeasDev md{w_off.p, w_node.p, w_val.p, colmap.p, colscale.p, S.D, S.M, S.Dp, S.Mp};
        launch_voltages(jobs.p, frames, md, mixed ? 1 : 0, d_U, d_U64, st0);
    }
    if (d_J) {
        JacDev jd;
        std::memset(&jd, 0, sizeof(jd));
        jd.elements = elements.p; jd.ek = ek.p; jd.em = em.p; jd.gamma = gamma.p;
        jd.order = jac_order.p;
        jd.colmap = colmap.p; jd.colscale = colscale.p;
        jd.E = S.E; jd.k = S.k; jd.D = S.D; jd.M = S.M; jd.Dp = S.Dp; jd.Mp = S.Mp;
        jd.exact = S.jacobian_exact; jd.logarithmic = S.logarithmic; jd.ref = S.ref;
        for (int c = 0; c < S.C; ++c) {
            jd.alpha[c] = (float)S.alpha[c];
            jd.weight[c] = S.jacobian_exact ? (float)(1.0 / S.beta[c]) : (float)S.height;
            // voltage sources: jacobian x height without the current sources' sign
            // (eit.cpp:202-205); the kernel applies -sigma' w
            if (S.fixed) jd.weight[c] = -(float)S.height;
            // coupled electrodes: the mode weights n_c of the symmetric block system
            if (S.cem_coupled) jd.weight[c] = (float)S.n_c[c];
        }
        // the Gauss-Newton step's own Jacobian also gets its BF16 copy for the Gram
        uint16_t* jb = (d_J == Jm.p && Jb.p && reg_type == MPF_REG_IDENTITY) ? Jb.p : nullptr;
        if (S.cem_coupled) launch_jacobian(jobs_phi.p, frames, jd, 1, d_J, st0, jb);
        else launch_jacobian(jobs.p, frames, jd, mixed ? 1 : 0, d_J, st0, jb);
    }
}

FixDev Model::fixdev() const {
    FixDev fd;
    fd.cpl_off = cpl_off.p; fd.cpl_node = cpl_node.p; fd.cpl_elem = cpl_elem.p; fd.cpl_k = cpl_k.p; fd.cpl_m = cpl_m.p;
    fd.drow_off = drow_off.p; fd.drow_node = drow_node.p; fd.drow_elem = drow_elem.p;
    fd.drow_k = drow_k.p; fd.drow_m = drow_m.p;
    fd.Fsh = F.p;
    fd.np = S.np; fd.Rp = S.Rp; fd.C = S.C;
    return fd;
}

KrylovDriver& Model::driver() {
    if (!drv) drv.reset(new KrylovDriver(ctx, {&hjobs}));
    return *drv;
}

int Model::forward(const mpf_solve_opts& o, float* d_U, double* d_U64, float* d_J, mpf_solve_stats* stats) {
    std::vector<Model*> ms{this};
    return forward_models(ms, driver(), o, &d_U, &d_U64, &d_J, stats);
}

// One forward solve of every frame of every model (models::EIT::solve,
// src/models/eit.cpp:139-213): the Krylov iterations of all models run in the same
// launches (drv covers exactly these models' jobs); mixed precision adds FP64 residual
// refinement passes.
int forward_models(std::vector<Model*>& ms, KrylovDriver& drv, const mpf_solve_opts& o, float* const* d_U,
                   double* const* d_U64, float* const* d_J, mpf_solve_stats* stats) {
    MPF_REQUIRE(!ms.empty(), "solve: no models");
    hipStream_t st0 = ms[0]->ctx->stream;
    for (Model* m : ms) m->prepare(o);
    bool capped = false;
    const bool mixed = o.precision == MPF_PREC_MIXED;
    bool warm = mixed && o.warm_start;
    for (Model* m : ms) warm = warm && m->x64_valid;
    auto jobs_of = [](Model* m) { return m->jobs.p; };
    if (warm) {
        // start from the previous FP64 fields: solve for the correction of the FP64
        // residual at the new sigma, to the same absolute accuracy tol * ||beta F||
        for (Model* m : ms) launch_residual64(jobs_of(m), m->tiles.p, m->ntiles, m->js.p, m->njs, m->hjobs[0].lg,
                                                  (size_t)m->S.np * m->S.Rp * sizeof(double), st0);
        capped |= drv.run(o, MPF_STOP_TOLERANCE, o.tol, true, true, stats);
        for (Model* m : ms) launch_x64_update(jobs_of(m), m->tiles.p, m->ntiles, 0, st0);
    } else {
        capped |= drv.run(o, o.stop_rule, o.tol, false, false, stats);
        if (mixed)
            for (Model* m : ms) launch_x64_update(jobs_of(m), m->tiles.p, m->ntiles, 1, st0);
    }
    for (Model* m : ms) m->x64_valid = mixed;
    if (mixed) {
        for (int k = 0; k < o.refine_steps; ++k) {
            for (Model* m : ms) launch_residual64(jobs_of(m), m->tiles.p, m->ntiles, m->js.p, m->njs, m->hjobs[0].lg,
                                                  (size_t)m->S.np * m->S.Rp * sizeof(double), st0);
            capped |= drv.run(o, MPF_STOP_TOLERANCE, o.refine_tol, true, false, stats);
            for (Model* m : ms) launch_x64_update(jobs_of(m), m->tiles.p, m->ntiles, 0, st0);
        }
        // final FP64 residual for the report
        for (Model* m : ms) launch_residual64(jobs_of(m), m->tiles.p, m->ntiles, m->js.p, m->njs, m->hjobs[0].lg,
                                                  (size_t)m->S.np * m->S.Rp * sizeof(double), st0);
    }
    for (size_t i = 0; i < ms.size(); ++i)
        ms[i]->finish(o, d_U ? d_U[i] : nullptr, d_U64 ? d_U64[i] : nullptr, d_J ? d_J[i] : nullptr, stats);
    MPF_HIP(hipStreamSynchronize(st0));
    return capped ? MPF_ENOCONV : MPF_OK;
}

void Model::ensure_bicg() {
    if (rh.p) return;
    const size_t V = vec_elems() * frames;
    rh.alloc(V); yv.alloc(V); sv.alloc(V); zv.alloc(V); tv.alloc(V);
    const size_t Vf = vec_elems();
    for (int f = 0; f < frames; ++f) {
        JobDev& J = hjobs[f];
        J.rh = rh.p + Vf * f; J.yv = yv.p + Vf * f; J.sv = sv.p + Vf * f; J.zv = zv.p + Vf * f;
        J.tv = tv.p + Vf * f;
    }
}

void Model::ensure_gn() {
    const int m = S.mrows;
    if (Jm.p) return;
    Jm.alloc((size_t)frames * m * S.E);
    if (gram_bf16_applies(m, m)) Jb.alloc((size_t)frames * m * S.E);
    G.alloc((size_t)frames * m * m);
    A64.alloc((size_t)frames * m * m);
    U64.alloc((size_t)frames * m);
    y.alloc((size_t)frames * m); dy.alloc((size_t)frames * m); res.alloc((size_t)frames * m);
    ratio.alloc(frames);
    rgn.alloc((size_t)frames * m);
    vE.alloc((size_t)frames * S.E);
    work_n.alloc(gemv_n_workspace(m, S.E, frames) / sizeof(double) + 1);
    gram_work.alloc(std::max({gram_workspace_bytes(m, S.E), gram_nt_workspace_bytes(m, S.E, frames),
                              Jb.p ? gram_bf16_workspace_bytes(m, S.E, frames) : (size_t)0}));
    pairs.alloc(cholesky_pairs(m));
    info.alloc(1);
}

int Model::gn_step(float* d_gamma, const double* d_Umeas, double lambda, int gn_refine,
                   const mpf_solve_opts& o, double* d_Ucalc, mpf_solve_stats* stats) {
    gn_begin(d_gamma);
    const int rc = forward(o, nullptr, U64.p, Jm.p, stats);
    if (rc != MPF_OK && rc != MPF_ENOCONV) return rc;
    gn_dense(d_gamma, d_Umeas, lambda, gn_refine, d_Ucalc);
    return rc;
}

void Model::gn_begin(const float* d_gamma) {
    ensure_gn();
    MPF_HIP(hipMemcpyAsync(gamma.p, d_gamma, sizeof(float) * (size_t)S.E * frames, hipMemcpyDeviceToDevice,
                           ctx->stream));
}

// The dense part of Solver::solveAbsolute after the forward solve has filled U64 and Jm.
void Model::gn_dense(float* d_gamma, const double* d_Umeas, double lambda, int gn_refine, double* d_Ucalc) {
    hipStream_t st0 = ctx->stream;
    const int m = S.mrows, E = S.E;
    if (d_Ucalc)
        MPF_HIP(hipMemcpyAsync(d_Ucalc, U64.p, sizeof(double) * (size_t)m * frames, hipMemcpyDeviceToDevice, st0));
    // r = U_meas - U(gamma)  (Inverse::calcExcitation, inverse.cpp:163-193)
    launch_sub(m * frames, d_Umeas, U64.p, rgn.p, st0);
    // noser: (J^T J + lam^2 W^-2) delta = J^T r is solved as J W = Js, delta = W Js^T y
    if (reg_type != MPF_REG_IDENTITY) launch_scale_cols(m, E, Jm.p, m, (size_t)m * E, frames, reg_w.p, st0);
    // J J^T on MFMA, then (J J^T + lam I) by FP64 Cholesky
    if (gram_tile() == 256)
        launch_gram_bf16(Jb.p, (size_t)m * E, m, E, m, G.p, (size_t)m * m, frames, gram_work.p, st0);
    else if (gram_nt_applies(m, m))
        launch_gram_nt(Jm.p, (size_t)m * E, m, E, m, G.p, (size_t)m * m, frames, gram_work.p, st0);
    else
        for (int f = 0; f < frames; ++f)
            launch_gram(Jm.p + (size_t)f * m * E, 0, m, E, m, G.p + (size_t)f * m * m, gram_work.p, st0);
    const double lam2 = (S.ref * lambda) * (S.ref * lambda);
    MPF_HIP(hipMemsetAsync(info.p, 0, sizeof(int), st0));
    launch_cholesky(m, G.p, (size_t)m * m, lam2, A64.p, (size_t)m * m, frames, info.p, pairs.p, st0);
    normal_solve(Jm.p, (size_t)m * E, A64.p, (size_t)m * m, lam2, frames, gn_refine);
    // delta = J^T y ; gamma += delta   (solver.cpp:222)
    launch_gemv_t(m, E, Jm.p, m, (size_t)m * E, y.p, m, vE.p, E, frames, st0);
    if (reg_type == MPF_REG_LAPLACIAN) laplacian_cg(Jm.p, (size_t)m * E, A64.p, (size_t)m * m, lam2, frames);
    if (reg_type != MPF_REG_IDENTITY) launch_scale_vec(E, frames, vE.p, reg_w.p, st0);
    launch_gamma_update(E * frames, d_gamma, vE.p, st0);
    int hinfo = 0;
    MPF_HIP(hipMemcpyAsync(&hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost, st0));
    MPF_HIP(hipStreamSynchronize(st0));
    if (hinfo) throw Error(MPF_EINVAL, "gn_step: normal-equation matrix is not positive definite");
}

// Tile of the Gauss-Newton step's Gram: 256 (k_gram_bf16 on the BF16 copy of J that the
// Jacobian kernel writes with identity regularisation) or 128 (k_gram_nt, converting the
// FP32 J while staging).  Both round J to BF16 the same way; they differ in the order of
// the FP32 sums.  Where both apply, each is timed once per (rows, elements, frames) on the
// current J and the faster is kept in the context's cache.
int Model::gram_tile() {
    const int m = S.mrows, E = S.E;
    if (!(Jb.p && reg_type == MPF_REG_IDENTITY)) return 128;
    if (!gram_nt_applies(m, m)) return 256;
    const long long key = ((long long)m << 44) | ((long long)E << 16) | frames;
    {
        std::lock_guard<std::mutex> lk(ctx->tune.mu);
        auto it = ctx->tune.gram.find(key);
        if (it != ctx->tune.gram.end()) return it->second;
    }
    hipStream_t st0 = ctx->stream;
    hipEvent_t a, b;
    MPF_HIP(hipEventCreate(&a));
    MPF_HIP(hipEventCreate(&b));
    float ms[2] = {0.0f, 0.0f};
    for (int c = 0; c < 2; ++c) {
        auto run = [&]() {
            if (c == 0) launch_gram_nt(Jm.p, (size_t)m * E, m, E, m, G.p, (size_t)m * m, frames, gram_work.p, st0);
            else launch_gram_bf16(Jb.p, (size_t)m * E, m, E, m, G.p, (size_t)m * m, frames, gram_work.p, st0);
        };
        run();   // warm-up
        MPF_HIP(hipEventRecord(a, st0));
        for (int r = 0; r < 3; ++r) run();
        MPF_HIP(hipEventRecord(b, st0));
        MPF_HIP(hipEventSynchronize(b));
        MPF_HIP(hipEventElapsedTime(&ms[c], a, b));
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    const int tile = ms[1] < ms[0] ? 256 : 128;
    std::lock_guard<std::mutex> lk(ctx->tune.mu);
    ctx->tune.gram[key] = tile;
    return tile;
}

void Model::normal_solve(const float* Jp, size_t strideJ, const double* A, size_t strideA, double lam2, int nb,
                         int gn_refine) {
    hipStream_t st0 = ctx->stream;
    const int m = S.mrows, E = S.E;
    launch_potrs(m, A, strideA, rgn.p, m, y.p, m, nb, st0);
    // FP64 refinement of (J J^T + lam I) y = r with residuals from the FP32 Jacobian: at
    // most gn_refine passes, fewer once every frame's remaining error, predicted from the
    // last two corrections (|dy_k|^2 / |dy_k-1|, relative to |y|, the refinement converging
    // linearly), is below kRefineStop — far below what the FP32 Jacobian resolves
    constexpr double kRefineStop = 1e-10;
    const bool one_pass = gemv_tn_applies(m, m, strideJ, Jp);
    if ((int)ratio_h.size() < nb) ratio_h.assign(nb, 0.0);
    std::vector<double> prev(nb, 1.0);
    for (int k = 0; k < gn_refine; ++k) {
        if (one_pass) {
            launch_gemv_tn(m, E, Jp, m, strideJ, y.p, m, lam2, rgn.p, m, res.p, m, nb, work_n.p, st0);
        } else {
            launch_gemv_t(m, E, Jp, m, strideJ, y.p, m, vE.p, E, nb, st0);
            launch_gemv_n(m, E, Jp, m, strideJ, vE.p, E, lam2, y.p, m, rgn.p, m, res.p, m, nb, work_n.p, st0);
        }
        launch_potrs(m, A, strideA, res.p, m, dy.p, m, nb, st0);
        launch_add_to_ratio(m, nb, y.p, dy.p, ratio.p, st0);
        if (k + 1 == gn_refine) break;
        MPF_HIP(hipMemcpyAsync(ratio_h.data(), ratio.p, sizeof(double) * nb, hipMemcpyDeviceToHost, st0));
        MPF_HIP(hipStreamSynchronize(st0));
        bool done = true;
        for (int b = 0; b < nb; ++b) {
            const double r = ratio_h[b];
            done = done && std::isfinite(r) && r * r <= kRefineStop * prev[b];
            prev[b] = r > 0.0 ? r : 1e-300;
        }
        if (done) break;
    }
}

void Model::laplacian_cg(const float* K, size_t strideK, const double* A, size_t strideA, double mu, int nb) {
    hipStream_t st0 = ctx->stream;
    const int m = S.mrows, E = S.E;
    const size_t nE = (size_t)E * frames, nm = (size_t)m * frames;
    if (!cg_r.p) {
        cg_r.alloc(nE); cg_z.alloc(nE); cg_p.alloc(nE); cg_q.alloc(nE); cg_u.alloc(nE);
        cg_t.alloc(nm); cg_s.alloc(nm); cg_dots.alloc((size_t)4 * frames);
    }
    double* x = vE.p;
    const size_t NE = (size_t)E * nb;
    // q = K^T K v + mu (v + B v)
    auto apply = [&](const double* v, double* out) {
        launch_gemv_n(m, E, K, m, strideK, v, E, 0.0, nullptr, 0, nullptr, 0, cg_t.p, m, nb, work_n.p, st0);
        launch_gemv_t(m, E, K, m, strideK, cg_t.p, m, cg_u.p, E, nb, st0);
        launch_lap_combine(E, nb, cg_u.p, mu, v, lap_nb.p, lap_bval.p, out, st0);
    };
    // z = (K^T K + mu I)^-1 v = (v - K^T (K K^T + mu I)^-1 K v) / mu
    auto precond = [&](const double* v, double* z) {
        launch_gemv_n(m, E, K, m, strideK, v, E, 0.0, nullptr, 0, nullptr, 0, cg_t.p, m, nb, work_n.p, st0);
        launch_potrs(m, A, strideA, cg_t.p, m, cg_s.p, m, nb, st0);
        launch_gemv_t(m, E, K, m, strideK, cg_s.p, m, cg_u.p, E, nb, st0);
        launch_axmy(NE, 1.0 / mu, v, cg_u.p, z, st0);
    };
    std::vector<double> h((size_t)4 * nb);
    auto fetch = [&]() {
        MPF_HIP(hipMemcpyAsync(h.data(), cg_dots.p, sizeof(double) * 4 * nb, hipMemcpyDeviceToHost, st0));
        MPF_HIP(hipStreamSynchronize(st0));
    };
    // |K^T r|^2 per column (the stopping reference)
    launch_gemv_t(m, E, K, m, strideK, rgn.p, m, cg_q.p, E, nb, st0);
    launch_dots(E, nb, cg_q.p, cg_q.p, cg_dots.p + 3, 4, st0);
    // r = K^T (r_data - K x) - mu (x + B x)
    launch_gemv_n(m, E, K, m, strideK, x, E, 0.0, nullptr, 0, rgn.p, m, cg_t.p, m, nb, work_n.p, st0);
    launch_gemv_t(m, E, K, m, strideK, cg_t.p, m, cg_u.p, E, nb, st0);
    launch_lap_combine(E, nb, cg_u.p, -mu, x, lap_nb.p, lap_bval.p, cg_r.p, st0);
    precond(cg_r.p, cg_z.p);
    MPF_HIP(hipMemcpyAsync(cg_p.p, cg_z.p, sizeof(double) * NE, hipMemcpyDeviceToDevice, st0));
    launch_dots(E, nb, cg_r.p, cg_z.p, cg_dots.p + 0, 4, st0);
    launch_dots(E, nb, cg_r.p, cg_r.p, cg_dots.p + 1, 4, st0);
    fetch();
    constexpr double kTol = 1e-10;
    constexpr int kMaxIt = 2000;
    std::vector<double> rz(nb), bb(nb), coef(nb);
    std::vector<char> done(nb, 0);
    DBuf<double> dcoef(nb);
    int it = 0;
    auto converged = [&]() {
        bool all = true;
        for (int b = 0; b < nb; ++b) {
            if (!done[b] && !(h[4 * b + 1] > kTol * kTol * bb[b])) done[b] = 1;
            all = all && done[b];
        }
        return all;
    };
    for (int b = 0; b < nb; ++b) { rz[b] = h[4 * b]; bb[b] = h[4 * b + 3]; }
    while (!converged() && it < kMaxIt) {
        ++it;
        apply(cg_p.p, cg_q.p);
        launch_dots(E, nb, cg_p.p, cg_q.p, cg_dots.p + 2, 4, st0);
        fetch();
        for (int b = 0; b < nb; ++b) coef[b] = (done[b] || !(h[4 * b + 2] > 0.0)) ? 0.0 : rz[b] / h[4 * b + 2];
        dcoef.upload(coef.data(), nb, st0);
        launch_cg_xr(E, nb, dcoef.p, x, cg_p.p, cg_r.p, cg_q.p, st0);
        precond(cg_r.p, cg_z.p);
        launch_dots(E, nb, cg_r.p, cg_z.p, cg_dots.p + 0, 4, st0);
        launch_dots(E, nb, cg_r.p, cg_r.p, cg_dots.p + 1, 4, st0);
        fetch();
        for (int b = 0; b < nb; ++b) {
            const double rzn = h[4 * b];
            coef[b] = (done[b] || rz[b] == 0.0) ? 0.0 : rzn / rz[b];
            rz[b] = rzn;
        }
        dcoef.upload(coef.data(), nb, st0);
        launch_cg_p(E, nb, dcoef.p, cg_p.p, cg_z.p, st0);
    }
    MPF_HIP(hipStreamSynchronize(st0));
    lap_iterations = it;
}

int Model::forward_at(const float* d_gamma_ref, const mpf_solve_opts& o, mpf_solve_stats* stats) {
    ensure_gn();
    for (int f = 0; f < frames; ++f)
        MPF_HIP(hipMemcpyAsync(gamma.p + (size_t)f * S.E, d_gamma_ref, sizeof(float) * S.E, hipMemcpyDeviceToDevice,
                               ctx->stream));
    return forward(o, nullptr, U64.p, Jm.p, stats);
}

void Model::set_regularization(int type, const float* d_gamma_ref, const mpf_solve_opts& o,
                               mpf_solve_stats* stats) {
    MPF_REQUIRE(type == MPF_REG_IDENTITY || type == MPF_REG_NOSER || type == MPF_REG_LAPLACIAN,
                "regularization: unknown type");
    reg_type = MPF_REG_IDENTITY;
    diff_ready = false;
    if (type == MPF_REG_IDENTITY) return;
    hipStream_t st0 = ctx->stream;
    const int m = S.mrows, E = S.E;
    if (type == MPF_REG_LAPLACIAN) {
        // R = L^T L from the mesh alone: w = diag(R)^-1/2 and B = W (R - diag R) W
        std::vector<float> w(E);
        for (int e = 0; e < E; ++e) {
            MPF_REQUIRE(S.lap_diag[e] > 0.0, "regularization: degenerate element in the laplacian");
            w[e] = (float)(1.0 / std::sqrt(S.lap_diag[e]));
        }
        std::vector<double> bv((size_t)E * 3, 0.0);
        for (int e = 0; e < E; ++e)
            for (int k = 0; k < 3; ++k) {
                const int f = S.lap_nb[(size_t)e * 3 + k];
                if (f >= 0) bv[(size_t)e * 3 + k] = S.lap_val[(size_t)e * 3 + k] * (double)w[e] * (double)w[f];
            }
        if (!reg_w.p) reg_w.alloc(E);
        if (!lap_nb.p) { lap_nb.alloc((size_t)E * 3); lap_bval.alloc((size_t)E * 3); }
        reg_w.upload(w.data(), E, st0);
        lap_nb.upload(S.lap_nb.data(), (size_t)E * 3, st0);
        lap_bval.upload(bv.data(), (size_t)E * 3, st0);
        MPF_HIP(hipStreamSynchronize(st0));
        reg_type = type;
        return;
    }
    std::vector<float> g1;
    DBuf<float> gref;
    if (!d_gamma_ref) {
        // the Solver's reference distribution: 1 (linear) or 0 (logarithmic), solver.cpp:55-62
        g1.assign(E, S.logarithmic ? 0.0f : 1.0f);
        gref.alloc(E);
        gref.upload(g1.data(), E, st0);
        d_gamma_ref = gref.p;
    }
    const int rc = forward_at(d_gamma_ref, o, stats);
    MPF_REQUIRE(rc == MPF_OK, "regularization: the reference forward solve did not converge");
    if (has_mask) launch_mul((size_t)m * E, Jm.p, sens_mask.p, st0);
    DBuf<double> d(E);
    launch_col_sq(m, E, Jm.p, m, d.p, st0);
    std::vector<double> hd(E);
    MPF_HIP(hipMemcpyAsync(hd.data(), d.p, sizeof(double) * E, hipMemcpyDeviceToHost, st0));
    MPF_HIP(hipStreamSynchronize(st0));
    double dmax = 0.0;
    for (double v : hd) dmax = std::max(dmax, v);
    MPF_REQUIRE(dmax > 0.0, "regularization: zero Jacobian at the reference gamma");
    // an element without sensitivity would make R singular; its weight is capped
    const double floor = 1e-12 * dmax;
    std::vector<float> w(E);
    for (int e = 0; e < E; ++e) w[e] = (float)(1.0 / std::sqrt(std::max(hd[e], floor)));
    if (!reg_w.p) reg_w.alloc(E);
    reg_w.upload(w.data(), E, st0);
    MPF_HIP(hipStreamSynchronize(st0));
    reg_type = type;
}

void Model::diff_prepare(const float* d_gamma_ref, double lambda, const mpf_solve_opts& o, mpf_solve_stats* stats) {
    hipStream_t st0 = ctx->stream;
    const int m = S.mrows, E = S.E;
    diff_ready = false;
    const int rc = forward_at(d_gamma_ref, o, stats);
    MPF_REQUIRE(rc == MPF_OK, "differential: the reference forward solve did not converge");
    if (!diff_J.p) {
        diff_J.alloc((size_t)m * E);
        diff_A.alloc((size_t)m * m);
        diff_U.alloc(m);
        diff_gamma.alloc(E);
    }
    // frame 0's Jacobian (every frame was solved at the same gamma)
    MPF_HIP(hipMemcpyAsync(diff_J.p, Jm.p, sizeof(float) * (size_t)m * E, hipMemcpyDeviceToDevice, st0));
    MPF_HIP(hipMemcpyAsync(diff_U.p, U64.p, sizeof(double) * m, hipMemcpyDeviceToDevice, st0));
    MPF_HIP(hipMemcpyAsync(diff_gamma.p, d_gamma_ref, sizeof(float) * E, hipMemcpyDeviceToDevice, st0));
    if (has_mask) launch_mul((size_t)m * E, diff_J.p, sens_mask.p, st0);
    if (reg_type != MPF_REG_IDENTITY) launch_scale_cols(m, E, diff_J.p, m, 0, 1, reg_w.p, st0);
    if (gram_nt_applies(m, m))
        launch_gram_nt(diff_J.p, 0, m, E, m, G.p, 0, 1, gram_work.p, st0);
    else
        launch_gram(diff_J.p, 0, m, E, m, G.p, gram_work.p, st0);
    diff_lam2 = (S.ref * lambda) * (S.ref * lambda);
    MPF_HIP(hipMemsetAsync(info.p, 0, sizeof(int), st0));
    launch_cholesky(m, G.p, 0, diff_lam2, diff_A.p, 0, 1, info.p, pairs.p, st0);
    int hinfo = 0;
    MPF_HIP(hipMemcpyAsync(&hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost, st0));
    MPF_HIP(hipStreamSynchronize(st0));
    if (hinfo) throw Error(MPF_EINVAL, "differential: normal-equation matrix is not positive definite");
    diff_ready = true;
}

// Solver::fromConfig's noSensitivity mask (solver.cpp:107-113)
void Model::set_sensitivity_mask(const float* d_mask) {
    diff_ready = false;
    has_mask = false;
    if (!d_mask) return;
    const size_t n = (size_t)S.mrows * S.E;
    if (!sens_mask.p) sens_mask.alloc(n);
    MPF_HIP(hipMemcpyAsync(sens_mask.p, d_mask, sizeof(float) * n, hipMemcpyDeviceToDevice, ctx->stream));
    MPF_HIP(hipStreamSynchronize(ctx->stream));
    has_mask = true;
}

void Model::diff_solve(const double* d_Umeas, int gn_refine, float* d_gamma_out) {
    MPF_REQUIRE(diff_ready, "differential: call the preparation first");
    hipStream_t st0 = ctx->stream;
    const int m = S.mrows, E = S.E;
    launch_sub_bcast(m, frames, d_Umeas, diff_U.p, rgn.p, st0);
    normal_solve(diff_J.p, 0, diff_A.p, 0, diff_lam2, frames, gn_refine);
    launch_gemv_t(m, E, diff_J.p, m, 0, y.p, m, vE.p, E, frames, st0);
    if (reg_type == MPF_REG_LAPLACIAN) laplacian_cg(diff_J.p, 0, diff_A.p, 0, diff_lam2, frames);
    if (reg_type != MPF_REG_IDENTITY) launch_scale_vec(E, frames, vE.p, reg_w.p, st0);
    launch_gamma_from(E, frames, diff_gamma.p, vE.p, d_gamma_out, st0);
    MPF_HIP(hipStreamSynchronize(st0));
}

Batch::Batch(mpf_ctx* c, std::vector<Model*> ms) : ctx(c), models(std::move(ms)) {
    MPF_REQUIRE(!models.empty(), "batch: no models");
    std::vector<std::vector<JobDev>*> src;
    for (Model* m : models) {
        MPF_REQUIRE(m->ctx == ctx, "batch: every model must belong to the batch's context");
        src.push_back(&m->hjobs);
    }
    drv.reset(new KrylovDriver(ctx, src));
}

int Batch::solve(const float* const* d_gamma, float* const* d_U, double* const* d_U64, float* const* d_J,
                 const mpf_solve_opts& o, mpf_solve_stats* stats) {
    for (size_t i = 0; i < models.size(); ++i) {
        Model& M = *models[i];
        MPF_REQUIRE(d_gamma[i], "batch solve: null gamma");
        MPF_HIP(hipMemcpyAsync(M.gamma.p, d_gamma[i], sizeof(float) * (size_t)M.S.E * M.frames,
                               hipMemcpyDeviceToDevice, ctx->stream));
    }
    return forward_models(models, *drv, o, d_U, d_U64, d_J, stats);
}

int Batch::gn_step(float* const* d_gamma, const double* const* d_Umeas, double lambda, int gn_refine,
                   const mpf_solve_opts& o, double* const* d_Ucalc, mpf_solve_stats* stats) {
    std::vector<double*> U64s;
    std::vector<float*> Jms;
    for (size_t i = 0; i < models.size(); ++i) {
        MPF_REQUIRE(d_gamma[i] && d_Umeas[i], "batch gn_step: null argument");
        models[i]->gn_begin(d_gamma[i]);
        U64s.push_back(models[i]->U64.p);
        Jms.push_back(models[i]->Jm.p);
    }
    const int rc = forward_models(models, *drv, o, nullptr, U64s.data(), Jms.data(), stats);
    if (rc != MPF_OK && rc != MPF_ENOCONV) return rc;
    for (size_t i = 0; i < models.size(); ++i)
        models[i]->gn_dense(d_gamma[i], d_Umeas[i], lambda, gn_refine, d_Ucalc ? d_Ucalc[i] : nullptr);
    return rc;
}

}  // namespace mpf
