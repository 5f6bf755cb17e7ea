// Krylov driver (see krylov.h).
//
// ConjugateGradient::solve / BiCGSTAB::solve (src/numeric/conjugate_gradient.cpp:61-161,
// src/numeric/bicgstab.cpp:83-193) copy the residual matrix to the host and synchronise
// after every iteration, for one harmonic of one frame at a time.  Here all (job,
// harmonic) pairs iterate in the same launches; the host polls the device state every
// `check_every` iterations and then drops converged pairs from the work lists.  With
// opts.graphs the iterations between polls are captured once as a HIP graph and replayed
// while the lists do not change (lists are compacted only when a quarter of their tiles
// belong to finished pairs).
#include "krylov.h"

#include <algorithm>
#include <cstring>

#include "model.h"

namespace mpf {

constexpr int kXcds = 8;

// Workgroups are dispatched round-robin over the XCDs (block b runs on XCD b % 8), and
// each XCD has its own L2.  Reorder a work list so that XCD x receives a contiguous run
// of consecutive tiles: neighbouring rows gathered by the SpMV then hit the same L2.
void xcd_order(const int4* in, int n, int4* out) {
    const int base = n / kXcds, extra = n % kXcds;
    int start[kXcds + 1];
    start[0] = 0;
    for (int x = 0; x < kXcds; ++x) start[x + 1] = start[x] + base + (x < extra ? 1 : 0);
    for (int b = 0; b < n; ++b) out[b] = in[start[b % kXcds] + b / kXcds];
}

KrylovDriver::KrylovDriver(mpf_ctx* c, std::vector<std::vector<JobDev>*> sources)
    : ctx(c), src(std::move(sources)) {
    for (size_t m = 0; m < src.size(); ++m)
        for (size_t f = 0; f < src[m]->size(); ++f) {
            origin.emplace_back((int)m, (int)f);
            h.push_back((*src[m])[f]);
        }
    const int nj = (int)h.size();
    MPF_REQUIRE(nj > 0, "krylov: no jobs");
    jobs.alloc(nj);
    st.alloc((size_t)nj * kMaxSlices * kStInts);
    active.alloc(1);
    hipStream_t s0 = ctx->stream;
    MPF_HIP(hipMemsetAsync(st.p, 0, st.bytes(), s0));
    // groups by (lane count, padded width)
    std::vector<int4> all_tiles;
    std::vector<int2> all_js;
    for (int j = 0; j < nj; ++j) {
        const JobDev& J = h[j];
        Group* g = nullptr;
        for (auto& q : groups)
            if (q.lg == J.lg && q.wpad == J.width) g = &q;
        if (!g) {
            groups.emplace_back();
            g = &groups.back();
            g->lg = J.lg;
            g->wpad = J.width;
        }
        g->jobs.push_back(j);
        // the 128-bit gather kernels keep byte offsets of a slice vector in 31 bits
        if ((long long)J.np * J.Rp * 4 >= (1LL << 31) - (1LL << 16)) g->wide = true;
        for (int c = 0; c < J.C; ++c) {
            all_js.push_back(make_int2(j, c));
            for (int t = 0; t < J.ntiles; ++t) all_tiles.push_back(make_int4(j, c, t, 0));
        }
    }
    size_t total = 0;
    for (auto& g : groups) {
        std::vector<int4> lst, lst_t, flst;
        std::vector<int2> gjs;
        long long rows = 0;
        for (int j : g.jobs) {
            for (int c = 0; c < h[j].C; ++c) {
                gjs.push_back(make_int2(j, c));
                for (int t = 0; t < h[j].ntiles; ++t) lst.push_back(make_int4(j, c, t, 0));
            }
            for (int t = 0; t < h[j].ntiles; ++t) {
                for (int c = 0; c < h[j].C; ++c) lst_t.push_back(make_int4(j, c, t, 0));
                flst.push_back(make_int4(j, 0, t, 0));
            }
            rows += (long long)h[j].C * h[j].np;
            g.nchunk = std::max(g.nchunk, (h[j].ntiles + kChunkTiles - 1) / kChunkTiles);
            if (h[j].cinv) init_nchunk = std::max(init_nchunk, g.nchunk);
        }
        g.njs = (int)gjs.size();
        g.js.alloc(g.njs);
        g.js.upload(gjs.data(), gjs.size(), s0);
        g.nall = (int)lst.size();
        const std::vector<int4>* src_o[2] = {&lst, &lst_t};
        for (int o = 0; o < 2; ++o) {
            g.base_o[o] = *src_o[o];
            g.all_o[o].resize(lst.size());
            xcd_order(src_o[o]->data(), g.nall, g.all_o[o].data());
            g.tiles_o[o].alloc(g.nall);
            g.tiles_o[o].upload(g.all_o[o].data(), g.all_o[o].size(), s0);
        }
        g.act.alloc(g.nall);
        g.hoff = total;
        total += g.nall;
        g.nfall = (int)flst.size();
        g.fbase = flst;
        g.fall.resize(flst.size());
        xcd_order(flst.data(), g.nfall, g.fall.data());
        g.ftiles.alloc(g.nfall);
        g.ftiles.upload(g.fall.data(), g.fall.size(), s0);
        g.fact.alloc(g.nfall);
        g.fhoff = total;
        total += g.nfall;
        // autotune key: lane count, width, and the power-of-two bucket of the rows per
        // launch (small launches favour different kernels than large ones)
        int bucket = 0;
        while ((1LL << bucket) < rows) ++bucket;
        g.key = ((long long)g.lg << 40) | ((long long)g.wpad << 32) | bucket;
    }
    n_init_tiles = (int)all_tiles.size();
    n_init_js = (int)all_js.size();
    std::vector<int4> ordered(all_tiles.size());
    xcd_order(all_tiles.data(), n_init_tiles, ordered.data());
    init_tiles.alloc(n_init_tiles);
    init_tiles.upload(ordered.data(), ordered.size(), s0);
    init_js.alloc(n_init_js);
    init_js.upload(all_js.data(), all_js.size(), s0);
    MPF_HIP(hipHostMalloc((void**)&h_active, sizeof(int), hipHostMallocDefault));
    MPF_HIP(hipHostMalloc((void**)&h_st, st.bytes(), hipHostMallocDefault));
    std::memset(h_st, 0, st.bytes());
    h_tiles_cap = total;
    MPF_HIP(hipHostMalloc((void**)&h_tiles, sizeof(int4) * h_tiles_cap, hipHostMallocDefault));
    MPF_HIP(hipStreamSynchronize(s0));
}

KrylovDriver::~KrylovDriver() {
    drop_graph();
    if (cap) (void)hipStreamDestroy(cap);
    if (h_active) (void)hipHostFree(h_active);
    if (h_st) (void)hipHostFree(h_st);
    if (h_tiles) (void)hipHostFree(h_tiles);
}

void KrylovDriver::drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    gexec = nullptr;
}

void KrylovDriver::launch_group_iter(Group& g, int iter, bool bicg, hipStream_t s, hipEvent_t e0,
                                     hipEvent_t e1) {
    if (g.ncur == 0) return;
    if (bicg)
        launch_bicg_iter(jobs.p, g.cur, g.ncur, active.p, iter, g.lg, g.wpad, s, e0, e1);
    else
        launch_pcg_iter(jobs.p, g.cur, g.ncur, active.p, iter, g.lg, g.wpad, g.variant, s, e0, e1, g.js.p, g.njs,
                        g.nchunk, g.fcur, g.nfcur);
}

// Time each SpMV candidate that supports the group's width on the live job data (the
// SpMV only writes q and the SpMV partials, which the next iteration recomputes) and keep
// the fastest in the context's cache.  Candidates: the 128-bit gather kernel over
// (job, slice, tile) and over (job, tile, slice) work lists (in the second, the harmonics
// of one tile run back to back on one XCD and share its L2 copy of the tile's matrix),
// the 32-bit gather kernel, and the harmonic-fused kernel over (job, tile).
void KrylovDriver::tune(Group& g) {
    int choice = -1;
    {
        std::lock_guard<std::mutex> lk(ctx->tune.mu);
        auto it = ctx->tune.spmv.find(g.key);
        if (it != ctx->tune.spmv.end()) choice = it->second;
    }
    if (choice < 0) {
        hipStream_t s0 = ctx->stream;
        hipEvent_t a, b;
        MPF_HIP(hipEventCreate(&a));
        MPF_HIP(hipEventCreate(&b));
        float best = 1e30f;
        choice = g.wpad <= 24 && !g.wide ? 0 : 1;
        const int cand[4][2] = {{0, 0}, {0, 1}, {1, 0}, {2, 0}};
        for (const auto& c : cand) {
            const int v = c[0], o = c[1];
            if (v != 1 && (g.wpad > 24 || g.wide)) continue;
            const int4* lst = v == 2 ? g.ftiles.p : g.tiles_o[o].p;
            const int n = v == 2 ? g.nfall : g.nall;
            launch_pcg_spmv(jobs.p, lst, n, g.lg, g.wpad, v, s0);   // warm-up
            MPF_HIP(hipEventRecord(a, s0));
            for (int r = 0; r < 8; ++r) launch_pcg_spmv(jobs.p, lst, n, g.lg, g.wpad, v, s0);
            MPF_HIP(hipEventRecord(b, s0));
            MPF_HIP(hipEventSynchronize(b));
            float ms = 0.0f;
            MPF_HIP(hipEventElapsedTime(&ms, a, b));
            if (ms < best) { best = ms; choice = v | (o << 4); }
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        std::lock_guard<std::mutex> lk(ctx->tune.mu);
        ctx->tune.spmv[g.key] = choice;
    }
    g.variant = choice & 15;
    g.order = choice >> 4;
}

bool KrylovDriver::run(const mpf_solve_opts& o, int rule, double tol, bool use_rhs, bool ref_b,
                       mpf_solve_stats* stats) {
    hipStream_t s0 = ctx->stream;
    const bool bicg = o.krylov == MPF_KRYLOV_BICGSTAB;
    const int nj = (int)h.size();
    // the two-level preconditioner (jobs with a coarse level; CG only)
    const bool want_coarse = !bicg && o.preconditioner == MPF_PRECOND_TWO_LEVEL;
    if (want_coarse != coarse) drop_graph();
    coarse = want_coarse && init_nchunk > 0;
    for (int j = 0; j < nj; ++j) {
        JobDev& J = h[j];
        J = (*src[origin[j].first])[origin[j].second];
        if (!coarse) J.cinv = nullptr;
        J.st = st.p + (size_t)j * kMaxSlices * kStInts;
        J.rule = rule;
        J.tol = tol;
        J.maxit = o.max_iter > 0 ? o.max_iter : 10 * J.n;
        J.use_rhs = use_rhs ? 1 : 0;
        J.ref_b = ref_b ? 1 : 0;
    }
    jobs.upload(h.data(), h.size(), s0);
    MPF_HIP(hipMemsetAsync(active.p, 0, sizeof(int), s0));
    if (bicg) launch_bicg_init(jobs.p, init_tiles.p, n_init_tiles, init_js.p, n_init_js, active.p, 0, s0);
    else launch_pcg_init(jobs.p, init_tiles.p, n_init_tiles, init_js.p, n_init_js, active.p, s0,
                         coarse ? init_nchunk : 0);
    for (auto& g : groups) {
        if (!bicg) tune(g);
        else g.order = 0;
        g.cur = g.tiles_o[g.order].p;
        g.ncur = g.nall;
        g.fcur = g.ftiles.p;
        g.nfcur = g.nfall;
    }
    int every = o.check_every > 0 ? o.check_every : 16;
    every = (every + kRing - 1) / kRing * kRing;   // whole direction-ring cycles per poll
    const bool graphs = o.graphs != 0 && !ctx->prof;
    int maxit = 0;
    for (const auto& J : h) maxit = std::max(maxit, J.maxit);
    drop_graph();
    int it = 0;
    while (true) {
        if (graphs) {
            if (!gexec || gbicg != bicg || gevery != every) {
                drop_graph();
                if (!cap) MPF_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
                hipGraph_t graph;
                MPF_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
                for (int i = 0; i < every; ++i)
                    for (auto& g : groups) launch_group_iter(g, i, bicg, cap, nullptr, nullptr);
                MPF_HIP(hipStreamEndCapture(cap, &graph));
                const hipError_t e = hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                MPF_HIP(e);
                gbicg = bicg;
                gevery = every;
            }
            MPF_HIP(hipGraphLaunch(gexec, s0));
        } else {
            for (int i = 0; i < every; ++i)
                for (auto& g : groups) {
                    std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
                    if (ctx->prof && g.ncur) ev = ctx->next_events();
                    launch_group_iter(g, it + i, bicg, s0, ev.first, ev.second);
                }
        }
        it += every;
        MPF_HIP(hipMemcpyAsync(h_active, active.p, sizeof(int), hipMemcpyDeviceToHost, s0));
        MPF_HIP(hipMemcpyAsync(h_st, st.p, st.bytes(), hipMemcpyDeviceToHost, s0));
        MPF_HIP(hipStreamSynchronize(s0));
        if (ctx->prof) ctx->collect_events();
        if (*h_active <= 0 || it >= maxit + every) break;
        // compact the work lists: always without graphs, otherwise once a quarter of the
        // listed tiles belong to finished pairs (finished pairs' blocks exit at once)
        for (auto& g : groups) {
            // the survivors in (job, slice, tile) or (job, tile, slice) order (a filter of
            // the unstriped list keeps its order), then striped over the XCDs again so that
            // XCD x gets a contiguous run of them
            std::vector<int4> un;
            un.reserve(g.ncur);
            for (const int4& t : g.base_o[g.order])
                if (h_st[((size_t)t.x * kMaxSlices + t.y) * kStInts]) un.push_back(t);
            const int nt = (int)un.size();
            if (nt == g.ncur) continue;
            if (graphs && nt * 4 > g.ncur * 3 && nt > 0) continue;
            int4* ordered = h_tiles + g.hoff;
            xcd_order(un.data(), nt, ordered);
            if (nt) MPF_HIP(hipMemcpyAsync(g.act.p, ordered, sizeof(int4) * nt, hipMemcpyHostToDevice, s0));
            // the group's staging region is rewritten only after the next poll's
            // synchronisation
            g.cur = g.act.p;
            g.ncur = nt;
            // the fused SpMV's (job, tile) items: jobs with a live harmonic
            std::vector<int4> fun;
            fun.reserve(g.nfall);
            for (const int4& t : g.fbase) {
                bool live = false;
                for (int c = 0; c < h[t.x].C; ++c) live = live || h_st[((size_t)t.x * kMaxSlices + c) * kStInts];
                if (live) fun.push_back(t);
            }
            const int nf = (int)fun.size();
            int4* fordered = h_tiles + g.fhoff;
            xcd_order(fun.data(), nf, fordered);
            if (nf) MPF_HIP(hipMemcpyAsync(g.fact.p, fordered, sizeof(int4) * nf, hipMemcpyHostToDevice, s0));
            g.fcur = g.fact.p;
            g.nfcur = nf;
            drop_graph();
        }
    }
    drop_graph();
    if (!bicg) launch_pcg_flush(jobs.p, init_tiles.p, n_init_tiles, s0);
    bool capped = false;
    int64_t slice_iters = 0;
    double bytes = 0.0;
    for (int j = 0; j < nj; ++j)
        for (int c = 0; c < h[j].C; ++c) {
            const int* sp = &h_st[((size_t)j * kMaxSlices + c) * kStInts];
            slice_iters += sp[1];
            if (stats) {
                stats->steps += sp[1];
                stats->max_steps = std::max<int64_t>(stats->max_steps, sp[1]);
                stats->not_converged += sp[2];
            }
            if (sp[2]) capped = true;
            // algorithmic bytes of one harmonic's SpMV with R right-hand sides: A_c values
            // and column ids once, p read and q written once
            bytes += (double)sp[1] * (8.0 * h[j].nnz + 8.0 * h[j].n * h[j].R);
        }
    if (ctx->prof) ctx->prof_bytes += bytes;
    (void)slice_iters;
    return capped;
}

}  // namespace mpf
