// Krylov driver: PCG or BiCGSTAB over a table of jobs (the frames of one model, or the
// frames of several models of different sizes), with one set of launches per iteration
// for each group of jobs that share the lane count and padded ELL width.
#pragma once

#include <map>
#include <mutex>
#include <vector>

#include "jobs.h"

struct mpf_ctx;

namespace mpf {

// stripe a (job, slice, tile) work list over the 8 XCDs: block b runs on XCD b % 8, and
// XCD x receives a contiguous run of consecutive tiles (shared neighbours in its L2)
void xcd_order(const int4* in, int n, int4* out);

// Launch choices made by timing the candidates once on live data, shared by all models
// of a context: the SpMV kernel and work-list order per (lane count, padded width,
// row-count bucket) (value: variant | order << 4), and the Gram kernel's tile per
// (rows, contraction length, batch) of the Gauss-Newton step (value: 128 or 256).
struct AutotuneCache {
    std::mutex mu;
    std::map<long long, int> spmv;
    std::map<long long, int> gram;
};

class KrylovDriver {
  public:
    // `sources`: the host job tables of the participating models (their pointers and
    // per-model fields are re-read at every run); `ntile_job[i]`, `C[i]`, `wpad[i]` per model
    KrylovDriver(mpf_ctx* ctx, std::vector<std::vector<JobDev>*> sources);
    ~KrylovDriver();
    KrylovDriver(const KrylovDriver&) = delete;
    KrylovDriver& operator=(const KrylovDriver&) = delete;

    // One solve of every job.  rule/tol: stop rule; use_rhs: start from the jobs' rhs
    // buffer (else beta_c F); ref_b: stop relative to ||beta_c F|| (SC_X2).  Returns
    // true when a pair hit the iteration cap or broke down.  Adds step counts to stats.
    bool run(const mpf_solve_opts& o, int rule, double tol, bool use_rhs, bool ref_b,
             mpf_solve_stats* stats);

    // step count of job j, harmonic c after the last run
    int steps(int j, int c) const { return h_st[((size_t)j * kMaxSlices + c) * kStInts + 1]; }
    int njobs() const { return (int)h.size(); }

  private:
    struct Group {
        int lg = 0, wpad = 0, variant = 0;
        int order = 0;                // work-list order: 0 (job, slice, tile), 1 (job, tile, slice)
        bool wide = false;            // a slice vector of >= 2^31 - 2^16 bytes: 32-bit-gather SpMV only
        long long key = 0;
        std::vector<int> jobs;        // indices into h
        std::vector<int4> base_o[2];  // every (job, slice, tile) item in (job, slice, tile) /
                                      // (job, tile, slice) order
        std::vector<int4> all_o[2];   // the same, XCD-ordered
        DBuf<int4> tiles_o[2];
        DBuf<int4> act;               // the active items
        DBuf<int2> js;                // the group's (job, slice) pairs (coarse correction)
        int njs = 0, nchunk = 0;      // coarse row chunks of the group's largest job
        int nall = 0, ncur = 0;
        const int4* cur = nullptr;
        size_t hoff = 0;              // the group's region of the pinned staging buffer
        // (job, 0, tile) items of the harmonic-fused SpMV (variant 2)
        std::vector<int4> fbase, fall;   // in (job, tile) order; XCD-ordered
        DBuf<int4> ftiles, fact;
        int nfall = 0, nfcur = 0;
        const int4* fcur = nullptr;
        size_t fhoff = 0;
    };
    void launch_group_iter(Group& g, int iter, bool bicg, hipStream_t s, hipEvent_t e0, hipEvent_t e1);
    bool coarse = false;              // this run uses the two-level preconditioner
    int init_nchunk = 0;
    void tune(Group& g);
    void drop_graph();

    mpf_ctx* ctx;
    std::vector<std::vector<JobDev>*> src;
    std::vector<JobDev> h;            // combined table (copies; st repointed)
    std::vector<std::pair<int, int>> origin;   // (model, frame) of each job
    DBuf<JobDev> jobs;
    DBuf<int> st, active;
    DBuf<int4> init_tiles;            // every tile of every job (init kernels take any width)
    DBuf<int2> init_js;
    int n_init_tiles = 0, n_init_js = 0;
    std::vector<Group> groups;
    int* h_active = nullptr;          // pinned
    int* h_st = nullptr;              // pinned [njobs][kMaxSlices][kStInts]
    int4* h_tiles = nullptr;          // pinned staging of active lists
    size_t h_tiles_cap = 0;
    // iteration group replayed as a HIP graph (valid while the active lists are unchanged)
    hipGraphExec_t gexec = nullptr;
    hipStream_t cap = nullptr;        // capture stream (the context's may be the null stream)
    bool gbicg = false;
    int gevery = 0;
};

}  // namespace mpf
