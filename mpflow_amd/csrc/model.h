// Host orchestration of the EIT path: models::EIT and solver::Solver re-laid out for
// multi-frame batched device solves (model.cpp).
#pragma once

#include <chrono>
#include <memory>

#include "dense.h"
#include "krylov.h"

struct mpf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    // live timing of the SpMV launches with HIP events on `stream`
    bool prof = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    size_t ev_used = 0;
    double prof_ms = 0.0;
    int64_t prof_launches = 0;
    double prof_bytes = 0.0;       // algorithmic bytes of the timed launches
    std::pair<hipEvent_t, hipEvent_t> next_events();
    void collect_events();
    mpf::AutotuneCache tune;
};

namespace mpf {

struct Model {
    mpf_ctx* ctx = nullptr;
    MeshSetup S;
    int frames = 1;
    int quirks = 1;
    // mesh data on the device
    DBuf<int32_t> ids, diag_slot, ctb_off, ctb_elem, elements, w_off, w_node, colmap;
    // voltage sources (S.fixed): coupling / Dirichlet-row contributions, per-slice
    // excitations [frames][C][np][Rp] and electrode currents [frames][C][np][Rp]
    DBuf<int32_t> cpl_off, cpl_node, cpl_elem, drow_off, drow_node, drow_elem;
    DBuf<double> cpl_k, cpl_m, drow_k, drow_m, fslice, cur;
    DBuf<int32_t> node_of_row;   // [np] device row -> caller's node (identity past the mesh nodes)
    DBuf<uint8_t> gw, aggq;
    DBuf<double> ctb_k, ctb_m, F, w_val, colscale;
    DBuf<float> cinv, csc, cy, cg;          // two-level preconditioner (allocated if the mesh has one)
    DBuf<double> cd1;
    DBuf<float> ek, em;
    DBuf<int32_t> jac_order;     // elements sorted by their first device node (locality)
    // per-frame state
    DBuf<float> s, r, s_lo, r_lo, dinv, x, rv, p, q, pring, rhs;
    DBuf<float> rh, yv, sv, zv, tv;   // BiCGSTAB vectors (allocated on first use)
    DBuf<double> x64, part, sc;
    DBuf<float> gamma;
    // job tables
    std::vector<JobDev> hjobs;
    DBuf<JobDev> jobs;
    DBuf<int4> tiles, asm_items;
    DBuf<int2> js;
    int ntiles = 0, nasm = 0, njs = 0;
    // coupled complete electrode model (S.cem_coupled): sources, electrode lengths, Pi,
    // representative pattern columns, T [frames][C][L][pRp], the pattern fields
    // phi [frames][C][np][pRp] and a job table pointing the Jacobian at them
    DBuf<int32_t> bc_off, bc_row, prep;
    DBuf<double> bc_val, elen, pi, cem_T, cem_phi;
    std::vector<JobDev> hjobs_phi;
    DBuf<JobDev> jobs_phi;
    // Gauss-Newton buffers (allocated on first use)
    DBuf<float> Jm, G;
    DBuf<uint16_t> Jb;            // BF16 copy of Jm for the Gram (identity regularisation)
    DBuf<double> A64, U64, y, dy, res, rgn, vE, work_n;
    DBuf<double> ratio;                 // per frame: |dy| / |y| of the last refinement pass
    std::vector<double> ratio_h;
    DBuf<char> gram_work;
    DBuf<int2> pairs;
    DBuf<int> info;

    Model(mpf_ctx* c, const mpf_model_desc& d);
    ~Model();

    size_t vec_elems() const { return (size_t)S.C * S.np * S.Rp; }
    void upload_jobs();
    // one forward solve of every frame; gamma already in `gamma`
    int forward(const mpf_solve_opts& o, float* d_U, double* d_U64, float* d_J, mpf_solve_stats* stats);
    int gn_step(float* d_gamma, const double* d_Umeas, double lambda, int gn_refine,
                const mpf_solve_opts& o, double* d_Ucalc, mpf_solve_stats* stats);
    // phases of a forward solve (shared with batched solves of several models)
    void prepare(const mpf_solve_opts& o);
    void finish(const mpf_solve_opts& o, float* d_U, double* d_U64, float* d_J, mpf_solve_stats* stats);
    bool x64_valid = false;
    // Gauss-Newton step in two phases around the forward solve (batched GN)
    void gn_begin(const float* d_gamma);
    void gn_dense(float* d_gamma, const double* d_Umeas, double lambda, int gn_refine, double* d_Ucalc);
    // regularisation of the GN normal equations (Inverse::calcRegularizationMatrix,
    // inverse.cpp:83-135): identity; noser R = diag(J0^T J0) with J0 the Jacobian at a
    // reference gamma; laplacian R = L^T L.  Non-identity types keep per-element weights
    // w = diag(R)^-1/2; the laplacian also keeps W (R - diag R) W per element (lap_nb,
    // lap_bval) for its E-space CG
    int reg_type = MPF_REG_IDENTITY;
    DBuf<float> reg_w;
    DBuf<int32_t> lap_nb;
    DBuf<double> lap_bval;
    int lap_iterations = 0;           // CG iterations of the last laplacian step
    DBuf<float> sens_mask;            // m x E, applied to the preparation Jacobians
    bool has_mask = false;
    void set_regularization(int type, const float* d_gamma_ref, const mpf_solve_opts& o, mpf_solve_stats* stats);
    // differential solves (Solver::preSolve + solveDifferential, solver.cpp:136-188): U,
    // J and the factored normal matrix at a reference gamma, reused by every later call
    bool diff_ready = false;
    double diff_lam2 = 0.0;
    DBuf<float> diff_gamma, diff_J;
    DBuf<double> diff_U, diff_A;
    void diff_prepare(const float* d_gamma_ref, double lambda, const mpf_solve_opts& o, mpf_solve_stats* stats);
    void set_sensitivity_mask(const float* d_mask);
    void diff_solve(const double* d_Umeas, int gn_refine, float* d_gamma_out);

  private:
    void ensure_bicg();
    void ensure_gn();
    int gram_tile();
    // frames' gamma <- one reference gamma; forward solve with Jacobian into U64 / Jm
    int forward_at(const float* d_gamma_ref, const mpf_solve_opts& o, mpf_solve_stats* stats);
    // (J J^T + lam2 I) y = r for `nb` right-hand sides against one factored matrix A
    // (stride 0 when shared), with gn_refine FP64-residual passes using J
    void normal_solve(const float* J, size_t strideJ, const double* A, size_t strideA, double lam2, int nb,
                      int gn_refine);
    // laplacian: xi (in vE, started at K^T y) <- (K^T K + mu (I + B))^-1 K^T r for `nb`
    // columns, K = J W (strideK), B = W (R - diag R) W, preconditioned by
    // (K^T K + mu I)^-1 through the factored K K^T + mu I (A, strideA)
    void laplacian_cg(const float* K, size_t strideK, const double* A, size_t strideA, double mu, int nb);
    DBuf<double> cg_r, cg_z, cg_p, cg_q, cg_u, cg_t, cg_s, cg_dots;
    KrylovDriver& driver();
    std::unique_ptr<KrylovDriver> drv;
    struct FixDev fixdev() const;
};

int forward_models(std::vector<Model*>& ms, KrylovDriver& drv, const mpf_solve_opts& o, float* const* d_U,
                   double* const* d_U64, float* const* d_J, mpf_solve_stats* stats);

// several models solved together (one set of Krylov launches per iteration)
struct Batch {
    mpf_ctx* ctx = nullptr;
    std::vector<Model*> models;
    std::unique_ptr<KrylovDriver> drv;
    Batch(mpf_ctx* c, std::vector<Model*> ms);
    int solve(const float* const* d_gamma, float* const* d_U, double* const* d_U64, float* const* d_J,
              const mpf_solve_opts& o, mpf_solve_stats* stats);
    int gn_step(float* const* d_gamma, const double* const* d_Umeas, double lambda, int gn_refine,
                const mpf_solve_opts& o, double* const* d_Ucalc, mpf_solve_stats* stats);
};

}  // namespace mpf
