/*
 * mpflow_amd — MI355X-native EIT forward/inverse hot path behind a C ABI.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns an int
 * status code and never throws.  Device pointers are caller-owned unless stated.
 * Work is stream-ordered on the context's HIP stream; a call returns when its work is
 * enqueued unless it says it synchronises.
 *
 * Reference seams replaced (paths relative to the mpFlow source tree):
 *   numeric::SparseMatrix::multiply            src/numeric/sparse_matrix.cpp:262-288
 *   FEM::Equation::update / updateMatrix        src/fem/equation.cpp:261-274, 317-345
 *   numeric::ConjugateGradient::solve           src/numeric/conjugate_gradient.cpp:61-161
 *   numeric::BiCGSTAB::solve                    src/numeric/bicgstab.cpp:83-193
 *   preconditioner::diagonal                    src/numeric/preconditioner.cpp:59-76
 *   FEM::Equation::calcJacobian                 src/fem/equation.cpp:282-310
 *   models::EIT::EIT / solve / applyMeasurementPattern   src/models/eit.cpp:29-95, 139-245
 *   solver::Inverse::calcJacobianSquare / calcExcitation / solve
 *                                               src/solver/inverse.cpp:136-238
 *   solver::Solver::preSolve / solveAbsolute    src/solver/solver.cpp:142-225
 *   FEM::Ports::circularBoundary                src/fem/ports.cpp:32-77
 */
#ifndef MPFLOW_AMD_H
#define MPFLOW_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define MPF_OK        0
#define MPF_EINVAL    1   /* bad argument (null pointer, shape mismatch, ...)      */
#define MPF_ENOMEM    2   /* device or host allocation failed                      */
#define MPF_EHIP      3   /* HIP runtime error                                     */
#define MPF_ENOCONV   4   /* an iterative solve hit its iteration cap              */
#define MPF_EFULL     5   /* ELL row needs more slots than allowed ("sparse format full",
                             src/numeric/sparse_matrix.cpp:417-419)                   */

/* stop rules of the Krylov solvers */
#define MPF_STOP_REFERENCE 0  /* sqrt(|r.z|/(b.b)) < FLT_EPSILON on column 0 only
                                 (src/numeric/conjugate_gradient.cpp:145-151)        */
#define MPF_STOP_TOLERANCE 1  /* every column ||r|| <= tol * ||b||                     */

/* precision of the forward solve */
#define MPF_PREC_FP32   0     /* FP32 PCG, FP64 scalars (reference arithmetic)          */
#define MPF_PREC_MIXED  1     /* FP32 PCG + FP64 iterative refinement of the residual   */

/* Krylov method */
#define MPF_KRYLOV_CG       0
#define MPF_KRYLOV_BICGSTAB 1
#define MPF_KRYLOV_BLOCK_CG 2     /* mpf_block_pcg_solve only */

#define MPF_PRECOND_JACOBI    0
#define MPF_PRECOND_TWO_LEVEL 1

#define MPF_REG_IDENTITY 0
#define MPF_REG_NOSER    1
#define MPF_REG_LAPLACIAN 2

/* mpf_model_desc.source_type */
#define MPF_SOURCE_CURRENT 0
#define MPF_SOURCE_VOLTAGE 1

typedef struct mpf_ctx mpf_ctx;
typedef struct mpf_model mpf_model;
typedef struct mpf_batch mpf_batch;
typedef struct mpf_mlp mpf_mlp;

/* Library version string (replaces version::getVersionString, src/version.cpp:23-25). */
const char* mpf_version(void);

/* Message of the last failing call on this host thread ("" if none). */
const char* mpf_last_error(void);

/* Context: device, HIP stream (NULL = create a default-priority stream), scratch
 * arena, autotune cache.  One context per host thread or stream. */
int mpf_ctx_create(int device, void* hip_stream, mpf_ctx** out);
int mpf_ctx_destroy(mpf_ctx* ctx);
int mpf_ctx_synchronize(mpf_ctx* ctx);
/* hipStream_t of the context, as void*. */
void* mpf_ctx_stream(mpf_ctx* ctx);

/* Live timing of the dominant kernel (the ELL SpMV of the Krylov loop) with HIP events
 * recorded on the context's stream around every launch.  Enabling resets the counters.
 * No counterpart in the reference, which has no instrumentation; the timed launch is
 * the A*p product of the loop in src/numeric/conjugate_gradient.cpp:111-158.
 * out[0] = summed launch time in ms, out[1] = launches, out[2] = algorithmic bytes
 * (8 nnz + 8 N R per harmonic per launch). */
int mpf_ctx_profile(mpf_ctx* ctx, int32_t enable);
int mpf_ctx_profile_read(mpf_ctx* ctx, double out[3]);

/* HBM streaming reference for the rooflines (no counterpart in the reference): the best
 * rate in GB/s over `reps` timed launches (after one warm-up) of a grid-stride 128-bit
 * kernel of `blocks` 256-thread blocks with `loads` (1, 2, 4) float4 in flight per thread.
 * pattern 0: y = x (8 bytes moved per float); 1: y = x + 0.5 y (12 bytes per float, the
 * Krylov update kernels' two-reads-one-write pattern).  d_x, d_y: n floats (n a multiple
 * of 4, 16-byte aligned), on the context's device.  Synchronises. */
int mpf_ctx_stream_probe(mpf_ctx* ctx, int32_t pattern, int32_t loads, const float* d_x, float* d_y, int64_t n,
                         int32_t blocks, int32_t reps, double* gbps);
/* The context's autotuner choices so far (no counterpart: the reference's launch
 * geometry is fixed).  which = 0: SpMV, keys (lane-count log2 << 40 | padded ELL width
 * << 32 | row-count bucket) and values (kernel variant | work-list order << 4: variant 0
 * 128-bit gathers, 1 32-bit gathers, 2 harmonic-fused; order 0 (job, harmonic, tile),
 * 1 (job, tile, harmonic)); which = 1: Gauss-Newton Gram, keys (rows << 44 | elements
 * << 16 | frames) and values (tile edge 128 or 256).  Writes up to `cap` pairs and the
 * number of entries to *count. */
int mpf_ctx_autotune_read(mpf_ctx* ctx, int32_t which, int64_t* keys, int32_t* values, int32_t cap,
                          int32_t* count);

/* ---------------------------------------------------------------------------
 * Host helpers (no device work)
 * ------------------------------------------------------------------------- */

/* Ports::circularBoundary (src/fem/ports.cpp:32-77): for each of `count` electrodes of
 * angular `width` starting at offset + 2*pi*l/count, the indices (rows of `boundary`)
 * of the counter-clockwise boundary edges inside the arc.  `out` receives
 * count x max_edges entries (row-major, -1 = unused); `*max_edges_out` the width used.
 * Pass out=NULL to query the width first. */
int mpf_ports_circular(int32_t n_nodes, const double* nodes, int32_t n_elements,
                       int32_t nodes_per_element, const int32_t* elements,
                       int32_t n_boundary, int32_t nodes_per_edge, const int32_t* boundary,
                       int32_t count, double width, double offset, int32_t clockwise,
                       int32_t* out, int32_t out_capacity, int32_t* max_edges_out);

/* basis::Linear / basis::Quadratic integralA and integralB over one straight triangle
 * (src/fem/basis_linear.cpp:72-91, src/fem/basis_quadratic.cpp:75-97): K[i*k+j] =
 * int grad(phi_i).grad(phi_j), M[i*k+j] = int phi_i phi_j for k = 3 (P1) or 6 (P2,
 * nodes [v0 v1 v2 m01 m12 m20]).  xy holds k (x, y) pairs.  Host only. */
int mpf_element_integrals(int32_t nodes_per_element, const double* xy, double* K, double* M);

/* ---------------------------------------------------------------------------
 * EIT model: mesh + ports + sources + 2.5D set-up (models::EIT::EIT,
 * src/models/eit.cpp:29-95; FEM::Equation::Equation, src/fem/equation.cpp:29-68).
 * ------------------------------------------------------------------------- */
typedef struct mpf_model_desc {
    /* mesh (host arrays, row-major) */
    int32_t n_nodes;              /* N                                             */
    int32_t n_elements;           /* E                                             */
    int32_t nodes_per_element;    /* 3 = P1, 6 = P2 [v0 v1 v2 m01 m12 m20]          */
    const double* nodes;          /* N x 2                                          */
    const int32_t* elements;      /* E x nodes_per_element                          */
    int32_t n_boundary;           /* boundary edge count                            */
    int32_t nodes_per_edge;       /* 2 = P1 [a b], 3 = P2 [a mid b]; counter-clockwise */
    const int32_t* boundary;      /* n_boundary x nodes_per_edge                    */
    /* ports: rows of `boundary` per port, -1 = unused (Ports::edges, ports.h:26-46) */
    int32_t n_ports;
    int32_t port_max_edges;
    const int32_t* port_edges;    /* n_ports x port_max_edges                       */
    /* sources (Sources::Sources, src/fem/sources.cpp:27-75), Open/current type     */
    int32_t n_drive;              /* D                                              */
    int32_t n_meas;               /* M                                              */
    const double* drive_pattern;  /* n_ports x D, row-major                         */
    const double* meas_pattern;   /* n_ports x M, row-major                         */
    const double* drive_current;  /* D values (Sources::values)                     */
    /* material and 2.5D (EIT::fromConfig, src/models/eit.cpp:119-126)              */
    double reference_value;
    int32_t components;           /* harmonics C >= 1 (at most 8)                   */
    double height;
    double port_height;
    int32_t logarithmic;          /* Equation<..., logarithmic>                     */
    int32_t quirks;               /* 1: reproduce the reference's P2 boundary integral */
    int32_t jacobian_exact;       /* 0: reference Jacobian (K_e only, factor height);
                                     1: exact 2.5D derivative (K_e + alpha_c M_e, 1/beta_c) */
    int32_t ell_max_width;        /* 0 -> 32 (reference cap); at most 64            */
    int32_t frames;               /* frames solved together on this mesh (>= 1)     */
    int32_t keep_node_order;      /* 1: solve in the caller's node numbering; 0: renumber
                                     internally for gather locality (results are the same
                                     up to FP32 rounding; outputs never expose node order) */
    double contact_impedance;     /* > 0: complete electrode model with this contact
                                     impedance on every port (no counterpart in the reference,
                                     whose ports use the gap model); electrode potentials
                                     become extra unknowns, current enters through them and
                                     the voltages are the electrode potentials; each harmonic
                                     carries its own CEM system.  0: gap model (reference) */
    int32_t source_type;          /* MPF_SOURCE_CURRENT (Open, the default) or MPF_SOURCE_VOLTAGE
                                     (Fixed, Sources::Type, sources.h; eit.cpp:89-91,159-194):
                                     the rows of every node with a clipped excitation
                                     |B| >= 1e-9 become Dirichlet rows holding (B_clip Pi),
                                     no beta scaling, no dcFree harmonic, and the result is
                                     height * Mpat^T B_clip^T (A_c Phi_c) summed over the
                                     harmonics (the electrode currents); Jacobian x height
                                     (not -height).  Gap model only; jacobian_exact = 0 */
    int32_t cem_coupled;          /* with contact_impedance > 0.  0: harmonic-wise CEM (above).
                                     1: each electrode holds ONE potential V_l for all
                                     harmonics (the patch e_l x height port_height): mode c
                                     solves (A_c + Zb / z') u_c = (h beta_c / z') Bc V with
                                     z' = z height / port_height (the patch average of the
                                     mode's contact weight), and the electrodes close the system
                                     sum_c (n_c / z') [(h beta_c)^2 |e| V - h beta_c Bc^T u_c] = I
                                     (n_c = height, then height / 2); solved by one Krylov solve
                                     per harmonic with the L unit electrode sources Bc and a
                                     Schur complement on V.  Voltages: Mpat^T V.  Needs
                                     jacobian_exact = 1 (weights n_c); mpf_eit_measure and
                                     mpf_jacobian (per-harmonic seams) are refused. */
    int32_t complex_gamma;        /* 1: complex conductivity (models::EIT over
                                     Equation<thrust::complex<float>, ...>, eit.cpp:289-304;
                                     ConjugateGradient<thrust::complex<float>>,
                                     conjugate_gradient.cpp:231-232).  gamma is [frames][E]
                                     complex (re, im float pairs); mpf_eit_solve writes U, U64
                                     and J as (re, im) pairs of float / double; mpf_gn_step
                                     takes and returns them the same way and solves the complex
                                     symmetric J^T J of Inverse::solve (plain transposes,
                                     inverse.cpp:148-152, 196-204).  CG with the diagonal
                                     preconditioner, FP32 (opts.precision MPF_PREC_FP32),
                                     current sources, gap model; the other model entry points
                                     and batches refuse such a model. */
} mpf_model_desc;

typedef struct mpf_solve_opts {
    int32_t stop_rule;            /* MPF_STOP_*                                     */
    double tol;                   /* for MPF_STOP_TOLERANCE                         */
    int32_t max_iter;             /* 0 -> 10 * N (conjugate_gradient.cpp:78)         */
    int32_t precision;            /* MPF_PREC_*                                     */
    int32_t refine_steps;         /* FP64 refinement passes (MPF_PREC_MIXED)        */
    double refine_tol;            /* inner tolerance of each refinement solve       */
    int32_t krylov;               /* MPF_KRYLOV_*                                   */
    int32_t check_every;          /* iterations between host convergence polls (0 -> 16) */
    int32_t warm_start;           /* MPF_PREC_MIXED only: start from the model's previous FP64
                                     fields (a correction solve of the FP64 residual, stopped at
                                     tol * ||b||) instead of from zero; ignored on the first solve */
    int32_t bicgstab_ignore_dc;   /* BiCGSTAB only.  0: harmonic 0 uses the dcFree operator
                                     A + 1 1^T like CG (FP32 BiCGSTAB on the singular matrix
                                     stagnates on large meshes); 1: ignore dcFree as the
                                     reference does (bicgstab.cpp:87).  Voltages and Jacobian
                                     are the same either way. */
    int32_t graphs;               /* 1: capture the Krylov iterations between two convergence
                                     polls as a HIP graph and replay it (fewer launch
                                     overheads for small meshes); not with profiling */
    int32_t preconditioner;       /* MPF_PRECOND_*.  JACOBI: the reference's diagonal
                                     preconditioner (preconditioner.cpp:59-76).  TWO_LEVEL
                                     (CG on model solves; no counterpart in the reference):
                                     Jacobi plus an additive coarse correction over compact
                                     aggregates (the four 64-node bisection leaves of each
                                     256-node tile),
                                     M^-1 = D^-1 + W Ec^-1 W^T with Ec the aggregates'
                                     sigma = 1 coarse matrix scaled per aggregate to the
                                     frame's conductivity; several times fewer iterations on
                                     large meshes.  Models above 512 tiles (131072 unknowns)
                                     fall back to Jacobi. */
    int32_t complex_hermitian;    /* complex-conductivity models only.  0 (default): conjugate
                                     orthogonal CG (dot products r^T z, p^T q without
                                     conjugation), the Krylov method for complex symmetric
                                     systems.  1: the reference's iteration, whose dot products
                                     conjugate their second operand (vectorDotProductKernel,
                                     matrix_kernel.cu:405-413); it stalls or diverges when the
                                     conductivity's imaginary part is not small. */
} mpf_solve_opts;

typedef struct mpf_solve_stats {
    int64_t steps;                /* sum over frames and harmonics of Krylov steps    */
    int64_t max_steps;            /* largest per (frame, harmonic) step count        */
    int32_t not_converged;        /* (frame, harmonic) pairs that hit the cap          */
    double refine_residual;       /* max relative FP64 residual after refinement      */
    double seconds;               /* wall time of the call (synchronised)             */
} mpf_solve_stats;

void mpf_solve_opts_default(mpf_solve_opts* o);

int mpf_model_create(mpf_ctx* ctx, const mpf_model_desc* desc, mpf_model** out);
int mpf_model_destroy(mpf_model* m);

/* dims[0]=N, [1]=E, [2]=padded rows np, [3]=ELL width, [4]=Jacobian rows
 * m = ceil16(M)*ceil16(D), [5]=ceil16(M), [6]=ceil16(D), [7]=R=D+M, [8]=RHS lanes,
 * [9]=frames, [10]=C, [11]=nnz, [12]=unique RHS columns solved (columns of Pi that are
 * scalar multiples of an earlier column share one solve), [13]=padded ELL width */
int mpf_model_dims(const mpf_model* m, int64_t dims[16]);

/* models::EIT::solve (src/models/eit.cpp:139-213) for every frame.
 *   d_gamma : frames x E   (device, float)   material parameter gamma
 *   d_U     : frames x m   (device, float)   voltages, column-major M x D with
 *             leading dimension ceil16(M), padded rows zero (EIT::result layout)
 *   d_U64   : frames x m   (device, double)  same in FP64 (nullable)
 *   d_J     : frames x m x E (device, float) Jacobian, column-major (nullable)
 * Synchronises.  Returns MPF_ENOCONV if any (frame, harmonic) hit the cap. */
int mpf_eit_solve(mpf_model* m, const float* d_gamma, float* d_U, double* d_U64, float* d_J,
                  const mpf_solve_opts* opts, mpf_solve_stats* stats);

/* One Solver::solveAbsolute per frame (src/solver/solver.cpp:195-225):
 * forward solve + Jacobian at gamma, r = U_meas - U(gamma), then
 * gamma += (J^T J + (ref*lambda)^2 I)^-1 J^T r, solved in the m x m data space
 * (J J^T on MFMA, BF16 in / FP32 accumulate, FP64 Cholesky, at most `gn_refine`
 * FP64-residual refinement passes, fewer once every frame's predicted remaining error
 * is below 1e-10 of its solution).
 *   d_gamma  : frames x E (device float, in/out)
 *   d_Umeas  : frames x m (device double)
 *   d_Ucalc  : frames x m (device double, out, nullable) voltages at the input gamma */
int mpf_gn_step(mpf_model* m, float* d_gamma, const double* d_Umeas, double lambda,
                int32_t gn_refine, const mpf_solve_opts* opts, double* d_Ucalc,
                mpf_solve_stats* stats);

/* Equation::update of one frame (src/fem/equation.cpp:261-274, updateMatrixKernel and
 * updateSystemMatrixKernel in src/fem/equation_kernel.cu:60-157): the system matrix
 * A = S(gamma) + alpha R(gamma), S = sum_e sigma(gamma_e) K_e, R = sum_e sigma(gamma_e) M_e,
 * summed in FP64 and rounded once, written as a SparseMatrix (sparse_matrix.cpp:48-114):
 * rows x 32 slots, entry (row, slot) at row*32 + slot, empty slots with column id
 * 0xFFFFFFFF and value 0, slots in the reference's first-insertion order
 * (equation.cpp:131-144), rows and column ids in the caller's node numbering.
 *   d_gamma         : E floats (device)
 *   alpha           : the harmonic's (2 pi c / height)^2 (eit.cpp:154-155), 0 for 2D
 *   d_column_ids    : rows*32 (device, nullable: the pattern does not depend on gamma)
 *   d_values        : rows*32 (device); rows = dims[0] (N, plus the electrode potentials
 *                     of a complete-electrode model)
 * MPF_EINVAL if the model was built with rows wider than 32 entries.  Synchronises. */
int mpf_fem_assemble(mpf_model* m, const float* d_gamma, double alpha, uint32_t* d_column_ids,
                     float* d_values);

/* EIT::applyMeasurementPattern (src/models/eit.cpp:220-245):
 *   U (+)= W Phi[:, 0:D],  W = measurementPattern^T excitationMatrix^T
 *   d_phi : fields, column-major rows x (D + M) with leading dimension ldphi >= rows
 *   d_U   : ceil16(M) x D column-major (ld ceil16(M)); rows >= M get 0 (or + 0)
 *   additive : 0 overwrite, 1 accumulate (the reference's harmonic > 0 calls).  Synchronises. */
int mpf_eit_measure(mpf_model* m, const float* d_phi, int32_t ldphi, int32_t additive, float* d_U);

/* Equation::calcJacobian (src/fem/equation.cpp:282-310, calcJacobianKernel in
 * src/fem/equation_kernel.cu:197-245) for one harmonic's fields:
 *   J[d*ceil16(M) + m, e] (+)= sigma'(gamma_e) sum_ij Phi[n_i, d] K_e[i][j] Phi[n_j, D + m]
 * with K_e the element stiffness integrals; rows with d >= D or m >= M are 0.  J is
 * column-major ceil16(M) ceil16(D) x E.  The reference's scaling by -height
 * (eit.cpp:202-205) is the caller's.  Synchronises. */
int mpf_jacobian(mpf_model* m, const float* d_phi, int32_t ldphi, const float* d_gamma, int32_t additive,
                 float* d_J);

/* Inverse::setRegularizationParameter's type (src/solver/inverse.cpp:83-135,
 * src/solver/solver.cpp:116-132) for the model's later GN steps and differential
 * preparations; the normal matrix is J^T J + (ref lambda)^2 R.
 *   MPF_REG_IDENTITY : R = I (the default)
 *   MPF_REG_NOSER    : R = diag(J0^T J0), J0 the Jacobian at d_gamma_ref (E floats;
 *                      null: the reference distribution, 1, or 0 when logarithmic),
 *                      computed here with one forward solve (opts) and kept, as the
 *                      reference keeps the matrix it computed when the type was set.
 *                      Entries below 1e-12 of the largest are raised to that floor.
 *   MPF_REG_LAPLACIAN: R = L^T L, L[edge][element] = +-|edge| over the elements' vertex
 *                      edges (inverse.cpp:96-117); no forward solve.  The step is solved
 *                      in E-space by FP64 CG on (J^T J + (ref lambda)^2 R) preconditioned by
 *                      the data-space solve with R's diagonal, to a relative residual of
 *                      1e-10 (at most 2000 iterations).
 * Synchronises. */
int mpf_gn_set_regularization(mpf_model* m, int32_t type, const float* d_gamma_ref,
                              const mpf_solve_opts* opts, mpf_solve_stats* stats);

/* Solver::preSolve + Solver::solveDifferential (src/solver/solver.cpp:136-188).
 * mpf_diff_prepare: forward solve, Jacobian and the factored data-space normal matrix at
 * d_gamma_ref (E floats), once.  mpf_diff_solve: for every frame's measurement
 *   d_Umeas : frames x m (FP64, EIT::result layout)
 *   d_gamma : frames x E (out) = gamma_ref + (J^T J + (ref lambda)^2 R)^-1 J^T (U_meas - U(gamma_ref))
 * with at most gn_refine FP64-residual passes (fewer once converged, as in mpf_gn_step):
 * no forward solve; per frame and pass one triangular solve pair and one pass over J
 * (J J^T y in one sweep, J rows <= 1024).  Both synchronise. */
int mpf_diff_prepare(mpf_model* m, const float* d_gamma_ref, double lambda, const mpf_solve_opts* opts,
                     mpf_solve_stats* stats);

/* The configuration's noSensitivity mask (Solver::fromConfig, src/solver/solver.cpp:107-113:
 * jacobian->elementwiseMultiply(jacobian, mask) after preSolve, before the regularisation
 * is set).  d_mask: m x E FP32, column-major like the Jacobian (ld m), copied; null clears
 * it.  The Jacobians of later mpf_gn_set_regularization (noser) and mpf_diff_prepare calls
 * are multiplied by it elementwise; mpf_gn_step's Jacobian is not, as the reference's
 * solveAbsolute recomputes it unmasked.  Invalidates a previous mpf_diff_prepare. */
int mpf_set_sensitivity_mask(mpf_model* m, const float* d_mask);
int mpf_diff_solve(mpf_model* m, const double* d_Umeas, int32_t gn_refine, float* d_gamma);

/* ---------------------------------------------------------------------------
 * Batches: several models (meshes of any size, any electrode count and harmonic count)
 * solved together.  No counterpart in the reference, whose models solve one at a time;
 * each Krylov iteration is one set of launches per group of models that share the lane
 * count (16/32/64 right-hand sides) and padded ELL width.  Models keep their own
 * buffers; a model belongs to at most one batch; destroy the batch before its models.
 * ------------------------------------------------------------------------- */
int mpf_batch_create(mpf_ctx* ctx, mpf_model* const* models, int32_t n_models, mpf_batch** out);
int mpf_batch_destroy(mpf_batch* b);
/* mpf_eit_solve of every model; the arrays hold one device pointer per model (entries
 * of d_U64 / d_J and the arrays themselves may be null). */
int mpf_batch_solve(mpf_batch* b, const float* const* d_gamma, float* const* d_U, double* const* d_U64,
                    float* const* d_J, const mpf_solve_opts* opts, mpf_solve_stats* stats);
/* mpf_gn_step of every model, forward solves batched, dense steps per model. */
int mpf_batch_gn_step(mpf_batch* b, float* const* d_gamma, const double* const* d_Umeas, double lambda,
                      int32_t gn_refine, const mpf_solve_opts* opts, double* const* d_Ucalc,
                      mpf_solve_stats* stats);

/* ---------------------------------------------------------------------------
 * Low-level kernels on caller-owned device buffers
 * ------------------------------------------------------------------------- */

/* ConjugateGradient::solve / BiCGSTAB::solve with a SparseMatrix operator and the
 * diagonal preconditioner (src/numeric/conjugate_gradient.cpp:61-161,
 * src/numeric/bicgstab.cpp:83-193, src/numeric/preconditioner.cpp:59-76), all R <= 64
 * right-hand sides at once, starting from the caller's X as the reference does
 * (r = B - A X, conjugate_gradient.cpp:80-82, bicgstab.cpp:105-108).
 *   d_column_ids, d_values : SparseMatrix device arrays (sparse_matrix.cpp:48-114): 32
 *                            slots per row, entry (row, slot) at row*32 + slot, empty
 *                            slots with column id 0xFFFFFFFF; `density` used slots.  A
 *                            column id that is neither 0xFFFFFFFF nor below n is refused
 *                            (MPF_EINVAL); a row without a diagonal entry gets a Jacobi
 *                            inverse of 0 (preconditioner_kernel.cu:63-83)
 *   d_B                    : Matrix column-major n x R with leading dimension ldb
 *   d_X                    : in: start X0, out: solution; column-major, leading dimension ldx
 *   dc_free                : solve (A + 1 1^T) X = B (CG; BiCGSTAB as opts.bicgstab_ignore_dc).
 *                            The start residual is B - A X0 + 1 (1^T X0), as the reference
 *                            forms it (conjugate_gradient.cpp:84-88); X0 = 0 is unaffected
 * Both stop rules compare with B's column norms.  opts: stop rule, tol, max_iter,
 * precision (MPF_PREC_MIXED adds FP64 refinement), check_every, graphs; `krylov` is set
 * by the entry point.  Synchronises. */
int mpf_pcg_solve(mpf_ctx* ctx, int32_t n, int32_t density, const uint32_t* d_column_ids,
                  const float* d_values, int32_t R, const float* d_B, int32_t ldb, float* d_X,
                  int32_t ldx, int32_t dc_free, const mpf_solve_opts* opts, mpf_solve_stats* stats);
int mpf_bicgstab_solve(mpf_ctx* ctx, int32_t n, int32_t density, const uint32_t* d_column_ids,
                       const float* d_values, int32_t R, const float* d_B, int32_t ldb, float* d_X,
                       int32_t ldx, int32_t dc_free, const mpf_solve_opts* opts, mpf_solve_stats* stats);

/* mpf_pcg_solve's problem (same arguments and meaning) solved by block PCG: the R <= 64
 * columns share one Krylov space (O'Leary's block CG with the diagonal preconditioner;
 * the block of search directions is orthonormalised every iteration and directions that
 * have become linearly dependent are dropped, so right-hand sides that are linear
 * combinations of each other -- drive patterns that sum to zero -- are handled).  No
 * counterpart in the reference, whose ConjugateGradient::solve
 * (src/numeric/conjugate_gradient.cpp:61-161) iterates the columns independently; a
 * caller replaces that call with this one when its columns share the matrix.
 * stats->steps counts block iterations (one SpMV of all columns each); not_converged is 1
 * when the iteration cap was hit or no independent direction was left (MPF_ENOCONV).
 * opts.check_every: iterations between host convergence polls; opts.graphs is ignored. */
int mpf_block_pcg_solve(mpf_ctx* ctx, int32_t n, int32_t density, const uint32_t* d_column_ids,
                        const float* d_values, int32_t R, const float* d_B, int32_t ldb, float* d_X,
                        int32_t ldx, int32_t dc_free, const mpf_solve_opts* opts, mpf_solve_stats* stats);

/* SparseMatrix::multiply (src/numeric/sparse_matrix.cpp:262-288; the kernel wrapper
 * sparseMatrixKernel::multiply, include/mpflow/numeric/sparse_matrix_kernel.h:57-60,
 * src/numeric/sparse_matrix_kernel.cu:157-196) with the same arguments minus the launch
 * geometry and stream:
 *   result[row + col * result_rows] = sum over slot < density of
 *       values[row*32 + slot] * matrix[column_ids[row*32 + slot] + col * matrix_rows]
 * skipping slots whose column id is 0xFFFFFFFF.  d_values / d_column_ids: the
 * SparseMatrix arrays (result_rows x 32); d_matrix (matrix_rows x columns) and d_result
 * (result_rows x columns) column-major with the row count as leading dimension
 * (Matrix::dataRows).  A column id that is neither 0xFFFFFFFF nor below matrix_rows is
 * refused (MPF_EINVAL).  Runs on the Krylov SpMV's gather core (X copied once into a
 * row-interleaved layout).  Synchronises. */
int mpf_sparse_multiply(mpf_ctx* ctx, const float* d_values, const uint32_t* d_column_ids, const float* d_matrix,
                        uint32_t result_rows, uint32_t matrix_rows, uint32_t columns, uint32_t density,
                        float* d_result);

/* Multi-harmonic ELL SpMV on the product's own layout, the S/R form of
 * SparseMatrix::multiply (src/numeric/sparse_matrix_kernel.cu:157-196) fused with
 * Equation::update's A_c = S + alpha_c R (src/fem/equation_kernel.cu:137-157):
 *   Y[c][row][j] = sum_slot (S[slot][row] + alpha[c] R[slot][row]) * X[c][ids[slot][row]][j]
 * ELL is column-major (slot-major): entry (row, slot) at slot*np + row, width a multiple
 * of 8 (at most 64); padding slots hold id = row and value 0; an id outside [0, np) is
 * refused (MPF_EINVAL).  alpha is a HOST array of C floats; every other pointer is device
 * memory.  X/Y: C x np x ld, all ld lanes are computed, j < R <= ld, ld in {16,32,64}.
 * Synchronises. */
int mpf_ell_spmv(mpf_ctx* ctx, int32_t np, int32_t width, const int32_t* d_ids,
                 const float* d_S, const float* d_R, int32_t C, const float* alpha,
                 const float* d_X, float* d_Y, int32_t R, int32_t ld);

/* Symmetric Gram matrix on MFMA (BF16 inputs, FP32 accumulation):
 *   trans == 0: G (rows x rows) = A A^T, A rows x k column-major (lda >= rows)
 *   trans == 1: G (k x k)       = A^T A (J^T J of Inverse::calcJacobianSquare,
 *                                  src/solver/inverse.cpp:136-156)
 * G is written full (both triangles), column-major, ldg = its order. */
int mpf_gram_bf16(mpf_ctx* ctx, int32_t trans, int32_t rows, int32_t k, const float* d_A,
                  int32_t lda, float* d_G);

/* y = A^T x (FP64 accumulation): Inverse::calcExcitation's J^T (U_meas - U_calc)
 * (src/solver/inverse.cpp:163-205).  A rows x cols column-major (lda >= rows). */
int mpf_gemv_t(mpf_ctx* ctx, int32_t rows, int32_t cols, const float* d_A, int32_t lda,
               const double* d_x, double* d_y);

/* Solve (G + shift I) y = b for SPD G (n x n, column-major FP32) by FP64 Cholesky on
 * the device; b, y FP64. */
int mpf_spd_solve(mpf_ctx* ctx, int32_t n, const float* d_G, double shift, const double* d_b,
                  double* d_y);

/* Inverse::solve's Tikhonov system (src/solver/inverse.cpp:60-76, 163-238) for a
 * caller's Jacobian and residual:
 *   delta = (J^T J + shift I)^-1 J^T r = J^T (J J^T + shift I)^-1 r
 *   d_J   : rows x cols column-major FP32 (ldj >= rows), e.g. the EIT Jacobian m x E
 *   shift : (referenceValue * regularizationFactor)^2 > 0
 *   d_r   : rows (FP64),  d_delta : cols (FP64, out)
 * J J^T on MFMA (BF16 in, FP32 accumulate), FP64 Cholesky, then gn_refine passes of FP64
 * residual refinement with the FP32 J.  Allocates its work space; synchronises. */
int mpf_normal_solve(mpf_ctx* ctx, int32_t rows, int32_t cols, const float* d_J, int32_t ldj, double shift,
                     const double* d_r, int32_t gn_refine, double* d_delta);

/* mpf_normal_solve without any allocation.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, kernels and result as
 * mpf_normal_solve (the Gram kernel sums its contraction in at most 4 slabs, so results
 * agree to rounding), computed in a fixed 24 MiB work space that each GPU holds from the
 * library's first context on it until the process ends.  Limits: rows <= 1024,
 * cols <= 65536, rows and ldj multiples of 4; larger problems are refused (MPF_EINVAL).
 * Every call on one GPU shares that GPU's work space: calls that overlap on one GPU (two
 * streams, two host threads) race.  Synchronises. */
int mpf_normal_solve_noalloc(mpf_ctx* ctx, int32_t rows, int32_t cols, const float* d_J, int32_t ldj,
                             double shift, const double* d_r, int32_t gn_refine, double* d_delta);

/* ---------------------------------------------------------------------------
 * Learned start: a multilayer perceptron (dense layers, GELU between them, linear
 * output) trained with MSE and Adam.  No counterpart in the reference.  GEMMs run on
 * MFMA with BF16 operands and FP32 accumulation; parameters, gradients and Adam moments
 * are FP32.  Matrices are row-major: X batch x dims[0], Y batch x dims[n_layers],
 * layer l's weights dims[l] x dims[l+1].
 * ------------------------------------------------------------------------- */
int mpf_mlp_create(mpf_ctx* ctx, int32_t n_layers, const int32_t* dims, uint64_t seed, mpf_mlp** out);
int mpf_mlp_destroy(mpf_mlp* m);
/* Y (nullable) = net(X); keeps the activations for mpf_mlp_bwd (X must stay valid). */
int mpf_mlp_fwd(mpf_mlp* m, int32_t batch, const float* d_X, float* d_Y);
/* gradients of mean((Y - T)^2) for the last forward; *loss (nullable) receives the MSE */
int mpf_mlp_bwd(mpf_mlp* m, const float* d_T, double* loss);
/* one Adam step (bias-corrected) on the current gradients */
int mpf_mlp_adam(mpf_mlp* m, float lr, float beta1, float beta2, float eps);
/* number of parameters; flat order W_0, b_0, W_1, b_1, ... */
int64_t mpf_mlp_param_count(const mpf_mlp* m);
/* which 0: parameters (get or set), 1: gradients (get); host buffer of param_count floats */
int mpf_mlp_params(mpf_mlp* m, int32_t which, int32_t set, float* host);

#ifdef __cplusplus
}
#endif
#endif /* MPFLOW_AMD_H */
