// Device-side job descriptors shared by the sparse and dense kernels.
//
// A "job" is one frame on one mesh: its own sigma-weighted S/R values, harmonics
// ("slices"), Krylov vectors and scalars.  Jobs of one launch may differ in size;
// the kernels walk a flat table of (job, slice, row tile) work items.
//
// Vector layout (all Krylov vectors): [slice][row][lane], lane stride Rp in
// {16, 32, 64} >= R = D + M.  A wavefront covers 64/Rp rows x Rp right-hand sides,
// so every gathered neighbour row is one contiguous Rp*4-byte load.
#pragma once

#include "mpf_internal.h"

namespace mpf {

// scalar fields per (slice, lane), FP64
enum ScField { SC_RHO = 0, SC_PQ, SC_PSUM, SC_BB, SC_RR, SC_ALPHA, SC_BETA, SC_BZERO,
               SC_RHO0, SC_RS, SC_V, SC_W, SC_OMEGA, SC_X0, SC_X1, SC_X2, SC_RHO1, SC_MB,
               SC_PA0, SC_PA1, SC_PA2, SC_PA3, SC_PQ0, SC_PQ1, SC_PQ2, SC_PQ3, SC_NFIELDS };
// PCG keeps the last kRing search directions: p of iteration k in buffer k % kRing, and
// alpha / p.q of iteration k in SC_PA0 / SC_PQ0 + k % kRing; x moves every kRing iterations
constexpr int kRing = 4;
// per-slice integer state: [0] active, [1] steps, [2] hit cap or breakdown,
// [3] BiCGSTAB: the next iteration is the first; CG: [0] as the update kernel of the
// current iteration saw it (the stop rule clears [0] during the pupdate launch)
constexpr int kStInts = 4;
// partial-sum channels per tile
constexpr int kParts = 10;

struct JobDev {
    int n, np, width, C, R, Rp, lg, ntiles, dc, maxit, rule, col0;   // R = unique RHS solved
    int dc_rank1;                // 1: the dc slice's operator is A + 1 1^T (CG dcFree);
                                 // 0: A itself (BiCGSTAB, constants projected out)
    int nnz;                     // true nonzeros of the pattern (bandwidth accounting)
    double tol;
    const int32_t* ids;          // [width][np]
    float* s;  float* r;         // FP32 values (hi)        [width][np]
    float* s_lo; float* r_lo;    // FP32 low parts (FP64 = hi + lo)
    float* dinv;                 // [C][np]
    float alpha[kMaxSlices];     // alpha_c of A_c = S + alpha_c R
    float xscale[kMaxSlices];    // beta_c excitation scale
    double alpha64[kMaxSlices];  // the same two constants in FP64 for the refinement residual
    double xscale64[kMaxSlices];
    const double* F;             // [np][Rp] excitation B*Pi, FP64 (slice s at F + s * fstride)
    size_t fstride;              // 0: one excitation for every slice (Open sources)
    double* cur;                 // [C][np][Rp] voltage sources: A_c Phi on the Dirichlet rows
    float* rhs;                  // [C][np][Rp] FP32 right-hand side buffer (always valid)
    int use_rhs;                 // 1: PCG starts from rhs; 0: from beta_c * F
    int ref_b;                   // 1: stop relative to ||beta_c F|| (SC_X2) instead of ||rhs||
    float* x; float* rv; float* p; float* q;   // [C][np][Rp] (q = nu in BiCGSTAB)
    float* pr[kRing - 1];        // [C][np][Rp] PCG: p of iterations k % kRing = 1.. (0: in p)
    // BiCGSTAB only (null otherwise): r-hat, y = K^-1 p, s, z = K^-1 s, t   [C][np][Rp]
    float* rh; float* yv; float* sv; float* zv; float* tv;
    double* x64;                 // [C][np][Rp] FP64 refined solution
    double* part;                // [kParts][C][ntiles][Rp]
    double* sc;                  // [SC_NFIELDS][C][Rp]
    int* st;                     // [C][kStInts]
    int* active;                 // unconverged (job, harmonic) pairs of the launch
    // assembly
    const int32_t* ctb_off; const int32_t* ctb_elem;
    const double* ctb_k; const double* ctb_m;
    const float* gamma;          // [E]
    double ref; int logarithmic; int E;
    const int32_t* diag_slot;    // [np]
    const uint8_t* gw;           // [np / 16] entry count of the widest row of each 16-row group
    // two-level preconditioner (CG only; null: Jacobi alone).  Aggregate 4t + q is the set
    // of tile t's real rows with aggq = q (its four 64-node bisection leaves):
    // M^-1 r = D^-1 r + W Einv W^T r with W the aggregate indicator,
    // Einv = Sc^-1/2 cinv_s Sc^-1/2 (cinv_s: inverse of W^T A_s(1) W, sigma = 1, plus n n^T
    // on the dcFree slice; Sc: per-aggregate conductivity scale of this frame).  W^T r of
    // aggregate q of a tile is partial channel agg_channel(q).
    const uint8_t* aggq;         // [np] aggregate of each row within its tile
    const float* cinv;           // [C][4 ntiles][cld], rows zero-padded to cld = ceil4(4 ntiles)
    int cld;
    float* csc;                  // [4 ntiles] Sc^-1/2 (written after each assembly)
    float* cy;                   // [C][4 ntiles][Rp] coarse correction Einv W^T r
    float* cg;                   // [C][4 ntiles][Rp] Sc^-1/2 W^T r of the current iteration
};

// PCG search direction of iteration k (k4 = k % kRing)
__host__ __device__ inline float* pbuf(const JobDev& J, int k4) { return k4 == 0 ? J.p : J.pr[k4 - 1]; }

__host__ __device__ inline size_t vec_index(const JobDev& J, int s, int row, int lane) {
    return ((size_t)s * J.np + row) * J.Rp + lane;
}

// Pointers read from a job descriptor are generic; re-tag them as global so the
// compiler emits global_load/store (no flat path, no LDS counter waits).
#if defined(__HIP_DEVICE_COMPILE__)
#define MPF_GLOBAL __attribute__((address_space(1)))
#else
#define MPF_GLOBAL
#endif
template <class T>
__device__ __forceinline__ MPF_GLOBAL T* gp(T* p) { return (MPF_GLOBAL T*)p; }

// host launch wrappers (sparse.hip)
void launch_assemble(const JobDev* d_jobs, const int4* d_items, int nitems, hipStream_t st);
void launch_dinv(const JobDev* d_jobs, const int4* d_tiles, int ntiles, hipStream_t st);
// nchunk > 0: jobs with a coarse level start from the two-level preconditioned residual
void launch_pcg_init(const JobDev* d_jobs, const int4* d_tiles, int ntiles,
                     const int2* d_js, int njs, int* d_active, hipStream_t st, int nchunk = 0);
// BiCGSTAB (bicgstab.hip): init over all pairs, then one iteration per call (five
// launches: p-update, SpMV, s-update, SpMV, x/r-update); ev0/ev1 bracket the first SpMV
void launch_bicg_init(const JobDev* d_jobs, const int4* d_tiles, int ntiles,
                      const int2* d_js, int njs, int* d_active, int lg, hipStream_t st);
void launch_bicg_iter(const JobDev* d_jobs, const int4* d_tiles, int ntiles, int* d_active, int iter,
                      int lg, int wpad, hipStream_t st, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
// ev0/ev1 (nullable) are recorded right before and after the SpMV launch
// one PCG iteration (SpMV, update, pupdate); `iter` (iterations since the solve's init)
// selects the rho slot and the direction-ring entry.  `variant` picks the SpMV kernel
// (0: 128-bit gathers, width-binned, for padded widths <= 24; 1: 32-bit gathers; 2: as 0
// with all harmonics of a (job, tile) in one block, over the (job, -, tile) list d_ftiles).
constexpr int kSpmvVariants = 3;
// d_js lists the group's (job, slice) pairs and `nchunk` the largest job's 32-tile
// chunks: after the SpMV a small launch sums p.q and 1^T p per chunk, and between update
// and pupdate one launch applies the two-level coarse correction (jobs that have one) and
// sums r.z and r.r per chunk, so update and pupdate reduce chunk partials only.
void launch_pcg_iter(const JobDev* d_jobs, const int4* d_tiles, int ntiles, int* d_active, int iter,
                     int lg, int wpad, int variant, hipStream_t st, hipEvent_t ev0 = nullptr,
                     hipEvent_t ev1 = nullptr, const int2* d_js = nullptr, int njs = 0, int nchunk = 0,
                     const int4* d_ftiles = nullptr, int nftiles = 0);
// per-tile conductivity scale of the two-level preconditioner (after assembly); d1: the
// model's per-tile sigma = 1 diagonal sums; work items with slice != 0 are skipped
void launch_coarse_scale(const JobDev* d_jobs, const int4* d_tiles, int ntiles, const double* d1,
                         hipStream_t st);

// the SpMV of launch_pcg_iter alone (autotuning)
void launch_pcg_spmv(const JobDev* d_jobs, const int4* d_tiles, int ntiles, int lg, int wpad, int variant,
                     hipStream_t st, int par = 0);
// PCG: the x update that a slice stopping at an even iteration still owes
void launch_pcg_flush(const JobDev* d_jobs, const int4* d_tiles, int ntiles, hipStream_t st);
// every job of the tile list has 2^lg lanes; slice_bytes: the largest np * 2^lg * 8
void launch_residual64(const JobDev* d_jobs, const int4* d_tiles, int ntiles,
                       const int2* d_js, int njs, int lg, size_t slice_bytes, hipStream_t st);
void launch_x64_update(const JobDev* d_jobs, const int4* d_tiles, int ntiles, int first,
                       hipStream_t st);

}  // namespace mpf
