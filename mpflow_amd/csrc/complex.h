// Complex conductivity (models::EIT over Equation<thrust::complex<float>, ...>,
// src/models/eit.cpp:289-304): device job table and launch interface (complex.hip), and
// the host model (complex.cpp).
#pragma once

#include <memory>

#include "cplx.h"
#include "dense.h"

struct mpf_ctx;

namespace mpf {

constexpr int kCParts = 4;    // partial-sum channels per tile
// scalar fields per (slice, lane): rho of the previous step, b.b, p.q, 1^T p, rho, r.r,
// beta, and a flag that rho_old was zero (the reference then zeroes p)
enum ZField { ZF_RHO = 0, ZF_BB, ZF_PQ, ZF_PSUM, ZF_RHON, ZF_RR, ZF_BETA, ZF_BZERO, ZF_N };

// One frame of a complex-conductivity model.  Vectors are [slice][row][lane] of float2
// (re, im) with lane stride Rq (a power of two >= the unique right-hand sides, >= 2);
// a thread owns two consecutive lanes (one 16-byte load per neighbour row).
struct CJob {
    int n, np, width, C, Ru, Rq, ntiles, maxit, rule, col0, E, logarithmic;
    int herm;                    // 1: dot products conjugate their second operand (the
                                 // reference's iteration); 0: COCG, bilinear r^T z, p^T q
    double tol, ref;
    const int32_t* ids;          // [width][np], padding slots (row, 0)
    const int32_t* diag_slot;    // [np]
    float2* sv; float2* rv;      // S = sum sigma_e K_e, R = sum sigma_e M_e   [width][np]
    float2* dinv;                // [C][np] 1 / (S_ii + alpha_c R_ii)
    float alpha[kMaxSlices];
    float beta[kMaxSlices];
    const double* F;             // [np][Fp] real excitation B Pi, unique column u in lane u
    int Fp;
    float2* x; float2* r; float2* p; float2* q;   // [C][np][Rq]
    zd* part;                    // [kCParts][C][ntiles][Rq]
    zd* sc;                      // [ZF_N][C][Rq]
    int* st;                     // [C][4]: live, steps, capped, -
    const float2* gamma;         // [E]
    const int32_t* ctb_off; const int32_t* ctb_elem; const double* ctb_k; const double* ctb_m;
};

struct CMeasDev {
    const int32_t* w_off; const int32_t* w_node; const double* w_val;
    const int32_t* colmap; const double* colscale;
    int D, M, Dp, Mp;
};

struct CJacDev {
    const int32_t* elements;
    const float* ek; const float* em;
    const int32_t* colmap; const double* colscale;
    int E, k, D, M, Dp, Mp, exact;
    float alpha[kMaxSlices];
    float weight[kMaxSlices];
};

void launch_c_assemble(const CJob* d_jobs, const int4* d_items, int nitems, hipStream_t st);
void launch_c_dinv(const CJob* d_jobs, const int4* d_tiles, int ntiles, hipStream_t st);
void launch_c_init(const CJob* d_jobs, const int4* d_tiles, int ntiles, const int2* d_js, int njs, hipStream_t st);
// one CG iteration over the listed tiles of live (frame, slice) pairs
void launch_c_iter(const CJob* d_jobs, const int4* d_tiles, int ntiles, const int2* d_js, int njs, int* d_live,
                   hipStream_t st);
void launch_c_voltages(const CJob* d_jobs, int njobs, const CMeasDev& md, float2* U32, zd* U64, hipStream_t st);
void launch_c_jacobian(const CJob* d_jobs, int njobs, const CJacDev& jd, float2* J, hipStream_t st);
// G_b = J_b J_b^T + shift I (complex symmetric, FP64 sums of the FP32 J, both triangles);
// J_b column-major m x E
void launch_c_gram(int m, int E, const float2* J, size_t strideJ, double shift, zd* G, size_t strideG, int batch,
                   hipStream_t st);
// y_b = J_b^T x_b (FP64); gamma_b += y_b when gamma is given
void launch_c_gemv_t(int m, int E, const float2* J, size_t strideJ, const zd* x, size_t stridex, zd* y,
                     size_t stridey, int batch, float2* gamma, hipStream_t st);
// r = Um - U (complex FP64, n entries)
void launch_c_sub(int n, const zd* Um, const zd* U, zd* r, hipStream_t st);

// Host model of a complex-conductivity EIT: conjugate orthogonal CG (or the reference's
// iteration, whose dot products conjugate their second operand) with the diagonal
// preconditioner on every (frame, harmonic), complex voltages and Jacobian, and the
// Gauss-Newton step through the complex symmetric data-space normal equations.
struct ComplexModel {
    ComplexModel(mpf_ctx* ctx, const mpf_model_desc& d);
    ~ComplexModel();
    // d_gamma [frames][E] complex (float2); d_U [frames][m], d_U64 [frames][m] complex;
    // d_J [frames][E][m] complex (column-major m x E, the reference's Matrix layout)
    int forward(const mpf_solve_opts& o, const float2* d_gamma, float2* d_U, zd* d_U64, float2* d_J,
                mpf_solve_stats* stats);
    int gn_step(float2* d_gamma, const zd* d_Umeas, double lambda, const mpf_solve_opts& o, zd* d_Ucalc,
                mpf_solve_stats* stats);
    struct Impl;
    std::unique_ptr<Impl> impl;
    const MeshSetup& setup() const;
    int frames() const;
};

}  // namespace mpf
