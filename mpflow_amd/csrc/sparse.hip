// Sparse kernels of the forward solve on CDNA4 (gfx950):
//   sigma-weighted gather assembly of S and R, the Jacobi diagonal, and the
//   multi-job / multi-harmonic / multi-RHS Jacobi-PCG and BiCGSTAB building blocks.
//
// Work items are (job, slice, 64-row tile).  A 256-thread block covers one tile of one
// slice: lane = tid % Rp is the right-hand side, g = tid / Rp a row group, so one
// wavefront reads 64/Rp rows x Rp contiguous floats of every gathered neighbour.
// Dot products are accumulated in FP64 per lane, combined across row groups in a fixed
// order, written as one partial per tile, and summed per (job, slice) by a one-block
// reduction in tile order: every scalar is deterministic run to run.
#include <cfloat>
#include <cmath>

#include "jobs.h"

namespace mpf {

__device__ __forceinline__ double sigma_of(const JobDev& J, float g) {
    // matParam (src/fem/equation_kernel.cu:35-43)
    return J.logarithmic ? J.ref * exp(2.302585092994046 * (double)g / 10.0) : J.ref * (double)g;
}

// --------------------------------------------------------------------------------
// assembly: S = sum_e sigma_e K_e, R = sum_e sigma_e M_e on the ELL pattern
// (updateMatrixKernel, src/fem/equation_kernel.cu:60-82), FP64 sums split hi/lo
// --------------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_assemble(const JobDev* __restrict__ jobs,
                                                     const int4* __restrict__ items) {
    const int4 it = items[blockIdx.x];
    const JobDev& J = jobs[it.x];
    const size_t ent = (size_t)it.y * kBlock + threadIdx.x;
    if (ent >= (size_t)J.width * J.np) return;
    const int b = J.ctb_off[ent], e = J.ctb_off[ent + 1];
    double S = 0.0, R = 0.0;
    for (int i = b; i < e; ++i) {
        const double sg = sigma_of(J, J.gamma[J.ctb_elem[i]]);
        S += sg * J.ctb_k[i];
        R += sg * J.ctb_m[i];
    }
    const float sh = (float)S, rh = (float)R;
    J.s[ent] = sh; J.s_lo[ent] = (float)(S - (double)sh);
    J.r[ent] = rh; J.r_lo[ent] = (float)(R - (double)rh);
}

// Jacobi diagonal 1/(S_ii + alpha_c R_ii) in FP32 (diagonalSparseKernel,
// src/numeric/preconditioner_kernel.cu:63-83); padding rows 0.
__global__ void __launch_bounds__(kBlock) k_dinv(const JobDev* __restrict__ jobs,
                                                 const int4* __restrict__ tiles) {
    const int4 t = tiles[blockIdx.x];
    const JobDev& J = jobs[t.x];
    if (threadIdx.x >= kTileRows) return;
    const int row = t.z * kTileRows + threadIdx.x;
    const int s = t.y;
    float v = 0.0f;
    if (row < J.n) {
        const size_t e = (size_t)J.diag_slot[row] * J.np + row;
        const float a = J.s[e] + J.alpha[s] * J.r[e];
        v = a != 0.0f ? 1.0f / a : 0.0f;
    }
    J.dinv[(size_t)s * J.np + row] = v;
}

// --------------------------------------------------------------------------------
// helpers
// --------------------------------------------------------------------------------
struct Lanes {
    int lane, g, G, Rp, lg;
};

__device__ __forceinline__ Lanes lanes_of(const JobDev& J) {
    Lanes L;
    L.lg = J.lg;
    L.Rp = 1 << J.lg;
    L.lane = threadIdx.x & (L.Rp - 1);
    L.g = threadIdx.x >> J.lg;
    L.G = kBlock >> J.lg;
    return L;
}

__device__ __forceinline__ double* part_ptr(const JobDev& J, int ch, int s, int tile) {
    return J.part + (((size_t)ch * J.C + s) * J.ntiles + tile) * J.Rp;
}

__device__ __forceinline__ double* sc_ptr(const JobDev& J, int f, int s) {
    return J.sc + ((size_t)f * J.C + s) * J.Rp;
}

// Combine per-thread values of NV channels across the row groups of the block in a
// fixed order and write them as this tile's partials.
template <int NV>
__device__ __forceinline__ void block_partials(const JobDev& J, const Lanes& L, int s, int tile,
                                               const int* ch, double (&v)[NV]) {
    __shared__ double red[NV][kBlock];
#pragma unroll
    for (int c = 0; c < NV; ++c) red[c][threadIdx.x] = v[c];
    __syncthreads();
    if (L.g == 0) {
#pragma unroll
        for (int c = 0; c < NV; ++c) {
            double a = red[c][L.lane];
            for (int gg = 1; gg < L.G; ++gg) a += red[c][gg * L.Rp + L.lane];
            part_ptr(J, ch[c], s, tile)[L.lane] = a;
        }
    }
}

// Sum NV partial channels over the tiles of (job, slice) in tile order; the result is
// valid in threads with g == 0 (lane = column).
template <int NV>
__device__ __forceinline__ void reduce_tiles(const JobDev& J, const Lanes& L, int s,
                                             const int* ch, double (&out)[NV]) {
    __shared__ double red[NV][kBlock];
    double a[NV];
#pragma unroll
    for (int c = 0; c < NV; ++c) a[c] = 0.0;
    // each row group sums a contiguous range of tiles
    const int per = (J.ntiles + L.G - 1) / L.G;
    const int t0 = L.g * per, t1 = min(J.ntiles, t0 + per);
    for (int t = t0; t < t1; ++t) {
#pragma unroll
        for (int c = 0; c < NV; ++c) a[c] += part_ptr(J, ch[c], s, t)[L.lane];
    }
#pragma unroll
    for (int c = 0; c < NV; ++c) red[c][threadIdx.x] = a[c];
    __syncthreads();
    if (L.g == 0) {
#pragma unroll
        for (int c = 0; c < NV; ++c) {
            double r = red[c][L.lane];
            for (int gg = 1; gg < L.G; ++gg) r += red[c][gg * L.Rp + L.lane];
            out[c] = r;
        }
    }
}

__device__ __forceinline__ bool slice_active(const JobDev& J, int s) {
    return J.st[s * kStInts] != 0;
}

// y = A_s * v for one row and this lane (A_s = S + alpha_s R, FP32)
__device__ __forceinline__ float row_spmv(const JobDev& J, float al, const float* __restrict__ V,
                                          int row, int lane) {
    float acc = 0.0f;
    const int np = J.np;
    const int Rp = J.Rp;
    int q = 0;
    for (; q + 4 <= J.width; q += 4) {
        const size_t e = (size_t)q * np + row;
        const int i0 = J.ids[e], i1 = J.ids[e + np], i2 = J.ids[e + 2 * (size_t)np], i3 = J.ids[e + 3 * (size_t)np];
        const float v0 = J.s[e] + al * J.r[e];
        const float v1 = J.s[e + np] + al * J.r[e + np];
        const float v2 = J.s[e + 2 * (size_t)np] + al * J.r[e + 2 * (size_t)np];
        const float v3 = J.s[e + 3 * (size_t)np] + al * J.r[e + 3 * (size_t)np];
        const float x0 = V[(size_t)i0 * Rp + lane], x1 = V[(size_t)i1 * Rp + lane];
        const float x2 = V[(size_t)i2 * Rp + lane], x3 = V[(size_t)i3 * Rp + lane];
        acc += v0 * x0; acc += v1 * x1; acc += v2 * x2; acc += v3 * x3;
    }
    for (; q < J.width; ++q) {
        const size_t e = (size_t)q * np + row;
        acc += (J.s[e] + al * J.r[e]) * V[(size_t)J.ids[e] * Rp + lane];
    }
    return acc;
}

// --------------------------------------------------------------------------------
// PCG (ConjugateGradient::solve, src/numeric/conjugate_gradient.cpp:61-161)
// --------------------------------------------------------------------------------
// init: x = 0, r = b, z = D^-1 r, p = z; partials r.z and b.b
__global__ void __launch_bounds__(kBlock) k_pcg_init(const JobDev* __restrict__ jobs,
                                                     const int4* __restrict__ tiles) {
    const int4 t = tiles[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    const Lanes L = lanes_of(J);
    const float* dinv = J.dinv + (size_t)s * J.np;
    const float xs = J.xscale[s];
    double rz = 0.0, bb = 0.0;
    for (int row = t.z * kTileRows + L.g; row < (t.z + 1) * kTileRows; row += L.G) {
        const size_t i = vec_index(J, s, row, L.lane);
        float b = 0.0f;
        if (L.lane < J.R) b = J.use_rhs ? J.rhs[i] : (float)(xs * J.F[(size_t)row * J.Rp + L.lane]);
        const float z = dinv[row] * b;
        J.x[i] = 0.0f;
        J.rv[i] = b;
        J.p[i] = z;
        rz += (double)b * (double)z;
        bb += (double)b * (double)b;
    }
    const int ch[2] = {0, 1};
    double v[2] = {rz, bb};
    block_partials<2>(J, L, s, t.z, ch, v);
}

__global__ void __launch_bounds__(kBlock) k_pcg_init_reduce(const JobDev* __restrict__ jobs,
                                                            const int2* __restrict__ js,
                                                            int* __restrict__ active) {
    const int2 t = js[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    const Lanes L = lanes_of(J);
    const int ch[2] = {0, 1};
    double out[2];
    reduce_tiles<2>(J, L, s, ch, out);
    if (L.g == 0) {
        sc_ptr(J, SC_RHO, s)[L.lane] = out[0];
        sc_ptr(J, SC_BB, s)[L.lane] = out[1];
        sc_ptr(J, SC_RS, s)[L.lane] = out[0];
    }
    if (threadIdx.x == 0) {
        J.st[s * kStInts + 0] = 1;
        J.st[s * kStInts + 1] = 0;
        J.st[s * kStInts + 2] = 0;
        atomicAdd(active, 1);
    }
}

// q = A p (+ 1 1^T p when dcFree); partials p.q and sum(p)
__global__ void __launch_bounds__(kBlock) k_pcg_spmv(const JobDev* __restrict__ jobs,
                                                     const int4* __restrict__ tiles) {
    const int4 t = tiles[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    if (!slice_active(J, s)) return;
    const Lanes L = lanes_of(J);
    const float al = J.alpha[s];
    const float* P = J.p + (size_t)s * J.np * J.Rp;
    float* Q = J.q + (size_t)s * J.np * J.Rp;
    double pq = 0.0, ps = 0.0;
    for (int row = t.z * kTileRows + L.g; row < (t.z + 1) * kTileRows; row += L.G) {
        const float acc = row_spmv(J, al, P, row, L.lane);
        const size_t i = (size_t)row * J.Rp + L.lane;
        Q[i] = acc;
        const float pv = P[i];
        pq += (double)pv * (double)acc;
        ps += (double)pv;
    }
    const int ch[2] = {0, 1};
    double v[2] = {pq, ps};
    block_partials<2>(J, L, s, t.z, ch, v);
}

__global__ void __launch_bounds__(kBlock) k_pcg_reduce_pq(const JobDev* __restrict__ jobs,
                                                          const int2* __restrict__ js) {
    const int2 t = js[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    if (!slice_active(J, s)) return;
    const Lanes L = lanes_of(J);
    const int ch[2] = {0, 1};
    double out[2];
    reduce_tiles<2>(J, L, s, ch, out);
    if (L.g == 0) {
        double pq = out[0];
        const double ps = out[1];
        // with dcFree, q = A p + 1 (1^T p), hence p.q = p.Ap + (1^T p)^2
        if (s == J.dc) pq += ps * ps;
        const double rho = sc_ptr(J, SC_RHO, s)[L.lane];
        sc_ptr(J, SC_PQ, s)[L.lane] = pq;
        sc_ptr(J, SC_PSUM, s)[L.lane] = (s == J.dc) ? ps : 0.0;
        sc_ptr(J, SC_ALPHA, s)[L.lane] = pq != 0.0 ? rho / pq : 0.0;
    }
}

// x += alpha p, r -= alpha q (a zero p.q zeroes both, updateVectorKernel
// conjugate_gradient_kernel.cu:77-89); partials r.z and r.r with z = D^-1 r
__global__ void __launch_bounds__(kBlock) k_pcg_update(const JobDev* __restrict__ jobs,
                                                       const int4* __restrict__ tiles) {
    const int4 t = tiles[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    if (!slice_active(J, s)) return;
    const Lanes L = lanes_of(J);
    const double pq = sc_ptr(J, SC_PQ, s)[L.lane];
    const float a = (float)sc_ptr(J, SC_ALPHA, s)[L.lane];
    const float dcv = (float)sc_ptr(J, SC_PSUM, s)[L.lane];
    const bool zero = (pq == 0.0);
    const float* dinv = J.dinv + (size_t)s * J.np;
    double rz = 0.0, rr = 0.0;
    for (int row = t.z * kTileRows + L.g; row < (t.z + 1) * kTileRows; row += L.G) {
        const size_t i = vec_index(J, s, row, L.lane);
        const float pv = J.p[i], qv = J.q[i] + (row < J.n ? dcv : 0.0f);
        float xv = J.x[i], rv = J.rv[i];
        if (zero) { xv = 0.0f; rv = 0.0f; }
        else { xv = xv + a * pv; rv = rv - a * qv; }
        J.x[i] = xv;
        J.rv[i] = rv;
        const float z = dinv[row] * rv;
        rz += (double)rv * (double)z;
        rr += (double)rv * (double)rv;
    }
    const int ch[2] = {2, 3};
    double v[2] = {rz, rr};
    block_partials<2>(J, L, s, t.z, ch, v);
}

// rho = r.z, convergence test, beta = rho / rho_old
__global__ void __launch_bounds__(kBlock) k_pcg_reduce_rz(const JobDev* __restrict__ jobs,
                                                          const int2* __restrict__ js,
                                                          int* __restrict__ active) {
    const int2 t = js[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    if (!slice_active(J, s)) return;
    const Lanes L = lanes_of(J);
    const int ch[2] = {2, 3};
    double out[2];
    reduce_tiles<2>(J, L, s, ch, out);
    __shared__ int not_done;
    if (threadIdx.x == 0) not_done = 0;
    __syncthreads();
    double rho = 0.0;
    if (L.g == 0) {
        rho = out[0];
        const double rr = out[1];
        const double bb = sc_ptr(J, SC_BB, s)[L.lane];
        sc_ptr(J, SC_RR, s)[L.lane] = rr;
        if (J.rule == MPF_STOP_REFERENCE) {
            if (L.lane == 0) {
                // sqrt(|rho| / ref) >= eps keeps iterating; NaN (ref == 0) stops
                const double crit = sqrt(fabs(rho / bb));
                if (crit >= (double)FLT_EPSILON) not_done = 1;
            }
        } else if (L.lane < J.R) {
            if (bb > 0.0 && rr > J.tol * J.tol * bb) atomicOr(&not_done, 1);
        }
    }
    __syncthreads();
    if (L.g == 0) {
        const double rho_old = sc_ptr(J, SC_RHO, s)[L.lane];
        sc_ptr(J, SC_BETA, s)[L.lane] = rho_old != 0.0 ? rho / rho_old : 0.0;
        sc_ptr(J, SC_BZERO, s)[L.lane] = rho_old != 0.0 ? 0.0 : 1.0;
        sc_ptr(J, SC_RHO, s)[L.lane] = rho;
    }
    if (threadIdx.x == 0) {
        const int steps = J.st[s * kStInts + 1] + 1;
        J.st[s * kStInts + 1] = steps;
        if (!not_done) {
            J.st[s * kStInts + 0] = 0;
            atomicSub(active, 1);
        } else if (steps >= J.maxit) {
            J.st[s * kStInts + 0] = 0;
            J.st[s * kStInts + 2] = 1;
            atomicSub(active, 1);
        }
    }
}

// p = z + beta p (rho_old == 0 zeroes p)
__global__ void __launch_bounds__(kBlock) k_pcg_pupdate(const JobDev* __restrict__ jobs,
                                                        const int4* __restrict__ tiles) {
    const int4 t = tiles[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    if (!slice_active(J, s)) return;
    const Lanes L = lanes_of(J);
    const float b = (float)sc_ptr(J, SC_BETA, s)[L.lane];
    const bool zero = sc_ptr(J, SC_BZERO, s)[L.lane] != 0.0;
    const float* dinv = J.dinv + (size_t)s * J.np;
    for (int row = t.z * kTileRows + L.g; row < (t.z + 1) * kTileRows; row += L.G) {
        const size_t i = vec_index(J, s, row, L.lane);
        J.p[i] = zero ? 0.0f : dinv[row] * J.rv[i] + b * J.p[i];
    }
}

// --------------------------------------------------------------------------------
// FP64 refinement: residual r64 = beta F - (A64 + dc) x64 written as the FP32 RHS
// --------------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_colsum64(const JobDev* __restrict__ jobs,
                                                     const int4* __restrict__ tiles) {
    const int4 t = tiles[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    const Lanes L = lanes_of(J);
    double a = 0.0;
    if (s == J.dc)
        for (int row = t.z * kTileRows + L.g; row < (t.z + 1) * kTileRows; row += L.G)
            a += J.x64[vec_index(J, s, row, L.lane)];
    const int ch[1] = {0};
    double v[1] = {a};
    block_partials<1>(J, L, s, t.z, ch, v);
}

__global__ void __launch_bounds__(kBlock) k_colsum64_reduce(const JobDev* __restrict__ jobs,
                                                            const int2* __restrict__ js) {
    const int2 t = js[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    const Lanes L = lanes_of(J);
    const int ch[1] = {0};
    double out[1];
    reduce_tiles<1>(J, L, s, ch, out);
    if (L.g == 0) sc_ptr(J, SC_X0, s)[L.lane] = (s == J.dc) ? out[0] : 0.0;
}

__global__ void __launch_bounds__(kBlock) k_residual64(const JobDev* __restrict__ jobs,
                                                       const int4* __restrict__ tiles,
                                                       float* __restrict__ rhs_out_unused) {
    const int4 t = tiles[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    const Lanes L = lanes_of(J);
    const double al = (double)J.alpha[s];
    const double xs = (double)J.xscale[s];
    const double dcs = sc_ptr(J, SC_X0, s)[L.lane];
    const double* X = J.x64 + (size_t)s * J.np * J.Rp;
    float* out = J.rhs;
    double rr = 0.0, bb = 0.0;
    for (int row = t.z * kTileRows + L.g; row < (t.z + 1) * kTileRows; row += L.G) {
        double acc = 0.0;
        for (int q = 0; q < J.width; ++q) {
            const size_t e = (size_t)q * J.np + row;
            const double v = ((double)J.s[e] + (double)J.s_lo[e]) + al * ((double)J.r[e] + (double)J.r_lo[e]);
            acc += v * X[(size_t)J.ids[e] * J.Rp + L.lane];
        }
        double b = 0.0, r = 0.0;
        if (row < J.n && L.lane < J.R) {
            b = xs * J.F[(size_t)row * J.Rp + L.lane];
            r = b - acc - dcs;
        }
        out[vec_index(J, s, row, L.lane)] = (float)r;
        rr += r * r;
        bb += b * b;
    }
    const int ch[2] = {0, 1};
    double v[2] = {rr, bb};
    block_partials<2>(J, L, s, t.z, ch, v);
}

__global__ void __launch_bounds__(kBlock) k_residual64_reduce(const JobDev* __restrict__ jobs,
                                                              const int2* __restrict__ js) {
    const int2 t = js[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    const Lanes L = lanes_of(J);
    const int ch[2] = {0, 1};
    double out[2];
    reduce_tiles<2>(J, L, s, ch, out);
    if (L.g == 0) {
        sc_ptr(J, SC_X1, s)[L.lane] = out[0];
        sc_ptr(J, SC_X2, s)[L.lane] = out[1];
    }
}

// x64 = x (first) or x64 += x
__global__ void __launch_bounds__(kBlock) k_x64_update(const JobDev* __restrict__ jobs,
                                                       const int4* __restrict__ tiles, int first) {
    const int4 t = tiles[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    const Lanes L = lanes_of(J);
    for (int row = t.z * kTileRows + L.g; row < (t.z + 1) * kTileRows; row += L.G) {
        const size_t i = vec_index(J, s, row, L.lane);
        const double v = (double)J.x[i];
        J.x64[i] = first ? v : J.x64[i] + v;
    }
}

// --------------------------------------------------------------------------------
// stand-alone multi-harmonic ELL SpMV (no job table)
// --------------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_spmv_plain(int np, int width, const int32_t* __restrict__ ids,
                                                       const float* __restrict__ S, const float* __restrict__ Rv,
                                                       const float* __restrict__ alpha,
                                                       const float* __restrict__ X, float* __restrict__ Y,
                                                       int lg) {
    const int Rp = 1 << lg;
    const int lane = threadIdx.x & (Rp - 1), g = threadIdx.x >> lg, G = kBlock >> lg;
    const int c = blockIdx.y;
    const float al = alpha[c];
    const float* V = X + (size_t)c * np * Rp;
    float* Out = Y + (size_t)c * np * Rp;
    for (int row = blockIdx.x * kTileRows + g; row < (blockIdx.x + 1) * kTileRows; row += G) {
        float acc = 0.0f;
        for (int q = 0; q < width; ++q) {
            const size_t e = (size_t)q * np + row;
            acc += (S[e] + al * Rv[e]) * V[(size_t)ids[e] * Rp + lane];
        }
        Out[(size_t)row * Rp + lane] = acc;
    }
}

// --------------------------------------------------------------------------------
// launch wrappers
// --------------------------------------------------------------------------------
static inline void check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error(MPF_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
}

void launch_assemble(const JobDev* d_jobs, const int4* d_items, int nitems, hipStream_t st) {
    if (!nitems) return;
    k_assemble<<<nitems, kBlock, 0, st>>>(d_jobs, d_items);
    check_launch();
}

void launch_dinv(const JobDev* d_jobs, const int4* d_tiles, int ntiles, hipStream_t st) {
    k_dinv<<<ntiles, kBlock, 0, st>>>(d_jobs, d_tiles);
    check_launch();
}

void launch_pcg_init(const JobDev* d_jobs, const int4* d_tiles, int ntiles, const int2* d_js,
                     int njs, int* d_active, hipStream_t st) {
    k_pcg_init<<<ntiles, kBlock, 0, st>>>(d_jobs, d_tiles);
    k_pcg_init_reduce<<<njs, kBlock, 0, st>>>(d_jobs, d_js, d_active);
    check_launch();
}

void launch_pcg_iter(const JobDev* d_jobs, const int4* d_tiles, int ntiles, const int2* d_js,
                     int njs, int* d_active, hipStream_t st) {
    k_pcg_spmv<<<ntiles, kBlock, 0, st>>>(d_jobs, d_tiles);
    k_pcg_reduce_pq<<<njs, kBlock, 0, st>>>(d_jobs, d_js);
    k_pcg_update<<<ntiles, kBlock, 0, st>>>(d_jobs, d_tiles);
    k_pcg_reduce_rz<<<njs, kBlock, 0, st>>>(d_jobs, d_js, d_active);
    k_pcg_pupdate<<<ntiles, kBlock, 0, st>>>(d_jobs, d_tiles);
    check_launch();
}

void launch_residual64(const JobDev* d_jobs, const int4* d_tiles, int ntiles, const int2* d_js,
                       int njs, int first, hipStream_t st) {
    (void)first;
    k_colsum64<<<ntiles, kBlock, 0, st>>>(d_jobs, d_tiles);
    k_colsum64_reduce<<<njs, kBlock, 0, st>>>(d_jobs, d_js);
    k_residual64<<<ntiles, kBlock, 0, st>>>(d_jobs, d_tiles, nullptr);
    k_residual64_reduce<<<njs, kBlock, 0, st>>>(d_jobs, d_js);
    check_launch();
}

void launch_x64_update(const JobDev* d_jobs, const int4* d_tiles, int ntiles, int first,
                       hipStream_t st) {
    k_x64_update<<<ntiles, kBlock, 0, st>>>(d_jobs, d_tiles, first);
    check_launch();
}

void launch_spmv(int np, int width, const int32_t* ids, const float* S, const float* R, int C,
                 const float* alpha, const float* X, float* Y, int Rp, hipStream_t st) {
    const int lg = Rp == 16 ? 4 : (Rp == 32 ? 5 : 6);
    dim3 grid(np / kTileRows, C);
    k_spmv_plain<<<grid, kBlock, 0, st>>>(np, width, ids, S, R, alpha, X, Y, lg);
    check_launch();
}

}  // namespace mpf
