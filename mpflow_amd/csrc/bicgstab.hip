// Preconditioned BiCGSTAB for all (job, harmonic, right-hand side) columns at once
// (BiCGSTAB::solve, src/numeric/bicgstab.cpp:83-193; updateVector,
// src/numeric/bicgstab_kernel.cu:32-42), on the tile lists and FP64 tile partials of the
// PCG.  K^-1 is the Jacobi diagonal.
//
// The reference ignores dcFree (bicgstab.cpp:87) and iterates on the singular harmonic-0
// Neumann matrix.  In FP32 the rounding of A y and A z feeds a constant component into
// the Krylov vectors that grows until the method breaks down (NaN) on meshes of a few
// thousand nodes.  By default this build keeps nu and t of the job's dc slice (J.dc >= 0)
// orthogonal to the constant vector (nu = A y - mean(A y) 1, likewise t) and starts from
// r = r^ = b - mean(b) 1 (the consistent part of b), which changes nothing in exact
// arithmetic (1^T A = 0, 1^T b = 0) and keeps FP32 iterations stable; J.dc = -1
// runs the reference's iteration unchanged.  The constant mode of the solution reaches
// neither the voltages nor the Jacobian.
//
// Iteration i (reference step i), per column:
//   B1  rho_i = (r, r^); beta = (rho_i / rho_{i-1}) (alpha / omega) with x/0 = 0;
//       p = r + beta (p - omega nu); y = K^-1 p.  Tile 0's block also applies the stop
//       rule to r.r of the previous step and counts it.
//   B2  nu = A y; partials (nu, r^)
//   B3  alpha = rho_i / (nu, r^); s = r - alpha nu; z = K^-1 s
//   B4  t = A z; w = K^-1 t; partials (z, w), (w, w)
//   B5  omega = (z, w) / (w, w); x = (x + alpha y) + omega z; r = s - omega t;
//       partials (r, r^) and (r, r) for the next B1
// Each consumer block sums the partials it needs over all tiles itself, in one fixed
// order (bit-identical scalars in every block).  Partial channels: ch0 (nu, r^) in B2 and
// (r, r) in B5, ch1 (z, w), ch2 (w, w), ch3 (r, r^).  On the dc slice nu and t are stored
// raw and their means m_nu, m_t are applied by the consumers: B2 adds ch4 = 1^T nu; B4
// adds ch4 = 1^T t, ch5 = (z, K^-1 1), ch6 = (K^-1 t, K^-1 1), from which B5 forms
// (z, w) and (w, w) of the projected t.  rho lives in two slots by parity.
#include "krylov_dev.h"

namespace mpf {

namespace {

template <int LG>
struct Lay {
    static constexpr int Rp = 1 << LG, LPR = Rp / 4, ROWS = kEwBlock / LPR;
};

__device__ __forceinline__ float4 f4(const MPF_GLOBAL float4* p, size_t i) { return p[i]; }
__device__ __forceinline__ const MPF_GLOBAL float4* cv4(const float* p, size_t base) {
    return (const MPF_GLOBAL float4*)gp(p) + base;
}
__device__ __forceinline__ MPF_GLOBAL float4* mv4(float* p, size_t base) {
    return (MPF_GLOBAL float4*)gp(p) + base;
}
__device__ __forceinline__ float el(const float4& v, int j) {
    return j == 0 ? v.x : (j == 1 ? v.y : (j == 2 ? v.z : v.w));
}

}  // namespace

// init: x = 0, r = r^ = b, p = nu = 0; partials (r, r^) ch3, (r, r) ch0, (b, b) ch1;
// on the dc slice also 1^T r^ (ch4) and (K^-1 1, K^-1 1) (ch5, the same in every column)
__global__ void __launch_bounds__(kEwBlock) k_bicg_init(const JobDev* __restrict__ jobs,
                                                        const int4* __restrict__ tiles) {
    const int4 t = tiles[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    const Ew L = ew_of(J);
    const float xs = J.xscale[s];
    const float* dinv = J.dinv + (size_t)s * J.np;
    double bb = 0.0, sb = 0.0, dd = 0.0;
    for (int row = t.z * kTileRows + L.g; row < (t.z + 1) * kTileRows; row += L.G) {
        const size_t i = vec_index(J, s, row, L.lane);
        float b = 0.0f;
        if (L.lane < J.R) b = J.use_rhs ? J.rhs[i] : (float)(xs * J.F[s * J.fstride + (size_t)row * J.Rp + L.lane]);
        J.x[i] = 0.0f;
        J.rv[i] = b;
        J.rh[i] = b;
        J.p[i] = 0.0f;
        J.q[i] = 0.0f;
        bb += (double)b * (double)b;
        sb += (double)b;
        dd += (double)dinv[row] * (double)dinv[row];
    }
    const int ch[5] = {3, 0, 1, 4, 5};
    const double v[5] = {bb, bb, bb, sb, dd};
    block_partials<5, kEwBlock>(J, s, t.z, ch, v);
}

__global__ void __launch_bounds__(kEwBlock) k_bicg_init_reduce(const JobDev* __restrict__ jobs,
                                                               const int2* __restrict__ js,
                                                               int* __restrict__ active) {
    const int2 t = js[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    const int ch[3] = {1, 4, 5};
    double out[3];
    reduce_tiles<3>(J, s, ch, out);
    if ((int)threadIdx.x < J.Rp) {
        const int lane = threadIdx.x;
        // a correction solve started from a previous solution stops relative to the full
        // right-hand side ||beta_c F||^2 left in SC_X2 by the FP64 residual
        sc_ptr(J, SC_BB, s)[lane] = J.ref_b ? sc_ptr(J, SC_X2, s)[lane] : out[0];
        sc_ptr(J, rho_slot(1), s)[lane] = 1.0;   // rho_{-1}
        sc_ptr(J, SC_ALPHA, s)[lane] = 1.0;
        sc_ptr(J, SC_OMEGA, s)[lane] = 1.0;
        sc_ptr(J, SC_W, s)[lane] = out[1];       // 1^T r^
        sc_ptr(J, SC_RHO0, s)[lane] = out[2];    // (K^-1 1, K^-1 1)
        sc_ptr(J, SC_V, s)[lane] = 0.0;          // m_nu
        sc_ptr(J, SC_RS, s)[lane] = 0.0;         // m_t
        sc_ptr(J, SC_MB, s)[lane] = out[1] / (double)J.n;   // mean(b)
    }
    if (threadIdx.x == 0) {
        J.st[s * kStInts + 0] = 1;
        J.st[s * kStInts + 1] = 0;
        J.st[s * kStInts + 2] = 0;
        J.st[s * kStInts + 3] = 1;   // the next B1 is the first: no residual to test yet
        atomicAdd(active, 1);
    }
}

// dc slice: r = r^ = b - mean(b) 1 (rows < n); partials (r, r^) ch3, (r, r) ch0, 1^T r^ ch4
__global__ void __launch_bounds__(kEwBlock) k_bicg_project(const JobDev* __restrict__ jobs,
                                                           const int4* __restrict__ tiles) {
    const int4 t = tiles[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    if (s != J.dc) return;
    const Ew L = ew_of(J);
    const float m = (float)sc_ptr(J, SC_MB, s)[L.lane];
    double rr = 0.0, sr = 0.0;
    for (int row = t.z * kTileRows + L.g; row < (t.z + 1) * kTileRows; row += L.G) {
        const size_t i = vec_index(J, s, row, L.lane);
        if (row < J.n) {
            const float v = J.rv[i] - m;
            J.rv[i] = v;
            J.rh[i] = v;
            rr += (double)v * (double)v;
            sr += (double)v;
        }
    }
    const int ch[3] = {3, 0, 4};
    const double v[3] = {rr, rr, sr};
    block_partials<3, kEwBlock>(J, s, t.z, ch, v);
}

__global__ void __launch_bounds__(kEwBlock) k_bicg_project_reduce(const JobDev* __restrict__ jobs,
                                                                  const int2* __restrict__ js) {
    const int2 t = js[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    if (s != J.dc) return;
    const int ch[1] = {4};
    double out[1];
    reduce_tiles<1>(J, s, ch, out);
    if ((int)threadIdx.x < J.Rp) sc_ptr(J, SC_W, s)[threadIdx.x] = out[0];
}

// B1
template <int LG>
__global__ void __launch_bounds__(kEwBlock) k_bicg_p(const JobDev* __restrict__ jobs,
                                                     const int4* __restrict__ tiles, int iter,
                                                     int* __restrict__ active) {
    using Ly = Lay<LG>;
    const int4 t = tiles[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    if (!slice_active(J, s)) return;
    const int par = iter & 1;
    const bool book = t.z == 0;
    // only tile 0's block reads or writes the first-iteration flag
    const bool first = book && J.st[s * kStInts + 3] != 0;
    __shared__ float sh_b[Ly::Rp], sh_om[Ly::Rp], sh_m[Ly::Rp];
    __shared__ int not_done;
    {
        double out[2] = {0.0, 0.0};
        if (book) {
            const int ch[2] = {3, 0};
            reduce_tiles<2>(J, s, ch, out);
        } else {
            const int ch[1] = {3};
            double o[1];
            reduce_tiles<1>(J, s, ch, o);
            out[0] = o[0];
        }
        if (threadIdx.x == 0) not_done = 0;
        __syncthreads();
        if ((int)threadIdx.x < Ly::Rp) {
            const int lane = threadIdx.x;
            const double rho = out[0];
            const double rho_old = sc_ptr(J, rho_slot(par ^ 1), s)[lane];
            const double al = sc_ptr(J, SC_ALPHA, s)[lane], om = sc_ptr(J, SC_OMEGA, s)[lane];
            const double t1 = rho_old != 0.0 ? rho / rho_old : 0.0;
            const double t2 = om != 0.0 ? al / om : 0.0;
            sh_b[lane] = (float)(t1 * t2);
            sh_om[lane] = (float)om;
            sh_m[lane] = s == J.dc ? (float)sc_ptr(J, SC_V, s)[lane] : 0.0f;
            if (book) {
                sc_ptr(J, rho_slot(par), s)[lane] = rho;
                if (!first) {
                    const double rr = out[1];
                    const double bb = sc_ptr(J, SC_BB, s)[lane];
                    sc_ptr(J, SC_RR, s)[lane] = rr;
                    if (J.rule == MPF_STOP_REFERENCE) {
                        // column 0 only: sqrt(r.r / b.b) >= eps keeps iterating
                        // (bicgstab.cpp:179-187); NaN (b.b == 0) stops
                        if (lane == J.col0 && fabs(sqrt(rr / bb)) >= (double)FLT_EPSILON) not_done = 1;
                    } else if (lane < J.R) {
                        if (bb > 0.0 && rr > J.tol * J.tol * bb) atomicOr(&not_done, 1);
                    }
                    // a NaN residual with a nonzero right-hand side is a breakdown
                    if (lane < J.R && bb > 0.0 && rr != rr) atomicOr(&not_done, 2);
                }
            }
        }
        __syncthreads();
        if (book && threadIdx.x == 0 && first) J.st[s * kStInts + 3] = 0;
        if (book && !first && threadIdx.x == 0) {
            const int steps = J.st[s * kStInts + 1] + 1;
            J.st[s * kStInts + 1] = steps;
            if (not_done & 2) {
                J.st[s * kStInts + 0] = 0;
                J.st[s * kStInts + 2] = 1;
                atomicSub(active, 1);
            } else if (!not_done) {
                J.st[s * kStInts + 0] = 0;
                atomicSub(active, 1);
            } else if (steps >= J.maxit) {
                J.st[s * kStInts + 0] = 0;
                J.st[s * kStInts + 2] = 1;
                atomicSub(active, 1);
            }
        }
    }
    const int c4 = threadIdx.x % Ly::LPR, r0 = threadIdx.x / Ly::LPR;
    float b[4], om[4], mn[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        b[j] = sh_b[4 * c4 + j]; om[j] = sh_om[4 * c4 + j]; mn[j] = sh_m[4 * c4 + j];
    }
    const size_t base = (size_t)s * J.np * Ly::LPR;
    const MPF_GLOBAL float* dinv = gp((const float*)J.dinv) + (size_t)s * J.np;
    const MPF_GLOBAL float4* Rr = cv4(J.rv, base);
    const MPF_GLOBAL float4* Nu = cv4(J.q, base);
    MPF_GLOBAL float4* P = mv4(J.p, base);
    MPF_GLOBAL float4* Y = mv4(J.yv, base);
    for (int row = t.z * kTileRows + r0; row < (t.z + 1) * kTileRows; row += Ly::ROWS) {
        const size_t i = (size_t)row * Ly::LPR + c4;
        const float4 r4 = f4(Rr, i), n4 = f4(Nu, i), p4 = P[i];
        const float d = dinv[row];
        const float live = row < J.n ? 1.0f : 0.0f;
        float pn[4], yn[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float nu = el(n4, j) - live * mn[j];
            const float tmp = el(p4, j) - om[j] * nu;
            pn[j] = el(r4, j) + b[j] * tmp;
            yn[j] = d * pn[j];
        }
        P[i] = make_float4(pn[0], pn[1], pn[2], pn[3]);
        Y[i] = make_float4(yn[0], yn[1], yn[2], yn[3]);
    }
}

// B2: nu = A y; partials (nu, r^) -> ch0 (+ 1^T nu -> ch4 on the dc slice)
template <int LG, int WMAX>
__global__ void __launch_bounds__(kEwBlock) k_bicg_spmv_nu(const JobDev* __restrict__ jobs,
                                                           const int4* __restrict__ tiles) {
    using Ly = Lay<LG>;
    const int4 t = tiles[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    if (!slice_active(J, s)) return;
    const int wave = threadIdx.x >> 6, l64 = threadIdx.x & 63;
    const int row0 = t.z * kTileRows + wave * kSpmvRowsPerWave;
    const size_t slice_off = (size_t)s * J.np * Ly::Rp;
    const MPF_GLOBAL float4* RH = cv4(J.rh, slice_off / 4);
    MPF_GLOBAL float4* Nu = mv4(J.q, slice_off / 4);
    double v[2][4] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
    __shared__ EllStage<WMAX> stage;
    ell_gather_wave<LG, WMAX>(J, s, row0, l64, J.yv + slice_off, stage, [&](int row, int c4, float4 acc) {
        const float4 rh = RH[(size_t)row * Ly::LPR + c4];
        Nu[(size_t)row * Ly::LPR + c4] = acc;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            v[0][j] += (double)el(acc, j) * (double)el(rh, j);
            v[1][j] += (double)el(acc, j);
        }
    });
    if (s == J.dc) {
        const int ch[2] = {0, 4};
        block_partials4<LG, 2>(J, s, t.z, ch, v);
    } else {
        const int ch[1] = {0};
        double v1[1][4] = {{v[0][0], v[0][1], v[0][2], v[0][3]}};
        block_partials4<LG, 1>(J, s, t.z, ch, v1);
    }
}

// B3
template <int LG>
__global__ void __launch_bounds__(kEwBlock) k_bicg_s(const JobDev* __restrict__ jobs,
                                                     const int4* __restrict__ tiles, int iter) {
    using Ly = Lay<LG>;
    const int4 t = tiles[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    if (!slice_active(J, s)) return;
    const int par = iter & 1;
    const bool dc = s == J.dc;
    __shared__ float sh_a[Ly::Rp], sh_m[Ly::Rp];
    {
        double out[2] = {0.0, 0.0};
        if (dc) {
            const int ch[2] = {0, 4};
            reduce_tiles<2>(J, s, ch, out);
        } else {
            const int ch[1] = {0};
            double o[1];
            reduce_tiles<1>(J, s, ch, o);
            out[0] = o[0];
        }
        if ((int)threadIdx.x < Ly::Rp) {
            const int lane = threadIdx.x;
            // projected nu = nu - m 1 (rows < n): (nu, r^) -= m 1^T r^
            const double m = dc ? out[1] / (double)J.n : 0.0;
            const double nr = out[0] - (dc ? m * sc_ptr(J, SC_W, s)[lane] : 0.0);
            const double al = nr != 0.0 ? sc_ptr(J, rho_slot(par), s)[lane] / nr : 0.0;
            sh_a[lane] = (float)al;
            sh_m[lane] = (float)m;
            if (t.z == 0) {
                sc_ptr(J, SC_ALPHA, s)[lane] = al;
                sc_ptr(J, SC_V, s)[lane] = m;
            }
        }
        __syncthreads();
    }
    const int c4 = threadIdx.x % Ly::LPR, r0 = threadIdx.x / Ly::LPR;
    float a[4], mn[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) { a[j] = sh_a[4 * c4 + j]; mn[j] = sh_m[4 * c4 + j]; }
    const size_t base = (size_t)s * J.np * Ly::LPR;
    const MPF_GLOBAL float* dinv = gp((const float*)J.dinv) + (size_t)s * J.np;
    const MPF_GLOBAL float4* Rr = cv4(J.rv, base);
    const MPF_GLOBAL float4* Nu = cv4(J.q, base);
    MPF_GLOBAL float4* Sv = mv4(J.sv, base);
    MPF_GLOBAL float4* Z = mv4(J.zv, base);
    for (int row = t.z * kTileRows + r0; row < (t.z + 1) * kTileRows; row += Ly::ROWS) {
        const size_t i = (size_t)row * Ly::LPR + c4;
        const float4 r4 = f4(Rr, i), n4 = f4(Nu, i);
        const float d = dinv[row];
        const float live = row < J.n ? 1.0f : 0.0f;
        float sn[4], zn[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            sn[j] = el(r4, j) - a[j] * (el(n4, j) - live * mn[j]);
            zn[j] = d * sn[j];
        }
        Sv[i] = make_float4(sn[0], sn[1], sn[2], sn[3]);
        Z[i] = make_float4(zn[0], zn[1], zn[2], zn[3]);
    }
}

// B4: t = A z; w = K^-1 t; partials (z, w) -> ch1, (w, w) -> ch2 (+ 1^T t -> ch4,
// (z, K^-1 1) -> ch5, (w, K^-1 1) -> ch6 on the dc slice)
template <int LG, int WMAX>
__global__ void __launch_bounds__(kEwBlock) k_bicg_spmv_t(const JobDev* __restrict__ jobs,
                                                          const int4* __restrict__ tiles) {
    using Ly = Lay<LG>;
    const int4 t = tiles[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    if (!slice_active(J, s)) return;
    const int wave = threadIdx.x >> 6, l64 = threadIdx.x & 63;
    const int row0 = t.z * kTileRows + wave * kSpmvRowsPerWave;
    const size_t slice_off = (size_t)s * J.np * Ly::Rp;
    const MPF_GLOBAL float* dinv = gp((const float*)J.dinv) + (size_t)s * J.np;
    const MPF_GLOBAL float4* Z = cv4(J.zv, slice_off / 4);
    MPF_GLOBAL float4* T = mv4(J.tv, slice_off / 4);
    double v[5][4] = {};
    __shared__ EllStage<WMAX> stage;
    ell_gather_wave<LG, WMAX>(J, s, row0, l64, J.zv + slice_off, stage, [&](int row, int c4, float4 acc) {
        const float4 z = Z[(size_t)row * Ly::LPR + c4];
        T[(size_t)row * Ly::LPR + c4] = acc;
        const double d = (double)dinv[row];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const double w = (double)(dinv[row] * el(acc, j));
            v[0][j] += (double)el(z, j) * w;
            v[1][j] += w * w;
            v[2][j] += (double)el(acc, j);
            v[3][j] += (double)el(z, j) * d;
            v[4][j] += w * d;
        }
    });
    if (s == J.dc) {
        const int ch[5] = {1, 2, 4, 5, 6};
        block_partials4<LG, 5>(J, s, t.z, ch, v);
    } else {
        const int ch[2] = {1, 2};
        double v2[2][4] = {{v[0][0], v[0][1], v[0][2], v[0][3]}, {v[1][0], v[1][1], v[1][2], v[1][3]}};
        block_partials4<LG, 2>(J, s, t.z, ch, v2);
    }
}

// B5
template <int LG>
__global__ void __launch_bounds__(kEwBlock) k_bicg_xr(const JobDev* __restrict__ jobs,
                                                      const int4* __restrict__ tiles) {
    using Ly = Lay<LG>;
    const int4 t = tiles[blockIdx.x];
    const JobDev& J = jobs[t.x];
    const int s = t.y;
    if (!slice_active(J, s)) return;
    const bool dc = s == J.dc;
    __shared__ float sh_a[Ly::Rp], sh_o[Ly::Rp], sh_m[Ly::Rp];
    {
        double out[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (dc) {
            const int ch[5] = {1, 2, 4, 5, 6};
            reduce_tiles<5>(J, s, ch, out);
        } else {
            const int ch[2] = {1, 2};
            double o[2];
            reduce_tiles<2>(J, s, ch, o);
            out[0] = o[0]; out[1] = o[1];
        }
        if ((int)threadIdx.x < Ly::Rp) {
            const int lane = threadIdx.x;
            double zw = out[0], ww = out[1], m = 0.0;
            if (dc) {
                // projected t = t - m 1 (rows < n), w = K^-1 t
                m = out[2] / (double)J.n;
                zw -= m * out[3];
                ww += -2.0 * m * out[4] + m * m * sc_ptr(J, SC_RHO0, s)[lane];
            }
            const double om = ww != 0.0 ? zw / ww : 0.0;
            sh_o[lane] = (float)om;
            sh_m[lane] = (float)m;
            sh_a[lane] = (float)sc_ptr(J, SC_ALPHA, s)[lane];
            if (t.z == 0) {
                sc_ptr(J, SC_OMEGA, s)[lane] = om;
                sc_ptr(J, SC_RS, s)[lane] = m;
            }
        }
        __syncthreads();
    }
    const int c4 = threadIdx.x % Ly::LPR, r0 = threadIdx.x / Ly::LPR;
    float a[4], o[4], mt[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) { a[j] = sh_a[4 * c4 + j]; o[j] = sh_o[4 * c4 + j]; mt[j] = sh_m[4 * c4 + j]; }
    const size_t base = (size_t)s * J.np * Ly::LPR;
    const MPF_GLOBAL float4* Y = cv4(J.yv, base);
    const MPF_GLOBAL float4* Z = cv4(J.zv, base);
    const MPF_GLOBAL float4* Sv = cv4(J.sv, base);
    const MPF_GLOBAL float4* T = cv4(J.tv, base);
    const MPF_GLOBAL float4* RH = cv4(J.rh, base);
    MPF_GLOBAL float4* X = mv4(J.x, base);
    MPF_GLOBAL float4* Rr = mv4(J.rv, base);
    double v[2][4] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
    for (int row = t.z * kTileRows + r0; row < (t.z + 1) * kTileRows; row += Ly::ROWS) {
        const size_t i = (size_t)row * Ly::LPR + c4;
        const float4 y4 = f4(Y, i), z4 = f4(Z, i), s4 = f4(Sv, i), t4 = f4(T, i), h4 = f4(RH, i);
        const float4 x4 = X[i];
        const float live = row < J.n ? 1.0f : 0.0f;
        float xn[4], rn[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float x1 = el(x4, j) + a[j] * el(y4, j);
            xn[j] = x1 + o[j] * el(z4, j);
            rn[j] = el(s4, j) - o[j] * (el(t4, j) - live * mt[j]);
            v[0][j] += (double)rn[j] * (double)el(h4, j);
            v[1][j] += (double)rn[j] * (double)rn[j];
        }
        X[i] = make_float4(xn[0], xn[1], xn[2], xn[3]);
        Rr[i] = make_float4(rn[0], rn[1], rn[2], rn[3]);
    }
    const int ch[2] = {3, 0};
    block_partials4<LG, 2>(J, s, t.z, ch, v);
}

// --------------------------------------------------------------------------------
// launchers
// --------------------------------------------------------------------------------
void launch_bicg_init(const JobDev* d_jobs, const int4* d_tiles, int ntiles, const int2* d_js,
                      int njs, int* d_active, int lg, hipStream_t st) {
    (void)lg;
    k_bicg_init<<<ntiles, kEwBlock, 0, st>>>(d_jobs, d_tiles);
    k_bicg_init_reduce<<<njs, kEwBlock, 0, st>>>(d_jobs, d_js, d_active);
    k_bicg_project<<<ntiles, kEwBlock, 0, st>>>(d_jobs, d_tiles);
    k_bicg_project_reduce<<<njs, kEwBlock, 0, st>>>(d_jobs, d_js);
    MPF_HIP(hipGetLastError());
}

template <int LG>
static void bicg_iter(const JobDev* d_jobs, const int4* d_tiles, int n, int* d_active, int iter, int wpad,
                      hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
    k_bicg_p<LG><<<n, kEwBlock, 0, st>>>(d_jobs, d_tiles, iter, d_active);
    if (ev0) MPF_HIP(hipEventRecord(ev0, st));
    switch (wpad) {
        case 8:  k_bicg_spmv_nu<LG, 8><<<n, kEwBlock, 0, st>>>(d_jobs, d_tiles); break;
        case 16: k_bicg_spmv_nu<LG, 16><<<n, kEwBlock, 0, st>>>(d_jobs, d_tiles); break;
        case 24: k_bicg_spmv_nu<LG, 24><<<n, kEwBlock, 0, st>>>(d_jobs, d_tiles); break;
        default: throw Error(MPF_EINVAL, "BiCGSTAB: padded ELL width above 24 is not supported");
    }
    if (ev1) MPF_HIP(hipEventRecord(ev1, st));
    k_bicg_s<LG><<<n, kEwBlock, 0, st>>>(d_jobs, d_tiles, iter);
    switch (wpad) {
        case 8:  k_bicg_spmv_t<LG, 8><<<n, kEwBlock, 0, st>>>(d_jobs, d_tiles); break;
        case 16: k_bicg_spmv_t<LG, 16><<<n, kEwBlock, 0, st>>>(d_jobs, d_tiles); break;
        default: k_bicg_spmv_t<LG, 24><<<n, kEwBlock, 0, st>>>(d_jobs, d_tiles); break;
    }
    k_bicg_xr<LG><<<n, kEwBlock, 0, st>>>(d_jobs, d_tiles);
}

void launch_bicg_iter(const JobDev* d_jobs, const int4* d_tiles, int ntiles, int* d_active, int iter,
                      int lg, int wpad, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
    if (lg == 6) bicg_iter<6>(d_jobs, d_tiles, ntiles, d_active, iter, wpad, st, ev0, ev1);
    else if (lg == 5) bicg_iter<5>(d_jobs, d_tiles, ntiles, d_active, iter, wpad, st, ev0, ev1);
    else if (lg == 4) bicg_iter<4>(d_jobs, d_tiles, ntiles, d_active, iter, wpad, st, ev0, ev1);
    else throw Error(MPF_EINVAL, "BiCGSTAB: unsupported lane count");
    MPF_HIP(hipGetLastError());
}

}  // namespace mpf
