// Complex-conductivity EIT kernels (gfx950).  The system A_c = S + alpha_c R with
// S = sum_e sigma_e K_e, R = sum_e sigma_e M_e is complex symmetric; fields, voltages and
// the Jacobian are complex.  CG follows ConjugateGradient<thrust::complex<float>>
// (src/numeric/conjugate_gradient.cpp:61-161, 231-232) step for step; its dot products
// either conjugate the second operand as the reference's do (vectorDotProductKernel,
// src/numeric/matrix_kernel.cu:405-413: rho = sum r conj(z), p.q = sum p conj(q)), or,
// by default, are the bilinear r^T z and p^T q of conjugate orthogonal CG, which
// converges on complex symmetric systems where the conjugated form stalls.  b.b and r.r
// stay Hermitian norms.  Vectors are FP32 complex, dot products and CG scalars FP64.
#include "complex.h"

namespace mpf {

static inline void check_launch() { MPF_HIP(hipGetLastError()); }

constexpr int kCBlock = 256;   // threads per block of the tile kernels (one 256-row tile)
constexpr double kLn10 = 2.302585092994046;

__device__ __forceinline__ bool c_live(const CJob& J, int s) { return J.st[s * 4] != 0; }
// one term a . b of a Krylov dot product: a conj(b) (reference) or a b (COCG)
__device__ __forceinline__ zd c_dot(const CJob& J, zd a, zd b) { return J.herm ? a * zconj(b) : a * b; }
__device__ __forceinline__ size_t c_idx(const CJob& J, int s, int row, int lane) {
    return ((size_t)s * J.np + row) * J.Rq + lane;
}
__device__ __forceinline__ zd* c_part(const CJob& J, int ch, int s, int tile) {
    return J.part + (((size_t)ch * J.C + s) * J.ntiles + tile) * J.Rq;
}
__device__ __forceinline__ zd* c_sc(const CJob& J, int f, int s) { return J.sc + ((size_t)f * J.C + s) * J.Rq; }

// sigma(gamma) (matParam, equation_kernel.cu:35-43) for complex gamma
__device__ __forceinline__ zd c_sigma(const CJob& J, float2 g) {
    const zd gz = zf(g);
    return J.logarithmic ? J.ref * zexp((kLn10 / 10.0) * gz) : J.ref * gz;
}

// thread layout of the tile kernels: TPR = Rq / 2 threads per row (two lanes each), RPP
// rows per pass, TPR passes over the tile's 256 rows
struct CLay {
    int tpr, rpp, rr, l2;
    __device__ CLay(int Rq) : tpr(Rq / 2), rpp(kCBlock / (Rq / 2)), rr(threadIdx.x / (Rq / 2)), l2(threadIdx.x % (Rq / 2)) {}
};

// per-thread FP64 values of NV channels (two lanes each) summed over the block's rows in
// a fixed order and written as the tile's partials
template <int NV>
__device__ __forceinline__ void c_block_partials(const CJob& J, int s, int tile, const int* ch, zd (&v)[NV][2]) {
    __shared__ zd red[NV][2 * kCBlock];
    const CLay ly(J.Rq);
#pragma unroll
    for (int c = 0; c < NV; ++c) {
        red[c][ly.rr * J.Rq + 2 * ly.l2] = v[c][0];
        red[c][ly.rr * J.Rq + 2 * ly.l2 + 1] = v[c][1];
    }
    __syncthreads();
    if ((int)threadIdx.x < J.Rq)
#pragma unroll
        for (int c = 0; c < NV; ++c) {
            zd a = zmake(0.0);
            for (int k = 0; k < ly.rpp; ++k) a += red[c][k * J.Rq + threadIdx.x];
            c_part(J, ch[c], s, tile)[threadIdx.x] = a;
        }
}

// S and R of every ELL entry from the gather plan (Equation::update, equation.cpp:261-274)
__global__ void __launch_bounds__(kCBlock) k_c_assemble(const CJob* __restrict__ jobs, const int4* __restrict__ items) {
    const int4 it = items[blockIdx.x];
    const CJob& J = jobs[it.x];
    const size_t ent = (size_t)it.y * kCBlock + threadIdx.x;
    if (ent >= (size_t)J.width * J.np) return;
    zd S = zmake(0.0), R = zmake(0.0);
    for (int i = J.ctb_off[ent]; i < J.ctb_off[ent + 1]; ++i) {
        const zd sg = c_sigma(J, J.gamma[J.ctb_elem[i]]);
        S += J.ctb_k[i] * sg;
        R += J.ctb_m[i] * sg;
    }
    J.sv[ent] = fz(S);
    J.rv[ent] = fz(R);
}

// diagonal preconditioner 1 / (S_ii + alpha_c R_ii) (diagonalSparseKernel,
// preconditioner_kernel.cu:63-83); padding rows 0
__global__ void __launch_bounds__(kCBlock) k_c_dinv(const CJob* __restrict__ jobs, const int4* __restrict__ tiles) {
    const int4 t = tiles[blockIdx.x];
    const CJob& J = jobs[t.x];
    const int s = t.y, row = t.z * kTileRows + threadIdx.x;
    zd v = zmake(0.0);
    if (row < J.n && J.diag_slot[row] >= 0) {
        const size_t e = (size_t)J.diag_slot[row] * J.np + row;
        const zd a = zf(J.sv[e]) + (double)J.alpha[s] * zf(J.rv[e]);
        if (!zzero(a)) v = zmake(1.0) / a;
    }
    J.dinv[(size_t)s * J.np + row] = fz(v);
}

// x = 0, r = beta_c F, z = D^-1 r, p = z; partials rho = r.z (ch 0), b.b (ch 1)
__global__ void __launch_bounds__(kCBlock) k_c_init(const CJob* __restrict__ jobs, const int4* __restrict__ tiles) {
    const int4 t = tiles[blockIdx.x];
    const CJob& J = jobs[t.x];
    const int s = t.y;
    const CLay ly(J.Rq);
    zd v[2][2] = {{zmake(0.0), zmake(0.0)}, {zmake(0.0), zmake(0.0)}};
    for (int ps = 0; ps < ly.tpr; ++ps) {
        const int row = t.z * kTileRows + ps * ly.rpp + ly.rr;
        const zd d = zf(J.dinv[(size_t)s * J.np + row]);
        float2 xs[2], rs[2], zs[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int lane = 2 * ly.l2 + h;
            const double b = lane < J.Ru ? (double)J.beta[s] * J.F[(size_t)row * J.Fp + lane] : 0.0;
            const zd z = d * zmake(b);
            xs[h] = make_float2(0.0f, 0.0f);
            rs[h] = make_float2((float)b, 0.0f);
            zs[h] = fz(z);
            v[0][h] += c_dot(J, zf(rs[h]), zf(zs[h]));
            v[1][h] += zmake(b * b);
        }
        const size_t i = c_idx(J, s, row, 2 * ly.l2);
        *reinterpret_cast<float4*>(J.x + i) = make_float4(xs[0].x, xs[0].y, xs[1].x, xs[1].y);
        *reinterpret_cast<float4*>(J.r + i) = make_float4(rs[0].x, rs[0].y, rs[1].x, rs[1].y);
        *reinterpret_cast<float4*>(J.p + i) = make_float4(zs[0].x, zs[0].y, zs[1].x, zs[1].y);
    }
    if (t.z == 0 && threadIdx.x == 0) {
        J.st[s * 4 + 0] = 1; J.st[s * 4 + 1] = 0; J.st[s * 4 + 2] = 0; J.st[s * 4 + 3] = 0;
    }
    const int ch[2] = {0, 1};
    c_block_partials<2>(J, s, t.z, ch, v);
}

// q = A_s p; partials p.q (ch 0) and 1^T p (ch 1)
__global__ void __launch_bounds__(kCBlock) k_c_spmv(const CJob* __restrict__ jobs, const int4* __restrict__ tiles) {
    const int4 t = tiles[blockIdx.x];
    const CJob& J = jobs[t.x];
    const int s = t.y;
    if (!c_live(J, s)) return;
    const CLay ly(J.Rq);
    const float al = J.alpha[s];
    const float2* P = J.p + (size_t)s * J.np * J.Rq;
    zd v[2][2] = {{zmake(0.0), zmake(0.0)}, {zmake(0.0), zmake(0.0)}};
    for (int ps = 0; ps < ly.tpr; ++ps) {
        const int row = t.z * kTileRows + ps * ly.rpp + ly.rr;
        float a0r = 0.0f, a0i = 0.0f, a1r = 0.0f, a1i = 0.0f;
        for (int q = 0; q < J.width; ++q) {
            const size_t e = (size_t)q * J.np + row;
            const int id = J.ids[e];
            const float2 sv = J.sv[e], rv = J.rv[e];
            const float ar = sv.x + al * rv.x, ai = sv.y + al * rv.y;
            const float4 x = *reinterpret_cast<const float4*>(P + (size_t)id * J.Rq + 2 * ly.l2);
            a0r += ar * x.x - ai * x.y; a0i += ar * x.y + ai * x.x;
            a1r += ar * x.z - ai * x.w; a1i += ar * x.w + ai * x.z;
        }
        const size_t i = c_idx(J, s, row, 2 * ly.l2);
        *reinterpret_cast<float4*>(J.q + i) = make_float4(a0r, a0i, a1r, a1i);
        const float4 pp = *reinterpret_cast<const float4*>(J.p + i);
        const zd p0 = zmake(pp.x, pp.y), p1 = zmake(pp.z, pp.w);
        v[0][0] += c_dot(J, p0, zmake(a0r, a0i));
        v[0][1] += c_dot(J, p1, zmake(a1r, a1i));
        v[1][0] += p0;
        v[1][1] += p1;
    }
    const int ch[2] = {0, 1};
    c_block_partials<2>(J, s, t.z, ch, v);
}

// alpha = rho / (p.q + (1^T p) . (1^T p) on the dcFree harmonic); r -= alpha (q + 1 1^T p),
// x += alpha p (a zero denominator zeroes both, updateVectorKernel,
// conjugate_gradient_kernel.cu:77-89); partials rho = r.z (ch 2) and r.r (ch 3)
__global__ void __launch_bounds__(kCBlock) k_c_update(const CJob* __restrict__ jobs, const int4* __restrict__ tiles) {
    const int4 t = tiles[blockIdx.x];
    const CJob& J = jobs[t.x];
    const int s = t.y;
    if (!c_live(J, s)) return;
    const CLay ly(J.Rq);
    const bool dc = s == 0;
    zd al[2], ps[2];
    bool zero[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int lane = 2 * ly.l2 + h;
        ps[h] = c_sc(J, ZF_PSUM, s)[lane];
        const zd pq = c_sc(J, ZF_PQ, s)[lane] + (dc ? c_dot(J, ps[h], ps[h]) : zmake(0.0));
        zero[h] = zzero(pq);
        al[h] = zero[h] ? zmake(0.0) : c_sc(J, ZF_RHO, s)[lane] / pq;
    }
    zd v[2][2] = {{zmake(0.0), zmake(0.0)}, {zmake(0.0), zmake(0.0)}};
    for (int pass = 0; pass < ly.tpr; ++pass) {
        const int row = t.z * kTileRows + pass * ly.rpp + ly.rr;
        const size_t i = c_idx(J, s, row, 2 * ly.l2);
        const float4 q4 = *reinterpret_cast<const float4*>(J.q + i);
        const float4 r4 = *reinterpret_cast<const float4*>(J.r + i);
        const float4 x4 = *reinterpret_cast<const float4*>(J.x + i);
        const float4 p4 = *reinterpret_cast<const float4*>(J.p + i);
        const zd d = zf(J.dinv[(size_t)s * J.np + row]);
        const zd qv[2] = {zmake(q4.x, q4.y), zmake(q4.z, q4.w)};
        const zd rv[2] = {zmake(r4.x, r4.y), zmake(r4.z, r4.w)};
        const zd xv[2] = {zmake(x4.x, x4.y), zmake(x4.z, x4.w)};
        const zd pv[2] = {zmake(p4.x, p4.y), zmake(p4.z, p4.w)};
        float2 ro[2], xo[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (zero[h]) {
                ro[h] = make_float2(0.0f, 0.0f);
                xo[h] = make_float2(0.0f, 0.0f);
            } else {
                const zd qf = (dc && row < J.n) ? qv[h] + ps[h] : qv[h];
                ro[h] = fz(rv[h] - al[h] * qf);
                xo[h] = fz(xv[h] + al[h] * pv[h]);
            }
            const zd rn = zf(ro[h]);
            const zd z = zf(fz(d * rn));
            v[0][h] += c_dot(J, rn, z);
            v[1][h] += zmake(zabs2(rn));
        }
        *reinterpret_cast<float4*>(J.r + i) = make_float4(ro[0].x, ro[0].y, ro[1].x, ro[1].y);
        *reinterpret_cast<float4*>(J.x + i) = make_float4(xo[0].x, xo[0].y, xo[1].x, xo[1].y);
    }
    const int ch[2] = {2, 3};
    c_block_partials<2>(J, s, t.z, ch, v);
}

// tile partials of channels ch_a / ch_b summed in tile order into scalar fields f_a /
// f_b; one block per listed (job, slice) pair; live_only skips finished pairs
__global__ void __launch_bounds__(128) k_c_reduce(const CJob* __restrict__ jobs, const int2* __restrict__ js, int ch_a,
                                                  int f_a, int ch_b, int f_b, int live_only) {
    const int2 pr = js[blockIdx.x];
    const CJob& J = jobs[pr.x];
    const int s = pr.y;
    if (live_only && !c_live(J, s)) return;
    const int which = threadIdx.x / J.Rq, lane = threadIdx.x % J.Rq;
    if (which >= 2) return;
    const int ch = which ? ch_b : ch_a, f = which ? f_b : f_a;
    zd a = zmake(0.0);
    for (int t = 0; t < J.ntiles; ++t) a += c_part(J, ch, s, t)[lane];
    c_sc(J, f, s)[lane] = a;
}

// stop rule and beta per live (job, slice): the reference rule stops when
// |sqrt(rho / b.b)| of the first pattern column is not >= FLT_EPSILON
// (conjugate_gradient.cpp:141-151); the tolerance rule when every column has
// ||r|| <= tol ||b||; a non-finite residual or the iteration cap stops and reports
__global__ void __launch_bounds__(64) k_c_state(const CJob* __restrict__ jobs, const int2* __restrict__ js,
                                                int* __restrict__ live) {
    const int2 pr = js[blockIdx.x];
    const CJob& J = jobs[pr.x];
    const int s = pr.y;
    if (!c_live(J, s)) return;
    __shared__ int unconv, bad;
    if (threadIdx.x == 0) { unconv = 0; bad = 0; }
    __syncthreads();
    const int lane = threadIdx.x;
    if (lane < J.Rq) {
        const zd rho = c_sc(J, ZF_RHON, s)[lane], ro = c_sc(J, ZF_RHO, s)[lane];
        const double rr = c_sc(J, ZF_RR, s)[lane].x, bb = c_sc(J, ZF_BB, s)[lane].x;
        if (lane < J.Ru) {
            if (!isfinite(rr) || !isfinite(rho.x) || !isfinite(rho.y)) atomicOr(&bad, 1);
            if (bb > 0.0 && !(rr <= J.tol * J.tol * bb)) atomicOr(&unconv, 1);
        }
        const bool z0 = zzero(ro);
        c_sc(J, ZF_BZERO, s)[lane] = zmake(z0 ? 1.0 : 0.0);
        c_sc(J, ZF_BETA, s)[lane] = z0 ? zmake(0.0) : rho / ro;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int* st = J.st + s * 4;
        const int steps = st[1] + 1;
        st[1] = steps;
        bool stop;
        if (J.rule == MPF_STOP_REFERENCE) {
            const zd rho = c_sc(J, ZF_RHON, s)[J.col0];
            const double bb = c_sc(J, ZF_BB, s)[J.col0].x;
            const double crit = sqrt(sqrt(zabs2(rho)) / bb);
            stop = !(crit >= 1.1920928955078125e-07);
        } else {
            stop = !unconv;
        }
        if (bad || steps >= J.maxit) { stop = true; st[2] = 1; }
        if (stop) {
            st[0] = 0;
            atomicSub(live, 1);
        }
    }
    __syncthreads();
    // rho of this step becomes rho_old
    if (lane < J.Rq) c_sc(J, ZF_RHO, s)[lane] = c_sc(J, ZF_RHON, s)[lane];
}

// p = z + beta p (0 when rho_old was 0), z = D^-1 r
__global__ void __launch_bounds__(kCBlock) k_c_pupdate(const CJob* __restrict__ jobs, const int4* __restrict__ tiles) {
    const int4 t = tiles[blockIdx.x];
    const CJob& J = jobs[t.x];
    const int s = t.y;
    if (!c_live(J, s)) return;
    const CLay ly(J.Rq);
    zd be[2];
    bool zero[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int lane = 2 * ly.l2 + h;
        be[h] = c_sc(J, ZF_BETA, s)[lane];
        zero[h] = c_sc(J, ZF_BZERO, s)[lane].x != 0.0;
    }
    for (int pass = 0; pass < ly.tpr; ++pass) {
        const int row = t.z * kTileRows + pass * ly.rpp + ly.rr;
        const size_t i = c_idx(J, s, row, 2 * ly.l2);
        const float4 r4 = *reinterpret_cast<const float4*>(J.r + i);
        const float4 p4 = *reinterpret_cast<const float4*>(J.p + i);
        const zd d = zf(J.dinv[(size_t)s * J.np + row]);
        const zd rv[2] = {zmake(r4.x, r4.y), zmake(r4.z, r4.w)};
        const zd pv[2] = {zmake(p4.x, p4.y), zmake(p4.z, p4.w)};
        float2 po[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const zd z = zf(fz(d * rv[h]));
            po[h] = zero[h] ? make_float2(0.0f, 0.0f) : fz(z + be[h] * pv[h]);
        }
        *reinterpret_cast<float4*>(J.p + i) = make_float4(po[0].x, po[0].y, po[1].x, po[1].y);
    }
}

// U[d Mp + m] = sum_c sum_n W[m, n] colscale_d Phi_c[n, colmap_d] (complex)
__global__ void __launch_bounds__(256) k_c_voltages(const CJob* __restrict__ jobs, CMeasDev md, float2* __restrict__ U32,
                                                    zd* __restrict__ U64) {
    const int j = blockIdx.x;
    const CJob& J = jobs[j];
    const int mrows = md.Mp * md.Dp;
    for (int idx = threadIdx.x; idx < mrows; idx += blockDim.x) {
        const int d = idx / md.Mp, m = idx % md.Mp;
        zd v = zmake(0.0);
        if (d < md.D && m < md.M)
            for (int s = 0; s < J.C; ++s)
                for (int q = md.w_off[m]; q < md.w_off[m + 1]; ++q)
                    v += (md.w_val[q] * md.colscale[d]) * zf(J.x[c_idx(J, s, md.w_node[q], md.colmap[d])]);
        if (U32) U32[(size_t)j * mrows + idx] = fz(v);
        if (U64) U64[(size_t)j * mrows + idx] = v;
    }
}

// J[d Mp + m, e] = -sigma'_e sum_c w_c Phi_c[el, d]^T Kc_e Phi_c[el, D + m] with complex
// fields and no conjugation (calcJacobianKernel, equation_kernel.cu:197-245); one block
// per (element, job), the column written once
constexpr int kCJacMaxRows = 4096;

__global__ void __launch_bounds__(256) k_c_jacobian(const CJob* __restrict__ jobs, CJacDev jd, float2* __restrict__ Jout) {
    const int e = blockIdx.x, j = blockIdx.y;
    const CJob& Jb = jobs[j];
    const int k = jd.k, D = jd.D, M = jd.M, Mp = jd.Mp, Dp = jd.Dp, R = D + M;
    const int mrows = Mp * Dp;
    __shared__ float2 phi[6][64];
    __shared__ float2 T[6][64];
    __shared__ float Kc[36];
    __shared__ int nodes[6];
    if ((int)threadIdx.x < k) nodes[threadIdx.x] = jd.elements[(size_t)e * k + threadIdx.x];
    float2 acc[kCJacMaxRows / 256];
#pragma unroll
    for (int q = 0; q < kCJacMaxRows / 256; ++q) acc[q] = make_float2(0.0f, 0.0f);
    __syncthreads();
    for (int s = 0; s < Jb.C; ++s) {
        const float al = jd.exact ? jd.alpha[s] : 0.0f;
        const float w = jd.weight[s];
        for (int i = threadIdx.x; i < k * k; i += blockDim.x)
            Kc[i] = jd.ek[(size_t)e * k * k + i] + al * jd.em[(size_t)e * k * k + i];
        for (int i = threadIdx.x; i < k * R; i += blockDim.x) {
            const int a = i / R, c = i % R;
            const float2 v = Jb.x[c_idx(Jb, s, nodes[a], jd.colmap[c])];
            const float sc = (float)jd.colscale[c];
            phi[a][c] = make_float2(sc * v.x, sc * v.y);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < k * M; i += blockDim.x) {
            const int a = i / M, m = i % M;
            float tr = 0.0f, ti = 0.0f;
            for (int b = 0; b < k; ++b) {
                tr += Kc[a * k + b] * phi[b][D + m].x;
                ti += Kc[a * k + b] * phi[b][D + m].y;
            }
            T[a][m] = make_float2(w * tr, w * ti);
        }
        __syncthreads();
        // the loop is unrolled to its compile-time bound so that acc stays in registers
#pragma unroll
        for (int q = 0; q < kCJacMaxRows / 256; ++q) {
            const int idx = threadIdx.x + 256 * q;
            const int d = idx / Mp, m = idx % Mp;
            if (idx < mrows && d < D && m < M)
                for (int a = 0; a < k; ++a) {
                    const float2 u = phi[a][d], t = T[a][m];
                    acc[q].x += u.x * t.x - u.y * t.y;
                    acc[q].y += u.x * t.y + u.y * t.x;
                }
        }
        __syncthreads();
    }
    const float2 g = Jb.gamma[e];
    const zd ds = Jb.logarithmic ? (kLn10 / 10.0) * (Jb.ref * zexp((kLn10 / 10.0) * zf(g))) : zmake(Jb.ref);
    const zd sc = -ds;
    float2* col = Jout + ((size_t)j * jd.E + e) * mrows;
#pragma unroll
    for (int q = 0; q < kCJacMaxRows / 256; ++q) {
        const int idx = threadIdx.x + 256 * q;
        if (idx < mrows) col[idx] = fz(sc * zf(acc[q]));
    }
}

// G = J J^T + shift I: 32 x 32 tiles of both triangles, FP64 sums of the FP32 products
__global__ void __launch_bounds__(256) k_c_gram(int m, int E, const float2* __restrict__ J, size_t strideJ, double shift,
                                                zd* __restrict__ G, size_t strideG) {
    const int I0 = blockIdx.x * 32, J0 = blockIdx.y * 32, b = blockIdx.z;
    const float2* A = J + b * strideJ;
    __shared__ float2 Ai[32][33];
    __shared__ float2 Aj[32][33];
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
    zd acc[4] = {zmake(0.0), zmake(0.0), zmake(0.0), zmake(0.0)};
    for (int e0 = 0; e0 < E; e0 += 32) {
        for (int i = threadIdx.x; i < 32 * 32; i += 256) {
            const int r = i % 32, ee = i / 32;
            const bool ok = e0 + ee < E;
            Ai[ee][r] = (ok && I0 + r < m) ? A[(size_t)(e0 + ee) * m + I0 + r] : make_float2(0.0f, 0.0f);
            Aj[ee][r] = (ok && J0 + r < m) ? A[(size_t)(e0 + ee) * m + J0 + r] : make_float2(0.0f, 0.0f);
        }
        __syncthreads();
        for (int ee = 0; ee < 32; ++ee) {
            const zd a = zf(Ai[ee][tx]);
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[c] += a * zf(Aj[ee][ty + 8 * c]);
        }
        __syncthreads();
    }
    zd* Gb = G + b * strideG;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int r = I0 + tx, col = J0 + ty + 8 * c;
        if (r < m && col < m) Gb[r + (size_t)col * m] = r == col ? acc[c] + zmake(shift) : acc[c];
    }
}

// y[e] = sum_i J[i, e] x[i] (one wavefront per element); gamma[e] += y[e] when given
__global__ void __launch_bounds__(256) k_c_gemv_t(int m, int E, const float2* __restrict__ J, size_t strideJ,
                                                  const zd* __restrict__ x, size_t stridex, zd* __restrict__ y,
                                                  size_t stridey, float2* __restrict__ gamma) {
    const int e = blockIdx.x * 4 + threadIdx.x / 64, lane = threadIdx.x % 64, b = blockIdx.y;
    if (e >= E) return;
    const float2* col = J + b * strideJ + (size_t)e * m;
    const zd* xv = x + b * stridex;
    zd a = zmake(0.0);
    for (int i = lane; i < m; i += 64) a += zf(col[i]) * xv[i];
    for (int off = 32; off > 0; off >>= 1) a += Scal<zd>::shfl_xor(a, off);
    if (lane == 0) {
        y[b * stridey + e] = a;
        if (gamma) {
            float2& g = gamma[(size_t)b * E + e];
            g = fz(zf(g) + a);
        }
    }
}

__global__ void __launch_bounds__(256) k_c_sub(int n, const zd* __restrict__ a, const zd* __restrict__ b, zd* __restrict__ r) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) r[i] = a[i] - b[i];
}

void launch_c_assemble(const CJob* d_jobs, const int4* d_items, int nitems, hipStream_t st) {
    if (!nitems) return;
    k_c_assemble<<<nitems, kCBlock, 0, st>>>(d_jobs, d_items);
    check_launch();
}

void launch_c_dinv(const CJob* d_jobs, const int4* d_tiles, int ntiles, hipStream_t st) {
    k_c_dinv<<<ntiles, kCBlock, 0, st>>>(d_jobs, d_tiles);
    check_launch();
}

void launch_c_init(const CJob* d_jobs, const int4* d_tiles, int ntiles, const int2* d_js, int njs, hipStream_t st) {
    k_c_init<<<ntiles, kCBlock, 0, st>>>(d_jobs, d_tiles);
    k_c_reduce<<<njs, 128, 0, st>>>(d_jobs, d_js, 0, ZF_RHO, 1, ZF_BB, 0);
    check_launch();
}

void launch_c_iter(const CJob* d_jobs, const int4* d_tiles, int ntiles, const int2* d_js, int njs, int* d_live,
                   hipStream_t st) {
    k_c_spmv<<<ntiles, kCBlock, 0, st>>>(d_jobs, d_tiles);
    k_c_reduce<<<njs, 128, 0, st>>>(d_jobs, d_js, 0, ZF_PQ, 1, ZF_PSUM, 1);
    k_c_update<<<ntiles, kCBlock, 0, st>>>(d_jobs, d_tiles);
    k_c_reduce<<<njs, 128, 0, st>>>(d_jobs, d_js, 2, ZF_RHON, 3, ZF_RR, 1);
    k_c_state<<<njs, 64, 0, st>>>(d_jobs, d_js, d_live);
    k_c_pupdate<<<ntiles, kCBlock, 0, st>>>(d_jobs, d_tiles);
    check_launch();
}

void launch_c_voltages(const CJob* d_jobs, int njobs, const CMeasDev& md, float2* U32, zd* U64, hipStream_t st) {
    k_c_voltages<<<njobs, 256, 0, st>>>(d_jobs, md, U32, U64);
    check_launch();
}

void launch_c_jacobian(const CJob* d_jobs, int njobs, const CJacDev& jd, float2* J, hipStream_t st) {
    MPF_REQUIRE(jd.Dp * jd.Mp <= kCJacMaxRows && jd.k <= 6 && jd.D + jd.M <= 64, "complex Jacobian: shape too large");
    k_c_jacobian<<<dim3(jd.E, njobs), 256, 0, st>>>(d_jobs, jd, J);
    check_launch();
}

void launch_c_gram(int m, int E, const float2* J, size_t strideJ, double shift, zd* G, size_t strideG, int batch,
                   hipStream_t st) {
    const int nt = (m + 31) / 32;
    k_c_gram<<<dim3(nt, nt, batch), 256, 0, st>>>(m, E, J, strideJ, shift, G, strideG);
    check_launch();
}

void launch_c_gemv_t(int m, int E, const float2* J, size_t strideJ, const zd* x, size_t stridex, zd* y,
                     size_t stridey, int batch, float2* gamma, hipStream_t st) {
    k_c_gemv_t<<<dim3((E + 3) / 4, batch), 256, 0, st>>>(m, E, J, strideJ, x, stridex, y, stridey, gamma);
    check_launch();
}

void launch_c_sub(int n, const zd* Um, const zd* U, zd* r, hipStream_t st) {
    if (n <= 0) return;
    k_c_sub<<<(n + 255) / 256, 256, 0, st>>>(n, Um, U, r);
    check_launch();
}

}  // namespace mpf
