// Multilayer perceptron for the learned Gauss-Newton start: dense layers with GELU,
// trained with MSE and Adam.  No counterpart in the reference.
//
// Every GEMM runs on MFMA (v_mfma_f32_32x32x16_bf16): FP32 operands in HBM are rounded to
// BF16 (nearest even) while staged in LDS, products accumulate in FP32.  One kernel covers
// the three products of a layer through element strides:
//   forward   Z = H W (+ b)        A = H  (batch x in),  B = W   (in x out)
//   backward  dW = H^T dZ          A = H^T,              B = dZ
//             dH = dZ W^T          A = dZ,               B = W^T
// Tiles of 64 x 64 outputs per 256-thread block (4 wavefronts x 32 x 32), K staged 32 at a
// time; the staging loop walks the unit-stride dimension so loads coalesce.
#include <cmath>

#include "mlp.h"

namespace mpf {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

constexpr int kMT = 64;       // output tile
constexpr int kMK = 32;       // K step
constexpr int kMLds = kMK + 8;

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

// C[m][n] = sum_k A(m, k) B(k, n) with A(m, k) = A[m*sam + k*sak], B(k, n) = B[k*sbk + n*sbn].
// epi 0: C = acc;  1: Z = acc + bias[n] (stored in C) and H = GELU(Z);  2: C = acc + bias[n]
__global__ void __launch_bounds__(256) k_mlp_gemm(int M, int N, int K, const float* __restrict__ A, long sam,
                                                  long sak, const float* __restrict__ B, long sbk, long sbn,
                                                  float* __restrict__ C, int ldc, int epi,
                                                  const float* __restrict__ bias, float* __restrict__ H) {
    __shared__ __attribute__((aligned(16))) __bf16 sa[kMT][kMLds];   // [m][k]
    __shared__ __attribute__((aligned(16))) __bf16 sb[kMT][kMLds];   // [n][k]
    const int m0 = blockIdx.y * kMT, n0 = blockIdx.x * kMT;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x16_t acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
    for (int k0 = 0; k0 < K; k0 += kMK) {
        for (int i = tid; i < kMT * kMK; i += 256) {
            int m, k;
            if (sak == 1) { m = i / kMK; k = i % kMK; } else { m = i % kMT; k = i / kMT; }
            const int gm = m0 + m, gk = k0 + k;
            sa[m][k] = (__bf16)((gm < M && gk < K) ? A[gm * sam + gk * sak] : 0.0f);
            int n, kb;
            if (sbn == 1) { n = i % kMT; kb = i / kMT; } else { n = i / kMK; kb = i % kMK; }
            const int gn = n0 + n, gkb = k0 + kb;
            sb[n][kb] = (__bf16)((gn < N && gkb < K) ? B[gkb * sbk + gn * sbn] : 0.0f);
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < kMK; ks += 16) {
            const int kq = ks + (lane >> 5) * 8;
            const bf16x8_t a = *reinterpret_cast<const bf16x8_t*>(&sa[wm + (lane & 31)][kq]);
            const bf16x8_t b = *reinterpret_cast<const bf16x8_t*>(&sb[wn + (lane & 31)][kq]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // accumulator layout: lane -> column (lane & 31) of the 32 x 32 sub-tile (stores
    // coalesce along rows of C), register v -> row 8*(v/4) + 4*(lane/32) + v%4
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int row = m0 + wm + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        const int col = n0 + wn + (lane & 31);
        if (row < M && col < N) {
            float z = acc[v];
            if (epi >= 1) z += bias[col];
            C[(size_t)row * ldc + col] = z;
            if (epi == 1) H[(size_t)row * ldc + col] = gelu(z);
        }
    }
}

// The same product for operands whose unit-stride dimensions, sizes and pitches are
// multiples of 4 floats: each 64 x 64 tile pair is staged 64 K at a time with 128-bit
// loads along each operand's unit-stride dimension (AK: A(m, k) contiguous in k, else in
// m; BK: B(k, n) contiguous in k, else in n), converted to BF16 and written to one of two
// LDS buffers while the MFMAs consume the other.  The MFMA sequence per output (k chunks
// of 16 in order) is the one k_mlp_gemm runs, so both give the same sums.
constexpr int kGK = 64;                // K step of the staged kernel
constexpr int kGL = kGK + 8;           // LDS row: 72 BF16 (144 bytes, 16-byte aligned)
template <bool AK, bool BK>
__global__ void __launch_bounds__(256) k_mlp_gemm4(int M, int N, int K, const float* __restrict__ A, long sam,
                                                   long sak, const float* __restrict__ B, long sbk, long sbn,
                                                   float* __restrict__ C, int ldc, int epi,
                                                   const float* __restrict__ bias, float* __restrict__ H) {
    __shared__ __attribute__((aligned(16))) __bf16 sa[2][kMT][kGL];   // [m][k]
    __shared__ __attribute__((aligned(16))) __bf16 sb[2][kMT][kGL];   // [n][k]
    const int m0 = blockIdx.y * kMT, n0 = blockIdx.x * kMT;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    // a tile is 64 x 64 floats = 1024 float4: four per thread.  Contiguous in k: float4 u
    // of thread t covers row (t + 256 u) / 16 and k from 4 ((t + 256 u) % 16).  Contiguous
    // in m (or n): thread t loads rows k = 4 (t / 16) + u, u = 0..3, of the four columns
    // from 4 (t % 16), a 4 x 4 block it writes to LDS transposed, four k per 8-byte store.
    float4 ra[4], rb[4];
    auto load = [&](int k0) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (AK) {
                const int f = tid + 256 * u, r = f >> 4, c = 4 * (f & 15);   // r = m, c = k
                if (m0 + r < M && k0 + c < K) v = *(const float4*)(A + (size_t)(m0 + r) * sam + (k0 + c));
            } else {
                const int k = 4 * (tid >> 4) + u, c = 4 * (tid & 15);        // c = m
                if (k0 + k < K && m0 + c < M) v = *(const float4*)(A + (size_t)(k0 + k) * sak + (m0 + c));
            }
            ra[u] = v;
            float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
            if (BK) {
                const int f = tid + 256 * u, r = f >> 4, c = 4 * (f & 15);   // r = n, c = k
                if (n0 + r < N && k0 + c < K) w = *(const float4*)(B + (size_t)(n0 + r) * sbn + (k0 + c));
            } else {
                const int k = 4 * (tid >> 4) + u, c = 4 * (tid & 15);        // c = n
                if (k0 + k < K && n0 + c < N) w = *(const float4*)(B + (size_t)(k0 + k) * sbk + (n0 + c));
            }
            rb[u] = w;
        }
    };
    typedef __bf16 bf16x4_t __attribute__((ext_vector_type(4)));
    auto store = [&](int buf) {
        if (AK) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int f = tid + 256 * u, r = f >> 4, c = 4 * (f & 15);
                *reinterpret_cast<bf16x4_t*>(&sa[buf][r][c]) =
                    bf16x4_t{(__bf16)ra[u].x, (__bf16)ra[u].y, (__bf16)ra[u].z, (__bf16)ra[u].w};
            }
        } else {
            const int kq = 4 * (tid >> 4), c = 4 * (tid & 15);
            *reinterpret_cast<bf16x4_t*>(&sa[buf][c][kq]) =
                bf16x4_t{(__bf16)ra[0].x, (__bf16)ra[1].x, (__bf16)ra[2].x, (__bf16)ra[3].x};
            *reinterpret_cast<bf16x4_t*>(&sa[buf][c + 1][kq]) =
                bf16x4_t{(__bf16)ra[0].y, (__bf16)ra[1].y, (__bf16)ra[2].y, (__bf16)ra[3].y};
            *reinterpret_cast<bf16x4_t*>(&sa[buf][c + 2][kq]) =
                bf16x4_t{(__bf16)ra[0].z, (__bf16)ra[1].z, (__bf16)ra[2].z, (__bf16)ra[3].z};
            *reinterpret_cast<bf16x4_t*>(&sa[buf][c + 3][kq]) =
                bf16x4_t{(__bf16)ra[0].w, (__bf16)ra[1].w, (__bf16)ra[2].w, (__bf16)ra[3].w};
        }
        if (BK) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int f = tid + 256 * u, r = f >> 4, c = 4 * (f & 15);
                *reinterpret_cast<bf16x4_t*>(&sb[buf][r][c]) =
                    bf16x4_t{(__bf16)rb[u].x, (__bf16)rb[u].y, (__bf16)rb[u].z, (__bf16)rb[u].w};
            }
        } else {
            const int kq = 4 * (tid >> 4), c = 4 * (tid & 15);
            *reinterpret_cast<bf16x4_t*>(&sb[buf][c][kq]) =
                bf16x4_t{(__bf16)rb[0].x, (__bf16)rb[1].x, (__bf16)rb[2].x, (__bf16)rb[3].x};
            *reinterpret_cast<bf16x4_t*>(&sb[buf][c + 1][kq]) =
                bf16x4_t{(__bf16)rb[0].y, (__bf16)rb[1].y, (__bf16)rb[2].y, (__bf16)rb[3].y};
            *reinterpret_cast<bf16x4_t*>(&sb[buf][c + 2][kq]) =
                bf16x4_t{(__bf16)rb[0].z, (__bf16)rb[1].z, (__bf16)rb[2].z, (__bf16)rb[3].z};
            *reinterpret_cast<bf16x4_t*>(&sb[buf][c + 3][kq]) =
                bf16x4_t{(__bf16)rb[0].w, (__bf16)rb[1].w, (__bf16)rb[2].w, (__bf16)rb[3].w};
        }
    };
    f32x16_t acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
    const int nk = (K + kGK - 1) / kGK;
    load(0);
    store(0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) load((kt + 1) * kGK);
#pragma unroll
        for (int ks = 0; ks < kGK; ks += 16) {
            const int kq = ks + (lane >> 5) * 8;
            const bf16x8_t a = *reinterpret_cast<const bf16x8_t*>(&sa[buf][wm + (lane & 31)][kq]);
            const bf16x8_t b = *reinterpret_cast<const bf16x8_t*>(&sb[buf][wn + (lane & 31)][kq]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        if (kt + 1 < nk) store(buf ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int row = m0 + wm + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        const int col = n0 + wn + (lane & 31);
        if (row < M && col < N) {
            float z = acc[v];
            if (epi >= 1) z += bias[col];
            C[(size_t)row * ldc + col] = z;
            if (epi == 1) H[(size_t)row * ldc + col] = gelu(z);
        }
    }
}

// dZ = dH * GELU'(Z) elementwise (in place on dH)
__global__ void k_mlp_gelu_back(size_t n, const float* __restrict__ Z, float* __restrict__ dH) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dH[i] *= gelu_grad(Z[i]);
}

// bias gradient: db[n] = sum over rows of dZ.  A 256-thread block owns 64 columns: row
// group g (0..3) sums rows g, g + 4, ... (loads coalesced along the columns), then the
// four group sums are added in order (a fixed summation order).
__global__ void __launch_bounds__(256) k_mlp_colsum(int M, int N, const float* __restrict__ dZ,
                                                    float* __restrict__ db) {
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int n = blockIdx.x * 64 + c;
    float s = 0.0f;
    if (n < N)
#pragma unroll 8
        for (int m = g; m < M; m += 4) s += dZ[(size_t)m * N + n];
    __shared__ float red[4][64];
    red[g][c] = s;
    __syncthreads();
    if (g == 0 && n < N) db[n] = ((red[0][c] + red[1][c]) + red[2][c]) + red[3][c];
}

// MSE: dY = 2 (Y - T) / (M N); per-block partial sums of (Y - T)^2 in FP64
__global__ void __launch_bounds__(256) k_mlp_mse(size_t n, const float* __restrict__ Y, const float* __restrict__ T,
                                                 float scale, float* __restrict__ dY, double* __restrict__ part) {
    __shared__ double red[256];
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const float d = Y[i] - T[i];
        dY[i] = 2.0f * d * scale;
        s += (double)d * (double)d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// out = scale * sum of the n partials: thread t sums t, t + 256, ..., then a tree sum
__global__ void __launch_bounds__(256) k_mlp_sum(int n, const double* __restrict__ part, double scale,
                                                 double* __restrict__ out) {
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0] * scale;
}

// Adam (Kingma & Ba 2015) with bias correction folded into the step size
__global__ void k_mlp_adam(size_t n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                           float* __restrict__ v, float lr_t, float b1, float b2, float eps) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr_t * mi / (sqrtf(vi) + eps);
}

// --------------------------------------------------------------------------------
// host
// --------------------------------------------------------------------------------
static void gemm(int M, int N, int K, const float* A, long sam, long sak, const float* B, long sbk, long sbn,
                 float* C, int ldc, int epi, const float* bias, float* H, hipStream_t st) {
    dim3 grid((N + kMT - 1) / kMT, (M + kMT - 1) / kMT);
    // the staged kernel needs each operand contiguous along k or along its other index,
    // with sizes and the other stride multiples of 4 floats and 16-byte aligned bases
    const bool ak = sak == 1, bk = sbk == 1;
    const long pa = ak ? sam : sak, pb = bk ? sbn : sbk;
    const bool vec = (ak || sam == 1) && (bk || sbn == 1) && M % 4 == 0 && N % 4 == 0 && K % 4 == 0 &&
                     pa % 4 == 0 && pb % 4 == 0 && ((uintptr_t)A & 15) == 0 && ((uintptr_t)B & 15) == 0;
    if (!vec)
        k_mlp_gemm<<<grid, 256, 0, st>>>(M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, epi, bias, H);
    else if (ak && bk)
        k_mlp_gemm4<true, true><<<grid, 256, 0, st>>>(M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, epi, bias, H);
    else if (ak)
        k_mlp_gemm4<true, false><<<grid, 256, 0, st>>>(M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, epi, bias, H);
    else if (bk)
        k_mlp_gemm4<false, true><<<grid, 256, 0, st>>>(M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, epi, bias, H);
    else
        k_mlp_gemm4<false, false><<<grid, 256, 0, st>>>(M, N, K, A, sam, sak, B, sbk, sbn, C, ldc, epi, bias, H);
    MPF_HIP(hipGetLastError());
}

Mlp::Mlp(hipStream_t s, const std::vector<int>& d, uint64_t seed) : st(s), dims(d) {
    MPF_REQUIRE(dims.size() >= 2, "mlp: need at least one layer");
    for (int v : dims) MPF_REQUIRE(v >= 1 && v <= (1 << 20), "mlp: layer width out of range");
    L = (int)dims.size() - 1;
    W.resize(L); b.resize(L); gW.resize(L); gb.resize(L); mW.resize(L); vW.resize(L); mb.resize(L); vb.resize(L);
    uint64_t state = seed;
    auto next = [&]() {   // SplitMix64
        state += 0x9E3779B97F4A7C15ULL;
        uint64_t z = state;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    };
    for (int l = 0; l < L; ++l) {
        const size_t nw = (size_t)dims[l] * dims[l + 1];
        // Glorot-uniform weights, zero biases
        const double lim = std::sqrt(6.0 / (dims[l] + dims[l + 1]));
        std::vector<float> hw(nw);
        for (auto& x : hw) x = (float)(lim * (2.0 * ((next() >> 11) * (1.0 / 9007199254740992.0)) - 1.0));
        W[l].alloc(nw); W[l].upload(hw.data(), nw, st);
        b[l].alloc(dims[l + 1]); MPF_HIP(hipMemsetAsync(b[l].p, 0, b[l].bytes(), st));
        gW[l].alloc(nw); gb[l].alloc(dims[l + 1]);
        mW[l].alloc(nw); vW[l].alloc(nw); mb[l].alloc(dims[l + 1]); vb[l].alloc(dims[l + 1]);
        MPF_HIP(hipMemsetAsync(mW[l].p, 0, mW[l].bytes(), st));
        MPF_HIP(hipMemsetAsync(vW[l].p, 0, vW[l].bytes(), st));
        MPF_HIP(hipMemsetAsync(mb[l].p, 0, mb[l].bytes(), st));
        MPF_HIP(hipMemsetAsync(vb[l].p, 0, vb[l].bytes(), st));
    }
    loss.alloc(1);
    MPF_HIP(hipStreamSynchronize(st));
}

void Mlp::ensure_batch(int B) {
    if (B <= cap) return;
    cap = B;
    Z.clear(); H.clear();
    Z.resize(L); H.resize(L);
    int wmax = 0;
    for (int l = 0; l < L; ++l) {
        Z[l].alloc((size_t)B * dims[l + 1]);
        H[l].alloc((size_t)B * dims[l + 1]);
        wmax = std::max(wmax, dims[l + 1]);
    }
    wmax = std::max(wmax, dims[0]);
    dA.alloc((size_t)B * wmax);
    dB.alloc((size_t)B * wmax);
    part.alloc(1024);
}

void Mlp::forward(int B, const float* X, float* Y) {
    MPF_REQUIRE(B >= 1, "mlp: batch < 1");
    ensure_batch(B);
    const float* in = X;
    for (int l = 0; l < L; ++l) {
        const int ni = dims[l], no = dims[l + 1];
        const bool last = l == L - 1;
        gemm(B, no, ni, in, ni, 1, W[l].p, no, 1, Z[l].p, no, last ? 2 : 1, b[l].p, H[l].p, st);
        in = H[l].p;
    }
    x_in = X;
    batch = B;
    if (Y) MPF_HIP(hipMemcpyAsync(Y, Z[L - 1].p, sizeof(float) * (size_t)B * dims[L], hipMemcpyDeviceToDevice, st));
}

double Mlp::backward(const float* T) {
    MPF_REQUIRE(batch > 0 && x_in, "mlp: backward needs a forward pass first");
    const int B = batch;
    const size_t ny = (size_t)B * dims[L];
    // dZ_L = 2 (Y - T) / (B * out) into dA
    const int nb = (int)std::min<size_t>(1024, (ny + 255) / 256);
    k_mlp_mse<<<nb, 256, 0, st>>>(ny, Z[L - 1].p, T, 1.0f / (float)ny, dA.p, part.p);
    k_mlp_sum<<<1, 256, 0, st>>>(nb, part.p, 1.0 / (double)ny, loss.p);
    float* dz = dA.p;
    float* dh = dB.p;
    for (int l = L - 1; l >= 0; --l) {
        const int ni = dims[l], no = dims[l + 1];
        const float* hin = l == 0 ? x_in : H[l - 1].p;
        // dW = H_in^T dZ ; db = column sums of dZ
        gemm(ni, no, B, hin, 1, ni, dz, no, 1, gW[l].p, no, 0, nullptr, nullptr, st);
        k_mlp_colsum<<<(no + 63) / 64, 256, 0, st>>>(B, no, dz, gb[l].p);
        if (l > 0) {
            // dH_in = dZ W^T, then dZ_in = dH_in * GELU'(Z_in)
            gemm(B, ni, no, dz, no, 1, W[l].p, 1, no, dh, ni, 0, nullptr, nullptr, st);
            const size_t n = (size_t)B * ni;
            k_mlp_gelu_back<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, Z[l - 1].p, dh);
            std::swap(dz, dh);
        }
    }
    MPF_HIP(hipGetLastError());
    double h = 0.0;
    MPF_HIP(hipMemcpyAsync(&h, loss.p, sizeof(double), hipMemcpyDeviceToHost, st));
    MPF_HIP(hipStreamSynchronize(st));
    return h;
}

void Mlp::adam(float lr, float b1, float b2, float eps) {
    ++t;
    const float lr_t = lr * (float)(std::sqrt(1.0 - std::pow((double)b2, (double)t)) / (1.0 - std::pow((double)b1, (double)t)));
    for (int l = 0; l < L; ++l) {
        const size_t nw = (size_t)dims[l] * dims[l + 1];
        k_mlp_adam<<<(unsigned)((nw + 255) / 256), 256, 0, st>>>(nw, W[l].p, gW[l].p, mW[l].p, vW[l].p, lr_t, b1, b2, eps);
        k_mlp_adam<<<(dims[l + 1] + 255) / 256, 256, 0, st>>>(dims[l + 1], b[l].p, gb[l].p, mb[l].p, vb[l].p, lr_t,
                                                             b1, b2, eps);
    }
    MPF_HIP(hipGetLastError());
}

int64_t Mlp::count() const {
    int64_t n = 0;
    for (int l = 0; l < L; ++l) n += (int64_t)dims[l] * dims[l + 1] + dims[l + 1];
    return n;
}

// flat order: W_0, b_0, W_1, b_1, ... (W row-major in x out)
void Mlp::params(int which, bool set, float* host) {
    MPF_REQUIRE(which == 0 || which == 1, "mlp: which must be 0 (parameters) or 1 (gradients)");
    MPF_REQUIRE(!(set && which == 1), "mlp: gradients are read-only");
    size_t off = 0;
    for (int l = 0; l < L; ++l) {
        const size_t nw = (size_t)dims[l] * dims[l + 1];
        DBuf<float>& w = which == 0 ? W[l] : gW[l];
        DBuf<float>& bb = which == 0 ? b[l] : gb[l];
        if (set) {
            w.upload(host + off, nw, st);
            bb.upload(host + off + nw, dims[l + 1], st);
        } else {
            MPF_HIP(hipMemcpyAsync(host + off, w.p, nw * sizeof(float), hipMemcpyDeviceToHost, st));
            MPF_HIP(hipMemcpyAsync(host + off + nw, bb.p, dims[l + 1] * sizeof(float), hipMemcpyDeviceToHost, st));
        }
        off += nw + dims[l + 1];
    }
    MPF_HIP(hipStreamSynchronize(st));
}

}  // namespace mpf
