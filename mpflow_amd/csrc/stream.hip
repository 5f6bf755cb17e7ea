// HBM streaming reference for the rooflines: grid-stride 128-bit kernels with the access
// patterns of the Krylov vector kernels, timed with HIP events on the context's stream.
#include <algorithm>

#include "numeric.h"

namespace mpf {

// PATTERN 0: y = x (one read, one write).  PATTERN 1: y = x + 0.5 y (two reads, one
// write: the update kernels' pattern).  U float4 per thread per iteration in flight.
template <int PATTERN, int U>
__global__ void __launch_bounds__(256) k_stream(const float4* __restrict__ x, float4* __restrict__ y, size_t n4) {
    const size_t stride = (size_t)gridDim.x * 256 * U;
    for (size_t base = (size_t)blockIdx.x * 256 * U + threadIdx.x; base < n4; base += stride) {
        float4 a[U], b[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t i = base + (size_t)u * 256;
            if (i < n4) {
                a[u] = x[i];
                if (PATTERN == 1) b[u] = y[i];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t i = base + (size_t)u * 256;
            if (i < n4) {
                float4 v = a[u];
                if (PATTERN == 1) {
                    v.x += 0.5f * b[u].x; v.y += 0.5f * b[u].y;
                    v.z += 0.5f * b[u].z; v.w += 0.5f * b[u].w;
                }
                y[i] = v;
            }
        }
    }
}

template <int PATTERN>
static void stream_launch(int U, int blocks, const float4* x, float4* y, size_t n4, hipStream_t st) {
    switch (U) {
        case 1: k_stream<PATTERN, 1><<<blocks, 256, 0, st>>>(x, y, n4); break;
        case 2: k_stream<PATTERN, 2><<<blocks, 256, 0, st>>>(x, y, n4); break;
        default: k_stream<PATTERN, 4><<<blocks, 256, 0, st>>>(x, y, n4); break;
    }
}

double stream_probe(int pattern, int loads, const float* x, float* y, size_t n, int blocks, int reps,
                    hipStream_t st) {
    MPF_REQUIRE(pattern == 0 || pattern == 1, "stream probe: pattern 0 (copy) or 1 (two reads, one write)");
    MPF_REQUIRE(loads == 1 || loads == 2 || loads == 4, "stream probe: 1, 2 or 4 loads in flight");
    MPF_REQUIRE(n >= 4 && n % 4 == 0 && blocks >= 1 && reps >= 1, "stream probe: n a positive multiple of 4");
    MPF_REQUIRE(((uintptr_t)x | (uintptr_t)y) % 16 == 0, "stream probe: 16-byte aligned buffers");
    const size_t n4 = n / 4;
    hipEvent_t e0, e1;
    MPF_HIP(hipEventCreate(&e0));
    MPF_HIP(hipEventCreate(&e1));
    float best = 0.0f;
    for (int r = 0; r <= reps; ++r) {   // the first launch is a warm-up
        MPF_HIP(hipEventRecord(e0, st));
        if (pattern == 0) stream_launch<0>(loads, blocks, (const float4*)x, (float4*)y, n4, st);
        else stream_launch<1>(loads, blocks, (const float4*)x, (float4*)y, n4, st);
        MPF_HIP(hipGetLastError());
        MPF_HIP(hipEventRecord(e1, st));
        MPF_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        MPF_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (r > 0 && (best == 0.0f || ms < best)) best = ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const double bytes = (double)n * sizeof(float) * (pattern == 0 ? 2.0 : 3.0);
    return bytes / (best * 1e-3) / 1e9;
}

}  // namespace mpf
