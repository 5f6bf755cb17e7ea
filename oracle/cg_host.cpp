// CPU baseline of the forward hot loop: multi-RHS Jacobi-PCG (optionally with the
// two-level coarse correction) on an ELL matrix in FP32 storage with FP64 scalars, the arithmetic of ConjugateGradient::solve
// (src/numeric/conjugate_gradient.cpp:61-161) with the dcFree 1 1^T term
// (conjugate_gradient_kernel.cu:34-44).  TEST INFRASTRUCTURE ONLY: used as the timed
// CPU baseline of bench.py and cross-checked against the FP64 oracle in tests/.
//
// Layout: ELL ids/vals row-major [row][width]; vectors [row][R] (RHS contiguous).
// OpenMP parallelises over rows.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

extern "C" {

int cgh_threads(void) { return omp_get_max_threads(); }

void cgh_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }

// Two-level coarse term: g = W^T r over the aggregates (aggregate t: rows astart[t] ..
// astart[t + 1] - 1), y = einv g (nt x nt, row-major, already scaled), returns
// sum_t g_t . y_t per column into gy.
static void coarse(int R, const int32_t* astart, int nt, const float* einv, const std::vector<float>& r,
                   std::vector<double>& g, std::vector<float>& y, std::vector<double>& gy) {
#pragma omp parallel for schedule(static)
    for (int t = 0; t < nt; ++t) {
        for (int c = 0; c < R; ++c) g[(size_t)t * R + c] = 0.0;
        for (int i = astart[t]; i < astart[t + 1]; ++i)
            for (int c = 0; c < R; ++c) g[(size_t)t * R + c] += r[(size_t)i * R + c];
    }
#pragma omp parallel for schedule(static)
    for (int t = 0; t < nt; ++t) {
        std::vector<float> a(R, 0.0f);
        for (int u = 0; u < nt; ++u) {
            const float e = einv[(size_t)t * nt + u];
            for (int c = 0; c < R; ++c) a[c] += e * (float)g[(size_t)u * R + c];
        }
        for (int c = 0; c < R; ++c) y[(size_t)t * R + c] = a[c];
    }
    for (int c = 0; c < R; ++c) gy[c] = 0.0;
    for (int t = 0; t < nt; ++t)
        for (int c = 0; c < R; ++c) gy[c] += g[(size_t)t * R + c] * (double)y[(size_t)t * R + c];
}

// Solves (A + dc 1 1^T) X = B for R columns.  Stops when every column has
// ||r|| <= tol ||b|| (tol > 0), or after max_iter steps.  Returns steps taken.
// einv (nullable): the two-level preconditioner's scaled coarse inverse over na
// aggregates of consecutive rows (aggregate t: rows astart[t] .. astart[t + 1] - 1,
// astart[na] = n), M^-1 r = D^-1 r + W einv W^T r.
int cgh_pcg(int n, int width, const int32_t* ids, const float* vals, int R, const float* B,
            float* X, int dc, double tol, int max_iter, const float* einv, int na, const int32_t* astart) {
    const int nt = einv ? na : 0;
    std::vector<int32_t> aof(nt ? n : 0);
    for (int t = 0; t < nt; ++t)
        for (int i = astart[t]; i < astart[t + 1]; ++i) aof[i] = t;
    std::vector<double> cg((size_t)nt * R), gy(R, 0.0);
    std::vector<float> cy((size_t)nt * R, 0.0f);
    std::vector<float> r((size_t)n * R), p((size_t)n * R), q((size_t)n * R), dinv(n);
    for (int i = 0; i < n; ++i) {
        float d = 0.0f;
        // first slot holding the diagonal (padding slots repeat the row id with value 0)
        for (int k = 0; k < width; ++k)
            if (ids[(size_t)i * width + k] == i) { d = vals[(size_t)i * width + k]; break; }
        dinv[i] = d != 0.0f ? 1.0f / d : 0.0f;
    }
    std::vector<double> rho(R), bb(R), pq(R), ps(R), rr(R), alpha(R), beta(R);
    std::memset(X, 0, sizeof(float) * (size_t)n * R);
    for (int c = 0; c < R; ++c) { rho[c] = 0; bb[c] = 0; }
#pragma omp parallel
    {
        std::vector<double> lrho(R, 0.0), lbb(R, 0.0);
#pragma omp for schedule(static)
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < R; ++c) {
                const size_t o = (size_t)i * R + c;
                const float b = B[o];
                r[o] = b;
                p[o] = dinv[i] * b;
                lrho[c] += (double)b * p[o];
                lbb[c] += (double)b * b;
            }
#pragma omp critical
        for (int c = 0; c < R; ++c) { rho[c] += lrho[c]; bb[c] += lbb[c]; }
    }
    if (nt) {
        coarse(R, astart, nt, einv, r, cg, cy, gy);
        for (int c = 0; c < R; ++c) rho[c] += gy[c];
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < R; ++c) p[(size_t)i * R + c] += cy[(size_t)aof[i] * R + c];
    }
    int step = 0;
    for (; step < max_iter; ++step) {
        for (int c = 0; c < R; ++c) { pq[c] = 0; ps[c] = 0; rr[c] = 0; }
#pragma omp parallel
        {
            std::vector<double> lpq(R, 0.0), lps(R, 0.0);
            std::vector<float> acc(R);
#pragma omp for schedule(static)
            for (int i = 0; i < n; ++i) {
                for (int c = 0; c < R; ++c) acc[c] = 0.0f;
                for (int k = 0; k < width; ++k) {
                    const float v = vals[(size_t)i * width + k];
                    const float* pr = &p[(size_t)ids[(size_t)i * width + k] * R];
                    for (int c = 0; c < R; ++c) acc[c] += v * pr[c];
                }
                for (int c = 0; c < R; ++c) {
                    const size_t o = (size_t)i * R + c;
                    q[o] = acc[c];
                    lpq[c] += (double)p[o] * acc[c];
                    lps[c] += p[o];
                }
            }
#pragma omp critical
            for (int c = 0; c < R; ++c) { pq[c] += lpq[c]; ps[c] += lps[c]; }
        }
        for (int c = 0; c < R; ++c) {
            if (dc) pq[c] += ps[c] * ps[c];
            alpha[c] = pq[c] != 0.0 ? rho[c] / pq[c] : 0.0;
            if (!dc) ps[c] = 0.0;
        }
        std::vector<double> nrho(R, 0.0);
#pragma omp parallel
        {
            std::vector<double> lrz(R, 0.0), lrr(R, 0.0);
#pragma omp for schedule(static)
            for (int i = 0; i < n; ++i)
                for (int c = 0; c < R; ++c) {
                    const size_t o = (size_t)i * R + c;
                    const float a = (float)alpha[c];
                    X[o] += a * p[o];
                    r[o] -= a * (q[o] + (float)ps[c]);
                    const float z = dinv[i] * r[o];
                    lrz[c] += (double)r[o] * z;
                    lrr[c] += (double)r[o] * r[o];
                }
#pragma omp critical
            for (int c = 0; c < R; ++c) { nrho[c] += lrz[c]; rr[c] += lrr[c]; }
        }
        if (nt) {
            coarse(R, astart, nt, einv, r, cg, cy, gy);
            for (int c = 0; c < R; ++c) nrho[c] += gy[c];
        }
        bool done = true;
        for (int c = 0; c < R; ++c) {
            if (bb[c] > 0.0 && rr[c] > tol * tol * bb[c]) done = false;
            beta[c] = rho[c] != 0.0 ? nrho[c] / rho[c] : 0.0;
            rho[c] = nrho[c];
        }
        if (tol > 0.0 && done) return step + 1;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < R; ++c) {
                const size_t o = (size_t)i * R + c;
                p[o] = dinv[i] * r[o] + (nt ? cy[(size_t)aof[i] * R + c] : 0.0f) + (float)beta[c] * p[o];
            }
    }
    return step;
}

}  // extern "C"
