// Complex scalars for the complex-conductivity path: `zd` (FP64) for sums, scalars and
// the dense normal equations, float2 (FP32 re, im) for stored vectors and matrices.
// A struct of its own rather than HIP's double2, whose operators act element-wise.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace mpf {

struct zd {
    double x, y;
};

__host__ __device__ inline zd zmake(double re, double im = 0.0) { return zd{re, im}; }
__host__ __device__ inline zd operator+(zd a, zd b) { return zd{a.x + b.x, a.y + b.y}; }
__host__ __device__ inline zd operator-(zd a, zd b) { return zd{a.x - b.x, a.y - b.y}; }
__host__ __device__ inline zd operator-(zd a) { return zd{-a.x, -a.y}; }
__host__ __device__ inline zd operator*(zd a, zd b) { return zd{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__host__ __device__ inline zd operator*(double s, zd a) { return zd{s * a.x, s * a.y}; }
__host__ __device__ inline zd& operator+=(zd& a, zd b) { a.x += b.x; a.y += b.y; return a; }
__host__ __device__ inline zd& operator-=(zd& a, zd b) { a.x -= b.x; a.y -= b.y; return a; }
__host__ __device__ inline zd& operator*=(zd& a, zd b) { a = a * b; return a; }
__host__ __device__ inline zd zconj(zd a) { return zd{a.x, -a.y}; }
__host__ __device__ inline double zabs2(zd a) { return a.x * a.x + a.y * a.y; }
__host__ __device__ inline bool zzero(zd a) { return a.x == 0.0 && a.y == 0.0; }
// a / b by Smith's algorithm (no overflow of |b|^2 for large b)
__host__ __device__ inline zd operator/(zd a, zd b) {
    if (fabs(b.x) >= fabs(b.y)) {
        const double r = b.y / b.x, d = b.x + b.y * r;
        return zd{(a.x + a.y * r) / d, (a.y - a.x * r) / d};
    }
    const double r = b.x / b.y, d = b.x * r + b.y;
    return zd{(a.x * r + a.y) / d, (a.y * r - a.x) / d};
}
__host__ __device__ inline zd& operator/=(zd& a, zd b) { a = a / b; return a; }
// principal square root
__host__ __device__ inline zd zsqrt(zd a) {
    const double m = hypot(a.x, a.y);
    if (m == 0.0) return zd{0.0, 0.0};
    const double t = sqrt(0.5 * (m + fabs(a.x)));
    return a.x >= 0.0 ? zd{t, 0.5 * a.y / t} : zd{0.5 * fabs(a.y) / t, a.y >= 0.0 ? t : -t};
}
__host__ __device__ inline zd zexp(zd a) {
    const double e = exp(a.x);
    return zd{e * cos(a.y), e * sin(a.y)};
}
__host__ __device__ inline zd zf(float2 v) { return zd{(double)v.x, (double)v.y}; }
__host__ __device__ inline float2 fz(zd v) { return make_float2((float)v.x, (float)v.y); }

// Scalar traits shared by the real and complex instantiations of the dense kernels: the
// formulas of an L L^T factorisation and its triangular solves are the same for real
// symmetric positive definite and for complex symmetric (not Hermitian) matrices.
template <class T> struct Scal;
template <> struct Scal<double> {
    __device__ static double zero() { return 0.0; }
    __device__ static double one() { return 1.0; }
    __device__ static double sqrt_(double a) { return sqrt(a); }
    // a usable pivot: positive for the real factorisation
    __device__ static bool pivot_ok(double a) { return a > 0.0; }
    __device__ static double tiny() { return 1e-300; }
    __device__ static double shfl(double v, int src) { return __shfl(v, src, 64); }
    __device__ static double shfl_xor(double v, int m) { return __shfl_xor(v, m, 64); }
};
template <> struct Scal<zd> {
    __device__ static zd zero() { return zd{0.0, 0.0}; }
    __device__ static zd one() { return zd{1.0, 0.0}; }
    __device__ static zd sqrt_(zd a) { return zsqrt(a); }
    // a usable pivot: finite and nonzero (complex symmetric matrices have no sign test)
    __device__ static bool pivot_ok(zd a) { return !zzero(a) && isfinite(a.x) && isfinite(a.y); }
    __device__ static zd tiny() { return zd{1e-300, 0.0}; }
    __device__ static zd shfl(zd v, int src) { return zd{__shfl(v.x, src, 64), __shfl(v.y, src, 64)}; }
    __device__ static zd shfl_xor(zd v, int m) { return zd{__shfl_xor(v.x, m, 64), __shfl_xor(v.y, m, 64)}; }
};

}  // namespace mpf
