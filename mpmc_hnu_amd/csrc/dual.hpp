// Forward-mode dual numbers for gfx950: a value and N directional derivatives held in
// registers of one lane. The derivative rules are those of the reference's Dual<T,N>
// (modules/ad/include/mpmc/ad/dual.hpp:139-183 sums, products and quotients;
// modules/ad/include/mpmc/ad/math.hpp:200-214 log, :258-272 log1p, :277-291 sqrt).
//
// The reference runs ceil(inputs / 4) passes of Dual<double, 4>, each re-evaluating the
// whole primal (ad/runtime_differentiate.hpp:87-168). Here one pass carries every
// direction of the chart at once (N = nc + 1 for (p, T, x_0..x_{nc-2})), so the primal,
// the cubic root solve included, runs once per phase. Domain conditions the reference
// turns into std::domain_error are checked by the callers before each operation.
#pragma once

#include <hip/hip_runtime.h>

namespace mpmc::dev {

template <int N>
struct Dual {
    double v;
    double d[N];

    __device__ __forceinline__ Dual() = default;
    __device__ __forceinline__ explicit Dual(double value) : v(value) {
#pragma unroll
        for (int k = 0; k < N; ++k) { d[k] = 0.0; }
    }
    // value with a unit derivative in direction `dir` (dir < 0: a constant)
    __device__ __forceinline__ static Dual seed(double value, int dir) {
        Dual r(value);
#pragma unroll
        for (int k = 0; k < N; ++k) { r.d[k] = k == dir ? 1.0 : 0.0; }
        return r;
    }
    __device__ __forceinline__ bool finite() const {
        bool ok = __builtin_isfinite(v);
#pragma unroll
        for (int k = 0; k < N; ++k) { ok = ok && __builtin_isfinite(d[k]); }
        return ok;
    }
};

template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v + b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) { r.d[k] = a.d[k] + b.d[k]; }
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    r.v = a.v - b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) { r.d[k] = a.d[k] - b.d[k]; }
    return r;
}
// product rule in the reference's operand order: d = da * b + a * db
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
#pragma unroll
    for (int k = 0; k < N; ++k) { r.d[k] = a.d[k] * b.v + a.v * b.d[k]; }
    r.v = a.v * b.v;
    return r;
}
// quotient rule: q = a / b, d = (da - q * db) / b
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    const double q = a.v / b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) { r.d[k] = (a.d[k] - q * b.d[k]) / b.v; }
    r.v = q;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator+(Dual<N> a, double s) { a.v += s; return a; }
template <int N>
__device__ __forceinline__ Dual<N> operator+(double s, const Dual<N>& a) { return Dual<N>(s) + a; }
template <int N>
__device__ __forceinline__ Dual<N> operator-(Dual<N> a, double s) { a.v -= s; return a; }
template <int N>
__device__ __forceinline__ Dual<N> operator-(double s, const Dual<N>& a) { return Dual<N>(s) - a; }
template <int N>
__device__ __forceinline__ Dual<N> operator*(Dual<N> a, double s) {
    a.v *= s;
#pragma unroll
    for (int k = 0; k < N; ++k) { a.d[k] *= s; }
    return a;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(double s, Dual<N> a) {
    a.v = s * a.v;
#pragma unroll
    for (int k = 0; k < N; ++k) { a.d[k] = s * a.d[k]; }
    return a;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(Dual<N> a, double s) {
    a.v /= s;
#pragma unroll
    for (int k = 0; k < N; ++k) { a.d[k] /= s; }
    return a;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(double s, const Dual<N>& a) { return Dual<N>(s) / a; }

template <int N>
__device__ __forceinline__ Dual<N> dlog(const Dual<N>& x) {
    Dual<N> r;
    r.v = log(x.v);
#pragma unroll
    for (int k = 0; k < N; ++k) { r.d[k] = x.d[k] / x.v; }
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> dlog1p(const Dual<N>& x) {
    Dual<N> r;
    r.v = log1p(x.v);
    const double den = 1.0 + x.v;
#pragma unroll
    for (int k = 0; k < N; ++k) { r.d[k] = x.d[k] / den; }
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> dsqrt(const Dual<N>& x) {
    Dual<N> r;
    r.v = __builtin_sqrt(x.v);
#pragma unroll
    for (int k = 0; k < N; ++k) { r.d[k] = (x.d[k] / r.v) / 2.0; }
    return r;
}

} // namespace mpmc::dev
