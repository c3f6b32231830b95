// Mixed-precision evaluation for the staged flash's searches on gfx950.
//
// Precision split, per evaluation:
//   FP64  the cubic's coefficients, discriminant, stationary points and root brackets,
//         the roots themselves (safeguarded Newton in the brackets, warm-started from the
//         lane's previous root), the derivative-validity test and the Gibbs comparison
//         between two admissible roots (root selection);
//   FP32  a_i(T) and sqrt(a_i(T)) (prepared once per lane in FP64, then rounded), the
//         mixing sums, ln(phi) of the selected root, the tangent-plane distance, the
//         Rachford-Rice solve away from its poles and the successive-substitution steps.
// Searches run in FP32 only far from the reference's termination tests; every decision
// close to one of them is reported as ambiguous and taken in FP64 by the caller.
//
// Reference quantities restated here (same mathematics, own arithmetic):
//   a_i(T), b_i             pr76_pure.hpp:144-165
//   mixing sums S_i, a, b   pr76_mixture.hpp:124-219
//   scaled cubic            pr76_roots.hpp:50-58 (scale 1: y = Z - B)
//   ln(phi)                 pr76_phase.hpp:193-258
//   least-Gibbs root        pr76_stability.hpp:36-119
//   liquid / vapor roots    pr76_split.hpp:38-84
#pragma once

#include "pr76_device.hpp"

namespace mpmc::dev {

constexpr float kEps32 = 1.1920928955078125e-07f;

// Outcome of one mixed evaluation: ISSUE_NONE, REJECTED (input contract), or
// M_AMBIGUOUS (this evaluation cannot be taken in reduced precision).
constexpr int M_AMBIGUOUS = 200;

// Per-lane constants of one cell: a_i(T) in FP32 through sqrt(a_i), b_i, and the
// pressure / temperature groups.
template <int NC>
struct Cell32 {
    float ra[NC];   // sqrt(a_i(T))
    float b[NC];    // b_i
    double rt;      // R T
    double prt;     // p / (R T)
    float prt2;     // p / (R T)^2
};

// Loads the lane's cell constants: sqrt(a_i(T)) prepared per cell in FP64 and rounded,
// b_i from the mixture, and the pressure / temperature groups.
template <int NC>
__device__ __forceinline__ bool cell32_load(const Mix<NC>& m, const float* ra32, int64_t n, int64_t c, double p,
                                            double t, Cell32<NC>& cl) {
    bool ok = finite(p) && p > 0 && finite(t) && t > 0;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        cl.ra[i] = ra32[i * n + c];
        cl.b[i] = static_cast<float>(m.b[i]);
        ok = ok && cl.ra[i] > 0.0f;
    }
    cl.rt = m.r_gas * t;
    cl.prt = p / cl.rt;
    cl.prt2 = static_cast<float>(cl.prt / cl.rt);
    return ok && finite(cl.rt) && cl.rt > 0;
}

// FP64 cubic h(y) = y^3 + c2 y^2 + c1 y + c0 in y = Z - B with c2 = 4B - 1,
// c1 = A + 2B^2 - 4B, c0 = -2B^2 (the reference's scaled cubic at scale 1).
struct Cubic64 {
    double c2, c1, c0;
    __device__ __forceinline__ double h(double y) const { return __builtin_fma(__builtin_fma(y + c2, y, c1), y, c0); }
    __device__ __forceinline__ double dh(double y) const { return __builtin_fma(__builtin_fma(3.0, y, 2.0 * c2), y, c1); }
    __device__ __forceinline__ double mag(double y) const {
        const double ay = __builtin_fabs(y);
        return ((ay + __builtin_fabs(c2)) * ay + __builtin_fabs(c1)) * ay + __builtin_fabs(c0);
    }
};

// Reciprocal to ~1 ulp without the IEEE division sequence (Newton steps only).
__device__ __forceinline__ double fast_rcp(double d) {
    double r = __builtin_amdgcn_rcp(d);
    r = __builtin_fma(__builtin_fma(-d, r, 1.0), r, r);
    return __builtin_fma(__builtin_fma(-d, r, 1.0), r, r);
}

// Root of h in the bracket (lo, hi) where h(lo) < 0 < h(hi) (h increases through it),
// safeguarded Newton from x0. Converges monotonically from the right in the convex part
// and from the left in the concave part; bisection guards every other step.
__device__ __forceinline__ double cubic_root(const Cubic64& c, double lo, double hi, double x0) {
    double x = (x0 > lo && x0 < hi) ? x0 : 0.5 * (lo + hi);
    for (int it = 0; it < 64; ++it) {
        const double f = c.h(x);
        if (f == 0) { break; }
        if (f < 0) { lo = x; } else { hi = x; }
        double xn = x - f * fast_rcp(c.dh(x));
        if (!(xn > lo && xn < hi)) { xn = 0.5 * (lo + hi); }
        const double step = __builtin_fabs(xn - x);
        x = xn;
        if (step <= 1e-13 * x || hi - lo <= 4.0 * kEps * x) { break; }
    }
    return x;
}

// Root structure of the cubic at (A, B): 1 or 3 real roots with y > 0. Returns
// ISSUE_NONE or M_AMBIGUOUS (near a double root, where the reference's root topology
// decision needs FP64 bisection).
struct Brackets {
    int count;          // 1 or 3
    double lo0, hi0;    // bracket of the smallest root
    double lo2, hi2;    // bracket of the largest root (== smallest when count == 1)
};

__device__ __forceinline__ int cubic_brackets(double A, double B, Cubic64& c, Brackets& br) {
    c.c2 = 4.0 * B - 1.0;
    c.c1 = A + 2.0 * B * B - 4.0 * B;
    c.c0 = -2.0 * B * B;
    if (!finite(c.c1) || !(c.c0 < 0)) { return M_AMBIGUOUS; }
    const double ub = 1.0 + smax(smax(__builtin_fabs(c.c2), __builtin_fabs(c.c1)), __builtin_fabs(c.c0));
    br.count = 1;
    br.lo0 = 0.0;
    br.hi0 = ub;
    const double disc = __builtin_fma(c.c2, c.c2, -3.0 * c.c1);
    if (disc > 0) {
        const double q = -c.c2 - __builtin_copysign(__builtin_sqrt(disc), c.c2);
        double s0 = q / 3.0, s1 = c.c1 / q;
        if (s1 < s0) { const double tmp = s0; s0 = s1; s1 = tmp; }
        if (s1 > 0) {
            const double h1 = c.h(s1);
            if (__builtin_fabs(h1) <= 1e-9 * c.mag(s1)) { return M_AMBIGUOUS; }
            if (s0 > 0) {
                const double h0 = c.h(s0);
                if (__builtin_fabs(h0) <= 1e-9 * c.mag(s0)) { return M_AMBIGUOUS; }
                if (h0 > 0 && h1 < 0) {
                    br.count = 3;
                    br.hi0 = s0;
                    br.lo2 = s1;
                    br.hi2 = ub;
                    return ISSUE_NONE;
                }
                if (h0 > 0) { br.hi0 = s0; } else { br.lo0 = s1; }
            } else {
                br.lo0 = s1;
            }
        }
    }
    br.lo2 = br.lo0;
    br.hi2 = br.hi0;
    return ISSUE_NONE;
}

// derivative_valid of the reference's root record (slope ratio > 8 sqrt(eps)).
__device__ __forceinline__ bool root_dvalid(const Cubic64& c, double y) {
    const double dhs = 3.0 * y * y + 2.0 * __builtin_fabs(c.c2) * y + __builtin_fabs(c.c1);
    return dhs > 0 && __builtin_fabs(c.dh(y)) > 8.0 * 1.4901161193847656e-08 * dhs;
}

// ln(phi)_mix of root y in FP64: (Z - 1) - ln(Z - B) - A L, with L the reference's
// log1p(u)/u / den group (sum_i x_i ln phi_i, pr76_phase.hpp:193-258).
__device__ __forceinline__ double gibbs_root(double A, double B, double y) {
    const double s2 = 1.4142135623730951;
    const double den = y + (2.0 - s2) * B;
    const double u = (2.0 * s2) * (B / den);
    return ((y + B) - 1.0) - log(y) - A * (log1p_over_x(u) / den);
}

// FP32 mixing sums at composition w: S_i = sqrt(a_i) sum_j f_ij w_j sqrt(a_j), a = sum w_i S_i.
template <int NC>
__device__ __forceinline__ void mix32(const Mix<NC>& m, const Cell32<NC>& c, const float* w, float* S, float& a,
                                      float& b) {
    float u[NC];
    b = 0.0f;
#pragma unroll
    for (int j = 0; j < NC; ++j) {
        u[j] = w[j] * c.ra[j];
        b = __builtin_fmaf(w[j], c.b[j], b);
    }
    a = 0.0f;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        float s = u[i];
#pragma unroll
        for (int j = 0; j < NC; ++j) {
            if (j != i) { s = __builtin_fmaf(static_cast<float>(m.factor[i * NC + j]), u[j], s); }
        }
        S[i] = c.ra[i] * s;
        a = __builtin_fmaf(w[i], S[i], a);
    }
}

__device__ __forceinline__ float log1p_over_x32(float u) {
    if (__builtin_fabsf(u) <= 0.125f) {
        float v = 1.0f / 11.0f;
#pragma unroll
        for (int k = 9; k >= 0; --k) { v = (k % 2 == 0 ? 1.0f : -1.0f) / static_cast<float>(k + 1) + u * v; }
        return v;
    }
    return __logf(1.0f + u) / u;
}

// ln(phi) of root y in FP32 (the scalar groups are formed in FP64 and rounded).
template <int NC>
__device__ __forceinline__ void lnphi32(const Cell32<NC>& c, const float* S, float a, float b, double A, double B,
                                        double y, float* out) {
    const double s2 = 1.4142135623730951;
    const double den = y + (2.0 - s2) * B;
    const float u = static_cast<float>((2.0 * s2) * B / den);
    const float lf = log1p_over_x32(u) / static_cast<float>(den);
    const float zm1 = static_cast<float>((y + B) - 1.0);
    const float lfree = __logf(static_cast<float>(y));
    const float invb = 1.0f / b;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        const float ratio = c.b[i] * invb;
        const float coef = (2.0f * S[i] - a * ratio) * c.prt2;
        out[i] = ratio * zm1 - lfree - coef * lf;
    }
}

// Mixed stability provider: least-Gibbs mechanically admissible root at w (FP64 root
// selection), ln(phi) in FP32. `yprev` carries the lane's previous smallest / largest
// roots as Newton starts. Returns ISSUE_NONE or M_AMBIGUOUS (near-multiple roots, a
// near tie between the two admissible roots, an ill-conditioned root, or non-finite
// values: the reference's decisions there need FP64).
template <int NC>
__device__ __forceinline__ int provider32(const Mix<NC>& m, const Cell32<NC>& c, const float* w, double* yprev,
                                          float* lnphi, int& branch, double& zout) {
    float S[NC], a, b;
    mix32<NC>(m, c, w, S, a, b);
    const double A = static_cast<double>(a) * c.prt / c.rt;
    const double B = static_cast<double>(b) * c.prt;
    if (!finite(A) || !finite(B) || !(B > 0) || !(A > 0)) { return M_AMBIGUOUS; }
    Cubic64 cu;
    Brackets br;
    if (cubic_brackets(A, B, cu, br) != ISSUE_NONE) { return M_AMBIGUOUS; }
    double y;
    if (br.count == 1) {
        y = cubic_root(cu, br.lo0, br.hi0, yprev[br.lo0 > 0 ? 1 : 0]);
        yprev[0] = y;
        yprev[1] = y;
        branch = 0;
    } else {
        const double y0 = cubic_root(cu, br.lo0, br.hi0, yprev[0]);
        const double y2 = cubic_root(cu, br.lo2, br.hi2, yprev[1]);
        yprev[0] = y0;
        yprev[1] = y2;
        const double g0 = gibbs_root(A, B, y0), g2 = gibbs_root(A, B, y2);
        const double d = g2 - g0;
        if (!finite(d) || __builtin_fabs(d) <= 1e-9 * (1.0 + __builtin_fabs(g0) + __builtin_fabs(g2))) {
            return M_AMBIGUOUS;
        }
        y = d < 0 ? y2 : y0;
        branch = d < 0 ? 2 : 0;
    }
    if (!root_dvalid(cu, y)) { return M_AMBIGUOUS; }
    lnphi32<NC>(c, S, a, b, A, B, y, lnphi);
    zout = y + B;
    bool fin = true;
#pragma unroll
    for (int i = 0; i < NC; ++i) { fin = fin && __builtin_isfinite(lnphi[i]); }
    return fin ? ISSUE_NONE : M_AMBIGUOUS;
}

// Mixed side provider: smallest (upper == 0) or largest (upper == 1) admissible root.
template <int NC>
__device__ __forceinline__ int side32(const Mix<NC>& m, const Cell32<NC>& c, const float* w, int upper,
                                      double& yprev, float* lnphi, double& zout) {
    float S[NC], a, b;
    mix32<NC>(m, c, w, S, a, b);
    const double A = static_cast<double>(a) * c.prt / c.rt;
    const double B = static_cast<double>(b) * c.prt;
    if (!finite(A) || !finite(B) || !(B > 0) || !(A > 0)) { return M_AMBIGUOUS; }
    Cubic64 cu;
    Brackets br;
    if (cubic_brackets(A, B, cu, br) != ISSUE_NONE) { return M_AMBIGUOUS; }
    const double y = upper ? cubic_root(cu, br.lo2, br.hi2, yprev) : cubic_root(cu, br.lo0, br.hi0, yprev);
    yprev = y;
    if (!root_dvalid(cu, y)) { return M_AMBIGUOUS; }
    lnphi32<NC>(c, S, a, b, A, B, y, lnphi);
    zout = y + B;
    bool fin = finite(zout);
#pragma unroll
    for (int i = 0; i < NC; ++i) { fin = fin && __builtin_isfinite(lnphi[i]); }
    return fin ? ISSUE_NONE : M_AMBIGUOUS;
}

// Rachford-Rice for the FP32 split: sum_i z_i t_i / (1 + beta t_i) = 0 on (0, 1) with
// t_i = K_i - 1. Safeguarded Newton in FP32 from `guess`; within 1e-3 of either pole end
// of (0, 1) the solve is repeated in FP64 ("near its poles"). Writes the normalized
// compositions and their logarithms (formed in log space, so a trace component below
// FP32's range keeps its logarithm). Returns 0 (interior root), 1 (no interior root: the
// reference's no-interior outcome) or M_AMBIGUOUS.
template <int NC>
__device__ __forceinline__ int rr32(const float* z, const float* lk, float guess, float& beta, float* x, float* y,
                                   float* lx, float* ly) {
    float t[NC];
    float f0 = 0.0f, f1 = 0.0f, s0 = 0.0f;
    bool fin = true;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        fin = fin && __builtin_fabsf(lk[i]) <= 80.0f;
        t[i] = expm1f(lk[i]);
        f0 = __builtin_fmaf(z[i], t[i], f0);
        f1 += z[i] * (t[i] / (1.0f + t[i]));
        s0 += __builtin_fabsf(z[i] * t[i]);
    }
    if (!fin || !__builtin_isfinite(f0) || !__builtin_isfinite(f1)) { return M_AMBIGUOUS; }
    const float margin = 64.0f * kEps32 * (1.0f + s0);
    if (f0 <= -margin || f1 >= margin) { return 1; }   // F(0) <= 0 or F(1) >= 0: no interior root
    if (f0 <= margin || f1 >= -margin) { return M_AMBIGUOUS; }
    float lo = 0.0f, hi = 1.0f;
    float b = (guess > 0.0f && guess < 1.0f) ? guess : 0.5f;
    for (int it = 0; it < 40; ++it) {
        float f = 0.0f, fp = 0.0f;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            const float r = t[i] * __builtin_amdgcn_rcpf(__builtin_fmaf(b, t[i], 1.0f));
            f = __builtin_fmaf(z[i], r, f);
            fp = __builtin_fmaf(-z[i] * r, r, fp);
        }
        if (f > 0) { lo = b; } else { hi = b; }
        float bn = b - f / fp;
        if (!(bn > lo && bn < hi)) { bn = 0.5f * (lo + hi); }
        const float step = __builtin_fabsf(bn - b);
        b = bn;
        if (step <= 4.0f * kEps32 * fminf(b, 1.0f - b) || hi - lo <= 4.0f * kEps32) { break; }
    }
    // x_i = z_i / (1 + beta t_i) / sx, y_i = K_i x_i sx / sy
    float den[NC];
    if (b < 1e-3f || b > 1.0f - 1e-3f) {
        // FP64 near the ends of (0, 1), where FP32 cancellation in F limits beta
        double td[NC], bd = b, lod = 0.0, hid = 1.0;
#pragma unroll
        for (int i = 0; i < NC; ++i) { td[i] = expm1(static_cast<double>(lk[i])); }
        for (int it = 0; it < 60; ++it) {
            double f = 0.0, fp = 0.0;
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                const double r = td[i] / __builtin_fma(bd, td[i], 1.0);
                f = __builtin_fma(static_cast<double>(z[i]), r, f);
                fp -= static_cast<double>(z[i]) * r * r;
            }
            if (f > 0) { lod = bd; } else { hid = bd; }
            double bn = bd - f / fp;
            if (!(bn > lod && bn < hid)) { bn = 0.5 * (lod + hid); }
            const double step = __builtin_fabs(bn - bd);
            bd = bn;
            if (step <= 4.0 * kEps * smin(bd, 1.0 - bd) || hid - lod <= 4.0 * kEps) { break; }
        }
        if (!(bd > 0.0 && bd < 1.0)) { return M_AMBIGUOUS; }
#pragma unroll
        for (int i = 0; i < NC; ++i) { den[i] = static_cast<float>(__builtin_fma(bd, td[i], 1.0)); }
        b = static_cast<float>(bd);
    } else {
#pragma unroll
        for (int i = 0; i < NC; ++i) { den[i] = __builtin_fmaf(b, t[i], 1.0f); }
    }
    float sx = 0.0f, sy = 0.0f;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        x[i] = z[i] / den[i];
        y[i] = x[i] * (t[i] + 1.0f);
        sx += x[i];
        sy += y[i];
    }
    const float isx = 1.0f / sx, isy = 1.0f / sy, lsx = __logf(sx), lsy = __logf(sy);
    bool ok = sx > 0.0f && sy > 0.0f && __builtin_isfinite(sx) && __builtin_isfinite(sy);
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        x[i] *= isx;
        y[i] *= isy;
        if (z[i] > 0.0f) {
            lx[i] = (__logf(z[i]) - __logf(den[i])) - lsx;
            ly[i] = (lx[i] + lk[i]) + (lsx - lsy);
            ok = ok && den[i] > 0.0f && __builtin_isfinite(lx[i]) && __builtin_isfinite(ly[i]);
        } else {
            lx[i] = 0.0f;
            ly[i] = 0.0f;
        }
    }
    beta = b;
    return ok ? 0 : M_AMBIGUOUS;
}

} // namespace mpmc::dev
