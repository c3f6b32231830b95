// PR76 single-phase evaluation on gfx950: mixing, scaled-cubic bisection roots and
// ln(phi), one composition per lane. FP64 throughout; all decisions follow the
// reference's thresholds and arithmetic order (compiled with -ffp-contract=off so
// that only the explicit fma calls fuse).
//
// Reference: modules/thermodynamics/include/mpmc/thermodynamics/
//   pr76_pure.hpp:144-165, pr76_mixture.hpp:124-219, pr76_roots.hpp:46-173,
//   pr76_phase.hpp:63-79,176-258.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

// Functions on the per-evaluation hot path are inlined into their kernels by default: a
// real call saves and restores the caller's registers through scratch on every
// evaluation. MPMC_HOT_INLINE=0 builds them as calls (smaller code, fewer registers).
#ifndef MPMC_HOT_INLINE
#define MPMC_HOT_INLINE 1
#endif
#if MPMC_HOT_INLINE
#define MPMC_HOT __device__ __forceinline__
#else
#define MPMC_HOT __device__ __noinline__
#endif

namespace mpmc::dev {

constexpr double kEps = DBL_EPSILON;
constexpr double kInf = __builtin_huge_val();

// Per-mixture constants as staged in LDS by each workgroup. Prepared on the host in
// the reference's arithmetic order (pr76_pure.hpp:187-194, pr76_mixture.hpp:236-250).
template <int NC>
struct Mix {
    double sqrt_tc[NC], ac[NC], b[NC], kappa[NC];
    double factor[NC * NC];
    double r_gas;
};

// Root-resolution status (pr76_roots.hpp:16) and the StabilityPropertyIssue that the
// providers raise (pt_stability.hpp:178-181). ISSUE_NONE means success.
enum RootStatus : int { RS_SUCCESS = 0, RS_NEAR_MULTIPLE, RS_ITERATION_LIMIT, RS_UNREPRESENTABLE };
enum Issue : int { ISSUE_NONE = -1, ROOT_TOPOLOGY = 0, ROOT_ITERATION_LIMIT, ROOT_RANGE, ILL_CONDITIONED,
                   NO_ADMISSIBLE, NONFINITE, REJECTED = 100 };

struct Root {
    double z, x;
    bool dvalid;
    int slope_sign;
};
// The roots are three named members, not an array: a per-lane array accessed with a
// run-time index lives in private (scratch) memory, and the compiler turns a chain of
// selects over array elements back into such an index.
struct RootSet {
    Root r0, r1, r2;
    int count;
    int status;
};

// Root k of a set, field by field through value selects (a select between member
// addresses, or a branch per member store, is merged back into an indexed access).
__device__ __forceinline__ Root root_at(const RootSet& rs, int k) {
    Root v;
    v.z = k == 0 ? rs.r0.z : (k == 1 ? rs.r1.z : rs.r2.z);
    v.x = k == 0 ? rs.r0.x : (k == 1 ? rs.r1.x : rs.r2.x);
    v.dvalid = k == 0 ? rs.r0.dvalid : (k == 1 ? rs.r1.dvalid : rs.r2.dvalid);
    v.slope_sign = k == 0 ? rs.r0.slope_sign : (k == 1 ? rs.r1.slope_sign : rs.r2.slope_sign);
    return v;
}
__device__ __forceinline__ void put_root_member(Root& dst, bool take, const Root& v) {
    dst.z = take ? v.z : dst.z;
    dst.x = take ? v.x : dst.x;
    dst.dvalid = take ? v.dvalid : dst.dvalid;
    dst.slope_sign = take ? v.slope_sign : dst.slope_sign;
}
__device__ __forceinline__ void put_root(RootSet& rs, int k, const Root& v) {
    put_root_member(rs.r0, k == 0, v);
    put_root_member(rs.r1, k == 1, v);
    put_root_member(rs.r2, k == 2, v);
}

__device__ __forceinline__ bool finite(double v) { return __builtin_isfinite(v); }
__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : (b > a ? b : a); }
__device__ __forceinline__ double dmin(double a, double b) { return b < a ? b : a; }
// std::max semantics (returns the first argument unless the second compares greater)
__device__ __forceinline__ double smax(double a, double b) { return a < b ? b : a; }
__device__ __forceinline__ double smin(double a, double b) { return b < a ? b : a; }

// Compensated (Kahan) accumulation in the reference's order (pt_stability.hpp:254-259).
__device__ __forceinline__ void kadd(double term, double& sum, double& c) {
    const double inc = term - c;
    const double next = sum + inc;
    c = (next - sum) - inc;
    sum = next;
}

struct Cubic { double c2, c1, c0; };

__device__ __forceinline__ double poly(double y, const Cubic& c) {
    return __builtin_fma(__builtin_fma(y + c.c2, y, c.c1), y, c.c0);
}
__device__ __forceinline__ double poly_scale(double y, const Cubic& c) {
    return ((__builtin_fabs(y) + __builtin_fabs(c.c2)) * __builtin_fabs(y) + __builtin_fabs(c.c1)) *
               __builtin_fabs(y) + __builtin_fabs(c.c0);
}

// All roots with Z > B of the PR76 cubic in y = (Z-B)/s: stationary-point brackets,
// bisection to 4 eps relative width (pr76_roots.hpp:77-173). With kNewton the root
// inside each bracket is found by safeguarded Newton instead of bisection (mixed
// precision mode): brackets, the near-multiple test at the stationary points, the
// endpoint sign tests and every acceptance check are the same FP64 decisions; only the
// root value can differ from the bisection's, by a few ulps.
template <bool kNewton = false>
MPMC_HOT RootSet pr76_roots(double a, double b) {
    RootSet rs;
    rs.r0 = Root{0.0, 0.0, false, 0};
    rs.r1 = rs.r0;
    rs.r2 = rs.r0;
    rs.count = 0;
    rs.status = RS_UNREPRESENTABLE;
    const double scale = smax(smax(1.0, b), __builtin_sqrt(__builtin_fabs(a)));
    Cubic c;
    {
        const double bs = b / scale;
        const double inv = 1.0 / scale;
        c.c2 = 4.0 * bs - inv;
        c.c1 = (a / scale) / scale + 2.0 * bs * bs - 4.0 * bs * inv;
        c.c0 = ((-2.0 * bs) * bs) * inv;
    }
    if (!finite(c.c2) || !finite(c.c1) || !finite(c.c0) || !(c.c0 < 0)) { return rs; }
    const double upper = 1.0 + smax(smax(__builtin_fabs(c.c2), __builtin_fabs(c.c1)), __builtin_fabs(c.c0));
    // Interval ends 0 < e1 [< e2 [< e3]]: the stationary points inside (0, upper), then
    // upper; held in scalars (no run-time indexed array).
    bool use0 = false, use1 = false;
    double s0 = 0.0, s1 = 0.0;
    const double disc = __builtin_fma(c.c2, c.c2, -3.0 * c.c1);
    if (disc > 0) {
        const double q = -c.c2 - __builtin_copysign(__builtin_sqrt(disc), c.c2);
        s0 = q / 3.0;
        s1 = c.c1 / q;
        if (s1 < s0) { const double tmp = s0; s0 = s1; s1 = tmp; }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const double pt = k == 0 ? s0 : s1;
            if (pt > 0 && pt < upper) {
                const double v = poly(pt, c);
                const double mag = poly_scale(pt, c);
                if (!finite(v) || !(mag > 0)) { return rs; }
                if (__builtin_fabs(v) <= 64.0 * kEps * mag) { rs.status = RS_NEAR_MULTIPLE; return rs; }
                if (k == 0) { use0 = true; } else { use1 = true; }
            }
        }
    }
    const int ne = 2 + (use0 ? 1 : 0) + (use1 ? 1 : 0);
    const double e1 = use0 ? s0 : (use1 ? s1 : upper);
    const double e2 = use0 && use1 ? s1 : upper;
    double last_z = 0.0;
    for (int iv = 1; iv < ne; ++iv) {
        double left = iv == 1 ? 0.0 : (iv == 2 ? e1 : e2);
        double right = iv == 1 ? e1 : (iv == 2 ? e2 : upper);
        double fl = poly(left, c);
        const double fr = poly(right, c);
        if (!finite(fl) || !finite(fr)) { rs.count = 0; rs.status = RS_UNREPRESENTABLE; return rs; }
        if (__builtin_signbit(fl) == __builtin_signbit(fr)) { continue; }
        double y = 0;
        bool conv = false;
        if (!kNewton) {
            // left moves only to points whose value has the sign bit of fl, so that bit
            // is fixed for the whole bisection.
            // The stopping tests are combined without short-circuit (each || is a branch
            // on the wave's execution mask) and the bracket is updated by selects: the
            // same arithmetic and decisions in about two thirds of the time.
            // y is the midpoint of two finite ends, so fmax(y, DBL_MIN) is the reference's
            // max(y, DBL_MIN) (one instruction); fy is nonzero past the stop test, so its
            // sign bit is fy < 0.
            const bool lneg = __builtin_signbit(fl);
            for (int it = 0; it < 2048; ++it) {
                y = (left + right) / 2;
                const double fy = poly(y, c);
                const bool stop = (fy == 0) | (y == left) | (y == right) |
                                  (right - left <= 4.0 * kEps * __builtin_fmax(y, DBL_MIN));
                if (stop) {
                    conv = true;
                    break;
                }
                const bool to_left = (fy < 0) == lneg;
                left = to_left ? y : left;
                right = to_left ? right : y;
            }
        } else {
            y = (left + right) / 2;
            for (int it = 0; it < 200; ++it) {
                const double fy = poly(y, c);
                if (fy == 0) { conv = true; break; }
                if (__builtin_signbit(fy) == __builtin_signbit(fl)) { left = y; fl = fy; } else { right = y; }
                const double d = __builtin_fma(3.0 * y + 2.0 * c.c2, y, c.c1);
                double yn = y - fy / d;
                if (!(yn > left && yn < right)) { yn = (left + right) / 2; }
                const double step = __builtin_fabs(yn - y);
                y = yn;
                if (step <= 2.0 * kEps * smax(y, DBL_MIN) || right - left <= 4.0 * kEps * smax(y, DBL_MIN)) {
                    conv = true;
                    break;
                }
            }
        }
        if (!conv) { rs.count = 0; rs.status = RS_ITERATION_LIMIT; return rs; }
        const double x = y * scale;
        const double z = x + b;
        const double mag = poly_scale(y, c);
        const double resid = __builtin_fabs(poly(y, c)) / mag;
        const double dh = __builtin_fma(3.0 * y + 2.0 * c.c2, y, c.c1);
        const double dhs = 3.0 * y * y + 2.0 * __builtin_fabs(c.c2) * y + __builtin_fabs(c.c1);
        const double sr = dhs > 0 ? __builtin_fabs(dh) / dhs : 0.0;
        const double bz = b / z;
        const double rep = 1.0 / x;
        const double att = ((a / z) / z) / (1.0 + 2.0 * bz - bz * bz);
        const double pres = __builtin_fabs(rep - att - 1.0) / (1.0 + __builtin_fabs(rep) + __builtin_fabs(att));
        if (!finite(z) || !(z > b) || !finite(resid) || !finite(pres) || resid > 128.0 * kEps ||
            pres > 512.0 * kEps || rs.count == 3) {
            rs.count = 0; rs.status = RS_UNREPRESENTABLE; return rs;
        }
        if (rs.count != 0 && !(z > last_z)) {
            rs.count = 0; rs.status = RS_NEAR_MULTIPLE; return rs;
        }
        // 8*sqrt(eps) = 2^-23 exactly.
        Root nr;
        nr.z = z;
        nr.x = x;
        nr.dvalid = sr > 8.0 * 1.4901161193847656e-08;
        nr.slope_sign = dh > 0 ? 1 : (dh < 0 ? -1 : 0);
        put_root(rs, rs.count, nr);
        last_z = z;
        ++rs.count;
    }
    rs.status = rs.count > 0 ? RS_SUCCESS : RS_UNREPRESENTABLE;
    return rs;
}

// Mixing + A, B + roots at one composition (pr76_mixture.hpp:124-219 full chart with
// row sums; pr76_phase.hpp:184-200). Returns ISSUE_NONE, ROOT_RANGE, a mapped root
// status, or REJECTED for the reference's domain_error contract checks.
template <int NC>
struct PhaseCtx {
    double amix, bmix, A, B;
    double S[NC];
    RootSet rs;
};

template <int NC>
__device__ __forceinline__ int check_fractions(const double* w) {
    // Compensated validation of sum(w) = 1 within 64 eps (pr76_mixture.hpp:146-158). The
    // reference accumulates in long double; a double-double sum carries more precision.
    double hi = 0, lo = 0;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        if (!finite(w[i]) || w[i] < 0 || w[i] > 1) { return REJECTED; }
        const double s = hi + w[i];
        const double bb = s - hi;
        lo += (hi - (s - bb)) + (w[i] - bb);
        hi = s;
    }
    const double dev = (hi - 1.0) + lo;
    if (__builtin_fabs(dev) > 64.0 * kEps) { return REJECTED; }
    return ISSUE_NONE;
}

template <int NC, bool kNewton = false>
MPMC_HOT int phase_prepare(const Mix<NC>& m, double p, double t, const double* w, PhaseCtx<NC>& ctx) {
    ctx.rs.status = RS_UNREPRESENTABLE;
    ctx.rs.count = 0;
    if (!finite(p) || !(p > 0) || !finite(t) || !(t > 0)) { return REJECTED; }
    if (check_fractions<NC>(w) != ISSUE_NONE) { return REJECTED; }
    double root[NC];
    double att = 0, cov = 0;
    const double st = __builtin_sqrt(t);
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        const double rr = st / m.sqrt_tc[i];
        const double br = 1.0 + m.kappa[i] * (1.0 - rr);
        if (!finite(rr) || !finite(br)) { return ROOT_RANGE; }
        const double a = (br * m.ac[i]) * br;
        if (!finite(a) || (a == 0 && br != 0)) { return ROOT_RANGE; }
        att += (w[i] * a) * w[i];
        cov += w[i] * m.b[i];
        ctx.S[i] = w[i] * a;
        if (NC > 1) {
            if (!(a > 0)) { return REJECTED; }
            root[i] = __builtin_sqrt(a);
        }
    }
#pragma unroll
    for (int i = 0; i < NC; ++i) {
#pragma unroll
        for (int j = i + 1; j < NC; ++j) {
            const double f = m.factor[i * NC + j];
            const double cross = ((w[i] * root[i]) * (w[j] * root[j])) * f;
            att += cross;
            att += cross;
            const double aij = (root[i] * f) * root[j];
            ctx.S[i] += w[j] * aij;
            ctx.S[j] += w[i] * aij;
        }
    }
    if (!finite(att) || !finite(cov) || !(cov > 0)) { return ROOT_RANGE; }
    ctx.amix = att;
    ctx.bmix = cov;
    const double rt = m.r_gas * t;
    if (!finite(rt) || !(rt > 0)) { return ROOT_RANGE; }
    const double prt = p / rt;
    ctx.A = (att / rt) * prt;
    ctx.B = cov * prt;
    if (!finite(ctx.A) || !finite(ctx.B) || !(ctx.B > 0) || (ctx.A == 0 && att != 0)) { return ROOT_RANGE; }
    ctx.rs = pr76_roots<kNewton>(ctx.A, ctx.B);
    switch (ctx.rs.status) {
    case RS_NEAR_MULTIPLE: return ROOT_TOPOLOGY;
    case RS_ITERATION_LIMIT: return ROOT_ITERATION_LIMIT;
    case RS_UNREPRESENTABLE: return ROOT_RANGE;
    default: return ISSUE_NONE;
    }
}

// log1p(u)/u: degree-19 alternating series for |u| <= 1/8 (pr76_phase.hpp:67-79).
__device__ __forceinline__ double log1p_over_x(double u) {
    if (__builtin_fabs(u) <= 0.125) {
        double v = -1.0 / 20.0;
#pragma unroll
        for (int k = 18; k >= 0; --k) { v = (k % 2 == 0 ? 1.0 : -1.0) / static_cast<double>(k + 1) + u * v; }
        return v;
    }
    return log1p(u) / u;
}

// ln(phi) of the root with free volume fv = Z - B (pr76_phase.hpp:193-258, double
// arithmetic).
template <int NC>
MPMC_HOT int phase_lnphi_x(const Mix<NC>& m, double p, double t, const PhaseCtx<NC>& ctx, double fv,
                                          double* out, double* zout) {
#pragma unroll
    for (int i = 0; i < NC; ++i) { if (!finite(ctx.S[i])) { return ROOT_RANGE; } }
    const double z = fv + ctx.B;
    const double s2 = 1.4142135623730951; // sqrt(2), correctly rounded
    const double den = fv + (2.0 - s2) * ctx.B;
    const double u = (2.0 * s2) * (ctx.B / den);
    const double lf = log1p_over_x(u) / den;
    double zm1 = z - 1.0;
    double lfree = log(fv);
    if (__builtin_fabs(zm1) < 0.25 && ctx.B < 0.125) {
        const double bz = ctx.B / z;
        zm1 = ctx.B / fv - (ctx.A / z) / (1.0 + 2.0 * bz - bz * bz);
        lfree = log1p(zm1 - ctx.B);
    }
    const double rt = m.r_gas * t;
    const double prt = p / rt;
    bool ok = true;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        const double ratio = m.b[i] / ctx.bmix;
        const double coef = ((2.0 * ctx.S[i] - ctx.amix * ratio) / rt) * prt;
        out[i] = ratio * zm1 - lfree - coef * lf;
        ok = ok && finite(out[i]);
    }
    if (!ok || !finite(z)) { return ROOT_RANGE; }
    if (zout) { *zout = z; }
    return ISSUE_NONE;
}

// ln(phi) of root k of the context's root set.
template <int NC>
MPMC_HOT int phase_lnphi(const Mix<NC>& m, double p, double t, const PhaseCtx<NC>& ctx, int k,
                                        double* out, double* zout) {
    return phase_lnphi_x<NC>(m, p, t, ctx, root_at(ctx.rs, k).x, out, zout);
}

// Activity of one candidate phase: ln(phi), provider branch (root index), smooth flag.
template <int NC>
struct Activity {
    double lnphi[NC];
    int branch;
    bool smooth;
};

// Stability provider: the least-Gibbs mechanically admissible root at w, with the
// near-tie smoothness flag (pr76_stability.hpp:36-119). The admissible roots (positive
// slope) are compared in root order. A PR76 cubic with c0 < 0 has one or three roots in
// y > 0 with slopes +, -, +, so at most two are admissible; they are held in two slots
// of registers rather than a run-time indexed array. A third admissible root cannot
// occur and is reported as a root-topology issue.
template <int NC, bool kNewton = false>
MPMC_HOT int stability_provider(const Mix<NC>& m, double p, double t, const double* w,
                                               Activity<NC>& out) {
    PhaseCtx<NC> ctx;
    const int pi = phase_prepare<NC, kNewton>(m, p, t, w, ctx);
    if (pi != ISSUE_NONE) { return pi; }
    const RootSet& rs = ctx.rs;
    const bool a0 = rs.count > 0 && rs.r0.slope_sign > 0;
    const bool a1 = rs.count > 1 && rs.r1.slope_sign > 0;
    const bool a2 = rs.count > 2 && rs.r2.slope_sign > 0;
    if (!a0 && !a1 && !a2) { return NO_ADMISSIBLE; }
    if (a0 && a1 && a2) { return ROOT_TOPOLOGY; }
    // first (f) and second (g) admissible root, in root order
    const int f = a0 ? 0 : (a1 ? 1 : 2);
    const bool two = (a0 && (a1 || a2)) || (a1 && a2);
    const int g = a0 ? (a1 ? 1 : 2) : 2;
    double c0[NC], c1[NC];
    {
        const double xf = f == 0 ? rs.r0.x : (f == 1 ? rs.r1.x : rs.r2.x);
        const int e = phase_lnphi_x<NC>(m, p, t, ctx, xf, c0, nullptr);
        if (e != ISSUE_NONE) { return e; }
    }
    if (two) {
        const double xg = g == 1 ? rs.r1.x : rs.r2.x;
        const int e = phase_lnphi_x<NC>(m, p, t, ctx, xg, c1, nullptr);
        if (e != ISSUE_NONE) { return e; }
    }
    // selection: the second candidate replaces the first when its Gibbs difference is negative
    bool second = false;
    if (two) {
        double s = 0, c = 0;
#pragma unroll
        for (int i = 0; i < NC; ++i) { kadd(w[i] * (c1[i] - c0[i]), s, c); }
        if (!finite(s)) { return NONFINITE; }
        second = s < 0;
    }
    const int bidx = second ? g : f;
    const bool dvalid = bidx == 0 ? rs.r0.dvalid : (bidx == 1 ? rs.r1.dvalid : rs.r2.dvalid);
    if (!dvalid) { return ILL_CONDITIONED; }
    bool smooth = true;
    if (two) {
        // the other candidate against the selected one
        double mag = 1;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            const double other = second ? c0[i] : c1[i];
            const double best = second ? c1[i] : c0[i];
            mag += w[i] * (__builtin_fabs(other) + __builtin_fabs(best));
        }
        if (!finite(mag)) { return NONFINITE; }
        double s = 0, c = 0;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            const double other = second ? c0[i] : c1[i];
            const double best = second ? c1[i] : c0[i];
            kadd(w[i] * (other - best), s, c);
        }
        if (!finite(s)) { return NONFINITE; }
        if (__builtin_fabs(s) <= 256.0 * kEps * mag) { smooth = false; }
    }
#pragma unroll
    for (int i = 0; i < NC; ++i) { out.lnphi[i] = second ? c1[i] : c0[i]; }
    out.branch = bidx;
    out.smooth = smooth;
    return ISSUE_NONE;
}

// Lowest (upper == 0) or highest (upper == 1) admissible root: the VLE liquid/vapor
// candidate roles (pr76_split.hpp:188-234) and the three-phase root sides
// (pr76_three_phase.hpp:39-102).
template <int NC, bool kNewton = false>
MPMC_HOT int side_provider(const Mix<NC>& m, double p, double t, const double* w, int upper,
                                          Activity<NC>& out, double& zout) {
    PhaseCtx<NC> ctx;
    const int pi = phase_prepare<NC, kNewton>(m, p, t, w, ctx);
    if (pi != ISSUE_NONE) { return pi; }
    const RootSet& rs = ctx.rs;
    const bool a0 = rs.count > 0 && rs.r0.slope_sign > 0;
    const bool a1 = rs.count > 1 && rs.r1.slope_sign > 0;
    const bool a2 = rs.count > 2 && rs.r2.slope_sign > 0;
    if (!a0 && !a1 && !a2) { return NO_ADMISSIBLE; }
    const int sel = upper ? (a2 ? 2 : (a1 ? 1 : 0)) : (a0 ? 0 : (a1 ? 1 : 2));
    const bool dvalid = sel == 0 ? rs.r0.dvalid : (sel == 1 ? rs.r1.dvalid : rs.r2.dvalid);
    if (!dvalid) { return ILL_CONDITIONED; }
    const double xs = sel == 0 ? rs.r0.x : (sel == 1 ? rs.r1.x : rs.r2.x);
    const int e = phase_lnphi_x<NC>(m, p, t, ctx, xs, out.lnphi, &zout);
    if (e != ISSUE_NONE) { return e; }
    out.branch = sel;
    out.smooth = true;
    return ISSUE_NONE;
}

} // namespace mpmc::dev
