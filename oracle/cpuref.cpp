// cpuref — the oracle. Test infrastructure only: an FP64 CPU restatement of the
// reference's PR76 PT flash (TPD multistart stability, bisection cubic roots,
// Rachford-Rice, logK-SSI split, three-phase SSI with generalized RR, max3
// orchestration and phase-set publication), of its forward-AD flash closure (IFT
// sensitivities of a two-phase split) and of its two-phase natural-variable TPFA
// assembly (property chart, mobility, accumulation, upwind faces, block Jacobian).
// It is loaded by tests/ and by bench.py's CPU-baseline leg and never by the product
// library.
//
// Each function cites the reference file:line it restates (paths relative to the
// reference checkout). The restatement keeps the reference's arithmetic order
// (compensated sums, fma, midpoint, log1p/expm1 branches) so that phase counts and
// labels can be compared exactly; it replaces the reference's exceptions by status
// codes and its vectors by fixed-capacity arrays (nc <= 12).
//
// Pinning: tests/test_oracle_*.py check this file against the Decimal(80) goldens the
// reference's own tests hold (tests/thermodynamics/pr76_pt/pt_test.cpp:90-150,
// tests/flash/pt_split/split_test.cpp:316-341, ternary_references.hpp,
// tests/flash/pr76_three_phase/sour_gas_references.hpp), against vectors written by
// the reference itself (tests/golden/*.npz, oracle/ref_driver.cpp; the closure and the
// assembly bit for bit), and against what the mathematics fixes (Cardano roots in long
// double, RR material balance, closure derivatives and assembly Jacobians against
// central differences, conservation of every face's fluxes between its two cells).
#include "cpuref.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <thread>
#include <type_traits>
#include <vector>

namespace cpuref {
namespace {

constexpr int NCMAX = 12;
constexpr int MAXTRIALS = 20;
constexpr double EPS = DBL_EPSILON;
constexpr double INF = std::numeric_limits<double>::infinity();
const double NaN = std::numeric_limits<double>::quiet_NaN();

// Input-contract violation that the reference raises as std::domain_error /
// invalid_argument outside the numerical-status channel. The whole cell is rejected.
struct Rejected {};

// ---------------------------------------------------------------- mixture constants
// pr76_pure.hpp:181-199 (kappa Eq.18 Horner order, ac and b split products),
// pr76_mixture.hpp:229-251 (factor 1-kij), pr76_pure.hpp:115-117 (R).
struct Mixture {
    int n;
    double sqrt_tc[NCMAX], ac[NCMAX], b[NCMAX], kappa[NCMAX];
    double factor[NCMAX * NCMAX];
};
const double R_GAS = static_cast<double>(8.31446261815324L);

Mixture make_mixture(int n, const double* tc, const double* pc, const double* omega, const double* kij) {
    Mixture m{};
    m.n = n;
    for (int i = 0; i < n; ++i) {
        m.sqrt_tc[i] = std::sqrt(tc[i]);
        m.kappa[i] = static_cast<double>(0.37464L) +
                     omega[i] * (static_cast<double>(1.54226L) - static_cast<double>(0.26992L) * omega[i]);
        const double rtc = R_GAS * tc[i];
        m.ac[i] = (static_cast<double>(0.45724L) * rtc) * (rtc / pc[i]);
        m.b[i] = (static_cast<double>(0.07780L) * rtc) / pc[i];
    }
    for (int k = 0; k < n * n; ++k) { m.factor[k] = 1.0 - kij[k]; }
    return m;
}

// Kahan step, pt_stability.hpp:254-259.
inline void kadd(double term, double& sum, double& c) {
    const double inc = term - c;
    const double next = sum + inc;
    c = (next - sum) - inc;
    sum = next;
}
inline double ksum(const double* v, int n) {
    double s = 0, c = 0;
    for (int i = 0; i < n; ++i) { kadd(v[i], s, c); }
    return s;
}
inline double midpoint(double a, double b) { return (a + b) / 2; } // |a|,|b| <= max/2 here

// ---------------------------------------------------------------- EOS phase
enum Issue { NONE = -1, ROOT_TOPOLOGY = 0, ROOT_ITERATION_LIMIT, ROOT_RANGE, ILL_CONDITIONED,
             NO_ADMISSIBLE, NONFINITE };
enum RootStatus { RS_SUCCESS = 0, RS_NEAR_MULTIPLE, RS_ITERATION_LIMIT, RS_UNREPRESENTABLE };

struct Root {
    double z, x, resid, presid, slope_ratio;
    bool dvalid;
    int slope_sign, iters;
};
struct RootSet {
    Root r[3];
    int count;
    double scale;
    int status;
};

struct Cubic { double c2, c1, c0; };
// pr76_roots.hpp:50-58
Cubic cubic(double a, double b, double s) {
    const double bs = b / s;
    const double inv = 1.0 / s;
    return {4.0 * bs - inv, (a / s) / s + 2.0 * bs * bs - 4.0 * bs * inv, ((-2.0 * bs) * bs) * inv};
}
inline double poly(double y, const Cubic& c) { return std::fma(std::fma(y + c.c2, y, c.c1), y, c.c0); }
inline double poly_scale(double y, const Cubic& c) {
    return ((std::fabs(y) + std::fabs(c.c2)) * std::fabs(y) + std::fabs(c.c1)) * std::fabs(y) + std::fabs(c.c0);
}

// pr76_roots.hpp:77-173 (bracketed bisection in y = (Z-B)/s).
RootSet roots(double a, double b) {
    RootSet rs{};
    rs.status = RS_UNREPRESENTABLE;
    rs.scale = std::max({1.0, b, std::sqrt(std::fabs(a))});
    const Cubic c = cubic(a, b, rs.scale);
    if (!std::isfinite(c.c2) || !std::isfinite(c.c1) || !std::isfinite(c.c0) || !(c.c0 < 0)) { return rs; }
    const double upper = 1.0 + std::max({std::fabs(c.c2), std::fabs(c.c1), std::fabs(c.c0)});
    double ends[4];
    int ne = 1;
    ends[0] = 0.0;
    const double disc = std::fma(c.c2, c.c2, -3.0 * c.c1);
    if (disc > 0) {
        const double q = -c.c2 - std::copysign(std::sqrt(disc), c.c2);
        double st[2] = {q / 3.0, c.c1 / q};
        if (st[1] < st[0]) { std::swap(st[0], st[1]); }
        for (double pt : st) {
            if (pt > 0 && pt < upper) {
                const double v = poly(pt, c);
                const double mag = poly_scale(pt, c);
                if (!std::isfinite(v) || !(mag > 0)) { return rs; }
                if (std::fabs(v) <= 64.0 * EPS * mag) { rs.status = RS_NEAR_MULTIPLE; return rs; }
                ends[ne++] = pt;
            }
        }
    }
    ends[ne++] = upper;
    for (int iv = 1; iv < ne; ++iv) {
        double left = ends[iv - 1], right = ends[iv];
        double fl = poly(left, c);
        const double fr = poly(right, c);
        if (!std::isfinite(fl) || !std::isfinite(fr)) { return rs; }
        if (std::signbit(fl) == std::signbit(fr)) { continue; }
        double y = 0;
        bool conv = false;
        int it = 0;
        for (; it < 2048; ++it) {
            y = midpoint(left, right);
            const double fy = poly(y, c);
            if (fy == 0 || y == left || y == right ||
                right - left <= 4.0 * EPS * std::max(y, DBL_MIN)) { conv = true; break; }
            if (std::signbit(fy) == std::signbit(fl)) { left = y; fl = fy; } else { right = y; }
        }
        if (!conv) { rs.count = 0; rs.status = RS_ITERATION_LIMIT; return rs; }
        const double x = y * rs.scale;
        const double z = x + b;
        const double mag = poly_scale(y, c);
        const double resid = std::fabs(poly(y, c)) / mag;
        const double dh = std::fma(3.0 * y + 2.0 * c.c2, y, c.c1);
        const double dhs = 3.0 * y * y + 2.0 * std::fabs(c.c2) * y + std::fabs(c.c1);
        const double sr = dhs > 0 ? std::fabs(dh) / dhs : 0.0;
        const double bz = b / z;
        const double rep = 1.0 / x;
        const double att = ((a / z) / z) / (1.0 + 2.0 * bz - bz * bz);
        const double pres = std::fabs(rep - att - 1.0) / (1.0 + std::fabs(rep) + std::fabs(att));
        if (!std::isfinite(z) || !(z > b) || !std::isfinite(resid) || !std::isfinite(pres) ||
            resid > 128.0 * EPS || pres > 512.0 * EPS || rs.count == 3) {
            rs.count = 0; rs.status = RS_UNREPRESENTABLE; return rs;
        }
        if (rs.count != 0 && !(z > rs.r[rs.count - 1].z)) {
            rs.count = 0; rs.status = RS_NEAR_MULTIPLE; return rs;
        }
        rs.r[rs.count++] = {z, x, resid, pres, sr, sr > 8.0 * std::sqrt(EPS),
                            dh > 0 ? 1 : (dh < 0 ? -1 : 0), it + 1};
    }
    rs.status = rs.count > 0 ? RS_SUCCESS : RS_UNREPRESENTABLE;
    return rs;
}

// Mixing, coefficients and roots at one composition: pr76_pure.hpp:144-165,
// pr76_mixture.hpp:124-219 (full coordinates, row sums), pr76_phase.hpp:184-200.
struct PhaseCtx {
    double amix, bmix, A, B;
    double S[NCMAX];
    RootSet rs;
};

void check_fractions(const double* w, int n) {
    long double sum = 0, comp = 0;
    for (int i = 0; i < n; ++i) {
        if (!std::isfinite(w[i]) || w[i] < 0 || w[i] > 1) { throw Rejected{}; }
        const long double inc = w[i] - comp;
        const long double next = sum + inc;
        comp = (next - sum) - inc;
        sum = next;
    }
    if (std::fabs(sum - 1.0L) > static_cast<long double>(64.0 * EPS)) { throw Rejected{}; }
}

// returns NONE, ROOT_RANGE (range_error) or the mapped root status
int prepare(const Mixture& m, double p, double t, const double* w, PhaseCtx& ctx) {
    const int n = m.n;
    ctx.rs.status = RS_UNREPRESENTABLE;
    ctx.rs.count = 0;
    if (!std::isfinite(p) || !(p > 0)) { throw Rejected{}; }
    if (!std::isfinite(t) || !(t > 0)) { throw Rejected{}; }
    check_fractions(w, n);
    double ai[NCMAX], root[NCMAX];
    double att = 0, cov = 0;
    for (int i = 0; i < n; ++i) {
        const double rr = std::sqrt(t) / m.sqrt_tc[i];
        const double br = 1.0 + m.kappa[i] * (1.0 - rr);
        if (!std::isfinite(rr) || !std::isfinite(br)) { return ROOT_RANGE; }
        const double a = (br * m.ac[i]) * br;
        if (!std::isfinite(a) || (a == 0 && br != 0)) { return ROOT_RANGE; }
        ai[i] = a;
        att += (w[i] * a) * w[i];
        cov += w[i] * m.b[i];
        ctx.S[i] = w[i] * a;
        if (n > 1) {
            if (!(a > 0)) { throw Rejected{}; }
            root[i] = std::sqrt(a);
            if (!std::isfinite(root[i])) { return ROOT_RANGE; }
        }
    }
    for (int i = 0; i < n; ++i) {
        for (int j = i + 1; j < n; ++j) {
            const double f = m.factor[i * n + j];
            const double cross = ((w[i] * root[i]) * (w[j] * root[j])) * f;
            att += cross;
            att += cross;
            const double aij = (root[i] * f) * root[j];
            ctx.S[i] += w[j] * aij;
            ctx.S[j] += w[i] * aij;
        }
    }
    if (!std::isfinite(att) || !std::isfinite(cov) || !(cov > 0)) { return ROOT_RANGE; }
    ctx.amix = att;
    ctx.bmix = cov;
    const double rt = R_GAS * t;
    if (!std::isfinite(rt) || !(rt > 0)) { return ROOT_RANGE; }
    const double prt = p / rt;
    ctx.A = (att / rt) * prt;
    ctx.B = cov * prt;
    if (!std::isfinite(ctx.A) || !std::isfinite(ctx.B) || !(ctx.B > 0) || (ctx.A == 0 && att != 0)) {
        return ROOT_RANGE;
    }
    ctx.rs = roots(ctx.A, ctx.B);
    switch (ctx.rs.status) {
    case RS_NEAR_MULTIPLE: return ROOT_TOPOLOGY;
    case RS_ITERATION_LIMIT: return ROOT_ITERATION_LIMIT;
    case RS_UNREPRESENTABLE: return ROOT_RANGE;
    default: return NONE;
    }
}

// log1p(u)/u, pr76_phase.hpp:67-79 (series of degree digits/3+2 = 19 for |u| <= 1/8).
double log1p_over_x(double u) {
    if (std::fabs(u) <= 0.125) {
        constexpr int degree = 53 / 3 + 2;
        double v = (degree % 2 == 0 ? 1.0 : -1.0) / static_cast<double>(degree + 1);
        for (int k = degree - 1; k >= 0; --k) { v = (k % 2 == 0 ? 1.0 : -1.0) / static_cast<double>(k + 1) + u * v; }
        return v;
    }
    return std::log1p(u) / u;
}

// ln(phi) of root k, pr76_phase.hpp:193-258 (double arithmetic, no IFT).
int lnphi(const Mixture& m, double p, double t, const PhaseCtx& ctx, int k, double* out, double* zout) {
    for (int i = 0; i < m.n; ++i) { if (!std::isfinite(ctx.S[i])) { return ROOT_RANGE; } }
    const Root& r = ctx.rs.r[k];
    const double fv = r.x;
    const double z = fv + ctx.B;
    const double s2 = std::sqrt(2.0);
    const double den = fv + (2.0 - s2) * ctx.B;
    const double u = (2.0 * s2) * (ctx.B / den);
    const double lf = log1p_over_x(u) / den;
    double zm1 = z - 1.0;
    double lfree = std::log(fv);
    if (std::fabs(zm1) < 0.25 && ctx.B < 0.125) {
        const double bz = ctx.B / z;
        zm1 = ctx.B / fv - (ctx.A / z) / (1.0 + 2.0 * bz - bz * bz);
        lfree = std::log1p(zm1 - ctx.B);
    }
    const double rt = R_GAS * t;
    const double prt = p / rt;
    for (int i = 0; i < m.n; ++i) {
        const double ratio = m.b[i] / ctx.bmix;
        const double coef = ((2.0 * ctx.S[i] - ctx.amix * ratio) / rt) * prt;
        out[i] = ratio * zm1 - lfree - coef * lf;
        if (!std::isfinite(out[i])) { return ROOT_RANGE; }
    }
    if (!std::isfinite(z)) { return ROOT_RANGE; }
    if (zout) { *zout = z; }
    return NONE;
}

struct Counters { uint64_t prop = 0, tpd = 0, rr = 0, rr3 = 0; };

// ---------------------------------------------------------------- providers
struct SPhase {
    double lnphi[NCMAX];
    int branch;
    bool smooth;
};
struct RPhase {
    SPhase act;
    double z;
};

// Pr76StabilityEvaluator::operator(), pr76_stability.hpp:36-119.
int stab_provider(const Mixture& m, double p, double t, const double* w, SPhase& out) {
    const int n = m.n;
    PhaseCtx ctx;
    const int pi = prepare(m, p, t, w, ctx);
    if (pi != NONE) { return pi; }
    double cand[3][NCMAX];
    int cb[3];
    int nc = 0;
    for (int k = 0; k < ctx.rs.count; ++k) {
        if (ctx.rs.r[k].slope_sign <= 0) { continue; }
        const int e = lnphi(m, p, t, ctx, k, cand[nc], nullptr);
        if (e != NONE) { return e; }
        cb[nc++] = k;
    }
    if (nc == 0) { return NO_ADMISSIBLE; }
    auto diff = [&](int a, int b, double& v) {
        double s = 0, c = 0;
        for (int i = 0; i < n; ++i) { kadd(w[i] * (cand[a][i] - cand[b][i]), s, c); }
        v = s;
        return std::isfinite(s);
    };
    int best = 0;
    for (int k = 1; k < nc; ++k) {
        double d;
        if (!diff(k, best, d)) { return NONFINITE; }
        if (d < 0) { best = k; }
    }
    if (!ctx.rs.r[cb[best]].dvalid) { return ILL_CONDITIONED; }
    bool smooth = true;
    for (int k = 0; k < nc; ++k) {
        if (k == best) { continue; }
        double mag = 1;
        for (int i = 0; i < n; ++i) { mag += w[i] * (std::fabs(cand[k][i]) + std::fabs(cand[best][i])); }
        if (!std::isfinite(mag)) { return NONFINITE; }
        double d;
        if (!diff(k, best, d)) { return NONFINITE; }
        if (std::fabs(d) <= 256.0 * EPS * mag) { smooth = false; }
    }
    for (int i = 0; i < n; ++i) { out.lnphi[i] = cand[best][i]; }
    out.branch = cb[best];
    out.smooth = smooth;
    return NONE;
}

// Candidate branch on one side of the admissible roots: Pr76VleEvaluator role call
// (pr76_split.hpp:188-234; liquid = lowest, vapor = highest admissible root) and
// Pr76ThreePhaseEvaluator (pr76_three_phase.hpp:39-102; lower/upper admissible).
int side_provider(const Mixture& m, double p, double t, const double* w, int upper, RPhase& out) {
    PhaseCtx ctx;
    const int pi = prepare(m, p, t, w, ctx);
    if (pi != NONE) { return pi; }
    int sel = -1;
    for (int k = 0; k < ctx.rs.count; ++k) {
        if (ctx.rs.r[k].slope_sign <= 0) { continue; }
        sel = k;
        if (!upper) { break; }
    }
    if (sel < 0) { return NO_ADMISSIBLE; }
    if (!ctx.rs.r[sel].dvalid) { return ILL_CONDITIONED; }
    const int e = lnphi(m, p, t, ctx, sel, out.act.lnphi, &out.z);
    if (e != NONE) { return e; }
    out.act.branch = sel;
    out.act.smooth = true;
    return NONE;
}

// ---------------------------------------------------------------- TPD / stability
struct StabOptions {
    double tpd_tol = 1e-10, stat_tol = 1e-8, armijo = 1e-4, log_step_limit = 4.0;
    int max_iterations = 512, max_backtracks = 32;
    uint64_t max_evaluations = 100000;
};
enum TrialStatus { T_NEGATIVE = 0, T_STATIONARY, T_ITERATION_LIMIT, T_EVALUATION_LIMIT,
                   T_LINE_SEARCH_FAILED, T_PROPERTY_FAILURE, T_NONSMOOTH, T_UNREPRESENTABLE };
enum StabStatus { S_UNSTABLE = 0, S_NO_INSTABILITY, S_INDETERMINATE };

struct TpdPoint {
    double w[NCMAX], res[NCMAX];
    double value, guard, stat;
    int branch;
    bool smooth;
};

double check_composition(const double* w, int n) { // pt_stability.hpp:265-277
    for (int i = 0; i < n; ++i) { if (!std::isfinite(w[i]) || w[i] < 0 || w[i] > 1) { throw Rejected{}; } }
    const double s = ksum(w, n);
    if (std::fabs(s - 1.0) > 64.0 * EPS) { throw Rejected{}; }
    return s;
}

// tangent_plane_distance, pt_stability.hpp:333-375.
int tpd(int n, const double* w, const double* z, const SPhase& trial, const SPhase& ref, TpdPoint& r) {
    const double sum = check_composition(w, n);
    (void)check_composition(z, n);
    for (int i = 0; i < n; ++i) {
        if (!std::isfinite(trial.lnphi[i]) || !std::isfinite(ref.lnphi[i])) { return NONFINITE; }
    }
    for (int i = 0; i < n; ++i) { r.w[i] = w[i]; r.res[i] = 0; }
    r.value = 0; r.stat = 0; r.branch = trial.branch; r.smooth = trial.smooth;
    double corr = 0, mag = 1, mc = 0;
    bool boundary = false;
    for (int i = 0; i < n; ++i) {
        if (z[i] == 0) { if (w[i] != 0) { throw Rejected{}; } continue; }
        if (w[i] == 0) { boundary = true; continue; }
        const double lw = std::log(w[i]), lz = std::log(z[i]);
        const double lr = std::fabs(w[i] - z[i]) < 0.5 * z[i] ? std::log1p((w[i] - z[i]) / z[i]) : lw - lz;
        const double q = lr + (trial.lnphi[i] - ref.lnphi[i]);
        r.res[i] = q;
        kadd(w[i] * q, r.value, corr);
        kadd(w[i] * (std::fabs(lw) + std::fabs(lz) + std::fabs(trial.lnphi[i]) + std::fabs(ref.lnphi[i])), mag, mc);
    }
    const double mean = r.value / sum;
    for (int i = 0; i < n; ++i) {
        if (z[i] == 0) { continue; }
        r.res[i] = w[i] == 0 ? -INF : r.res[i] - mean;
        r.stat = std::max(r.stat, std::fabs(r.res[i]));
    }
    r.guard = 256.0 * EPS * mag;
    if (!std::isfinite(r.value) || !std::isfinite(r.guard) || (!boundary && !std::isfinite(r.stat))) { return NONFINITE; }
    return NONE;
}

struct Trial {
    double init[NCMAX];
    bool has_point;
    TpdPoint pt;
    int status, iterations;
};
struct StabResult {
    int status;
    StabOptions opt;
    double feed[NCMAX];
    bool has_ref;
    SPhase ref;
    int ntrials;
    Trial trials[MAXTRIALS];
    uint64_t evaluations;
};

inline bool negative(const TpdPoint& p, const StabOptions& o) { return p.value < -o.tpd_tol - p.guard; }
// pt_stability.hpp:312-326
inline bool accept_tpd_step(const TpdPoint& pt, const TpdPoint& nx, double pred, double armijo, double rel) {
    const double g = pt.guard + nx.guard;
    if (pred <= g) { return nx.value <= pt.value + g && rel > 0 && nx.stat < pt.stat; }
    return nx.value <= pt.value - armijo * pred;
}

// detail::test_pt_stability_impl, pt_stability.hpp:383-616. imposed == nullptr is the
// feed-phase test; otherwise the common-tangent review of pt_stability.hpp:637-644.
void stability(const Mixture& m, double p, double t, const double* feed_in, const double* imposed,
               const double* const* extra, int nextra, const StabOptions& o, StabResult& R, Counters& cnt) {
    const int n = m.n;
    if (!std::isfinite(p) || p <= 0 || !std::isfinite(t) || t <= 0) { throw Rejected{}; }
    const double isum = check_composition(feed_in, n);
    if (imposed) { for (int i = 0; i < n; ++i) { if (!std::isfinite(imposed[i])) { throw Rejected{}; } } }
    int active = 0;
    for (int i = 0; i < n; ++i) { if (feed_in[i] > 0) { ++active; } }
    for (int e = 0; e < nextra; ++e) {
        (void)check_composition(extra[e], n);
        for (int i = 0; i < n; ++i) {
            if ((feed_in[i] == 0 && extra[e][i] != 0) || (feed_in[i] > 0 && extra[e][i] == 0)) { throw Rejected{}; }
        }
    }
    R.status = S_INDETERMINATE;
    R.opt = o;
    R.has_ref = false;
    R.ntrials = 0;
    R.evaluations = 0;
    for (int i = 0; i < n; ++i) { R.feed[i] = feed_in[i] / isum; }
    for (int i = 0; i < n; ++i) { if (feed_in[i] > 0 && R.feed[i] == 0) { return; } }
    auto add_start = [&](const double* v) {
        Trial& tr = R.trials[R.ntrials++];
        const double s = ksum(v, n);
        for (int i = 0; i < n; ++i) { tr.init[i] = v[i] / s; }
        tr.has_point = false; tr.status = T_EVALUATION_LIMIT; tr.iterations = 0;
    };
    add_start(R.feed);
    if (active > 1) {
        double u[NCMAX];
        for (int i = 0; i < n; ++i) { u[i] = R.feed[i] > 0 ? 1.0 / static_cast<double>(active) : 0.0; }
        add_start(u);
        for (int i = 0; i < n; ++i) {
            if (R.feed[i] == 0) { continue; }
            double s[NCMAX];
            for (int k = 0; k < n; ++k) { s[k] = R.feed[k] * 0.1; }
            s[i] += 0.9;
            add_start(s);
        }
    }
    for (int e = 0; e < nextra; ++e) { add_start(extra[e]); }
    SPhase imp;
    if (!imposed) {
        ++R.evaluations; ++cnt.prop;
        const int e = stab_provider(m, p, t, R.feed, R.ref);
        if (e != NONE) {
            for (int k = 0; k < R.ntrials; ++k) { R.trials[k].status = T_PROPERTY_FAILURE; }
            return;
        }
        R.has_ref = true;
    } else {
        for (int i = 0; i < n; ++i) { imp.lnphi[i] = R.feed[i] > 0 ? imposed[i] - std::log(R.feed[i]) : 0.0; }
        imp.branch = 0; imp.smooth = true;
        for (int i = 0; i < n; ++i) {
            if (!std::isfinite(imp.lnphi[i])) {
                for (int k = 0; k < R.ntrials; ++k) { R.trials[k].status = T_PROPERTY_FAILURE; }
                return;
            }
        }
    }
    const SPhase& dref = imposed ? imp : R.ref;
    bool neg_found = false, all_stat = true;
    for (int k = 0; k < R.ntrials; ++k) {
        Trial& tr = R.trials[k];
        auto evaluate = [&](const double* w, TpdPoint& out) -> int {
            ++R.evaluations; ++cnt.prop; ++cnt.tpd;
            SPhase ph;
            const int e = stab_provider(m, p, t, w, ph);
            if (e != NONE) { return e; }
            return tpd(n, w, R.feed, ph, dref, out);
        };
        bool lost = false;
        for (int i = 0; i < n; ++i) { lost = lost || (R.feed[i] > 0 && tr.init[i] == 0); }
        if (lost) { tr.status = T_UNREPRESENTABLE; all_stat = false; continue; }
        if (R.evaluations >= o.max_evaluations) { tr.status = T_EVALUATION_LIMIT; all_stat = false; continue; }
        if (evaluate(tr.init, tr.pt) != NONE) { tr.status = T_PROPERTY_FAILURE; all_stat = false; continue; }
        tr.has_point = true;
        for (;;) {
            if (negative(tr.pt, o)) { tr.status = T_NEGATIVE; neg_found = true; break; }
            if (!tr.pt.smooth) { tr.status = T_NONSMOOTH; break; }
            if (tr.pt.stat <= o.stat_tol) { tr.status = T_STATIONARY; break; }
            if (tr.iterations >= o.max_iterations) { tr.status = T_ITERATION_LIMIT; break; }
            const TpdPoint pt = tr.pt;
            const double scale = std::max(1.0, pt.stat / o.log_step_limit);
            double dir[NCMAX], cand[NCMAX];
            double slope = 0, sc = 0;
            for (int i = 0; i < n; ++i) {
                dir[i] = 0;
                if (R.feed[i] == 0) { continue; }
                dir[i] = -pt.res[i] / scale;
                kadd((pt.w[i] * pt.res[i]) * dir[i], slope, sc);
            }
            if (!std::isfinite(slope) || !(slope < 0)) { tr.status = T_LINE_SEARCH_FAILED; break; }
            bool accepted = false;
            double alpha = 1;
            tr.status = T_LINE_SEARCH_FAILED;
            for (int bt = 0; bt < o.max_backtracks; ++bt, alpha *= 0.5) {
                double largest = -INF;
                for (int i = 0; i < n; ++i) {
                    if (R.feed[i] > 0) {
                        cand[i] = std::log(pt.w[i]) + alpha * dir[i];
                        largest = std::max(largest, cand[i]);
                    }
                }
                for (int i = 0; i < n; ++i) { cand[i] = R.feed[i] > 0 ? std::exp(cand[i] - largest) : 0.0; }
                const double s = ksum(cand, n);
                bool positive = std::isfinite(s) && s > 0;
                for (int i = 0; i < n; ++i) {
                    cand[i] /= s;
                    positive = positive && (R.feed[i] == 0 || cand[i] > 0);
                }
                if (!positive) { continue; }
                if (R.evaluations >= o.max_evaluations) { tr.status = T_EVALUATION_LIMIT; break; }
                TpdPoint nx;
                if (evaluate(cand, nx) != NONE) { continue; }
                if (negative(nx, o) || accept_tpd_step(pt, nx, -alpha * slope, o.armijo, alpha / scale)) {
                    tr.pt = nx; accepted = true; break;
                }
            }
            if (!accepted) { break; }
            ++tr.iterations;
        }
        all_stat = all_stat && tr.status == T_STATIONARY;
    }
    R.status = neg_found ? S_UNSTABLE : (all_stat ? S_NO_INSTABILITY : S_INDETERMINATE);
}

// ---------------------------------------------------------------- Rachford-Rice
enum RRStatus { RR_INTERIOR = 0, RR_NO_INTERIOR, RR_DEGENERATE, RR_ITERATION_LIMIT, RR_UNREPRESENTABLE };
struct RRResult {
    int status;
    double beta;
    double x[NCMAX], y[NCMAX];
    double mass_abs, mass_rel;
    int iterations;
};

double rr_log_sum(int n, const double* z, const double* lk, double sign) { // rachford_rice.hpp:42-52
    double largest = -INF;
    for (int i = 0; i < n; ++i) { if (z[i] > 0) { largest = std::max(largest, std::log(z[i]) + sign * lk[i]); } }
    double s = 0, c = 0;
    for (int i = 0; i < n; ++i) { if (z[i] > 0) { kadd(std::exp(std::log(z[i]) + sign * lk[i] - largest), s, c); } }
    return largest + std::log(s);
}
double rr_value(double beta, int n, const double* z, const double* lk) { // rachford_rice.hpp:53-63
    double v = 0, c = 0;
    for (int i = 0; i < n; ++i) {
        if (z[i] == 0) { continue; }
        const double e = std::exp(-std::fabs(lk[i]));
        const double den = lk[i] >= 0 ? beta + (1 - beta) * e : (1 - beta) + beta * e;
        const double num = lk[i] >= 0 ? -std::expm1(-lk[i]) : std::expm1(lk[i]);
        kadd((z[i] / den) * num, v, c);
    }
    return v;
}

// solve_rachford_rice, rachford_rice.hpp:70-147.
void rachford_rice(int n, const double* z, const double* lk, RRResult& r, Counters& cnt) {
    (void)check_composition(z, n);
    double contrast = 0;
    for (int i = 0; i < n; ++i) {
        if (!std::isfinite(lk[i])) { throw Rejected{}; }
        if (z[i] > 0) { contrast = std::max(contrast, std::fabs(lk[i])); }
    }
    r.status = RR_UNREPRESENTABLE; r.beta = NaN; r.mass_abs = 0; r.mass_rel = 0; r.iterations = 0;
    if (contrast == 0) { r.status = RR_DEGENERATE; return; }
    const double ls = rr_log_sum(n, z, lk, 1), rs = rr_log_sum(n, z, lk, -1);
    if (!std::isfinite(ls) || !std::isfinite(rs)) { return; }
    if (ls <= 64 * EPS || rs <= 64 * EPS) { r.status = RR_NO_INTERIOR; return; }
    double left = 0, right = 1, beta = 0.5, value = 0;
    bool resolved = false;
    for (int it = 0; it < 192; ++it) {
        r.iterations = it + 1;
        ++cnt.rr;
        beta = midpoint(left, right);
        if (!(beta > 0 && beta < 1)) { break; }
        value = rr_value(beta, n, z, lk);
        if (!std::isfinite(value)) { return; }
        if (value == 0 || beta == left || beta == right || right - left <= 8 * EPS * std::min(beta, 1 - beta)) {
            resolved = true; break;
        }
        if (value > 0) { left = beta; } else { right = beta; }
    }
    if (!resolved) { r.status = RR_ITERATION_LIMIT; return; }
    r.beta = beta;
    if (std::fabs(value) > 64 * EPS) { return; }
    for (int i = 0; i < n; ++i) {
        r.x[i] = 0; r.y[i] = 0;
        if (z[i] == 0) { continue; }
        const double e = std::exp(-std::fabs(lk[i]));
        const double den = lk[i] >= 0 ? beta + (1 - beta) * e : (1 - beta) + beta * e;
        const double larger = z[i] / den;
        r.x[i] = lk[i] >= 0 ? larger * e : larger;
        r.y[i] = lk[i] >= 0 ? larger : larger * e;
        if (!std::isfinite(larger) || !(r.x[i] > 0 && r.y[i] > 0)) { return; }
    }
    const double sl = ksum(r.x, n), sv = ksum(r.y, n);
    if (std::fabs(sl - 1) > 64 * EPS || std::fabs(sv - 1) > 64 * EPS) { return; }
    for (int i = 0; i < n; ++i) {
        r.x[i] /= sl;
        r.y[i] /= sv;
        const double rec = std::fma(beta, r.y[i], (1 - beta) * r.x[i]);
        const double err = std::fabs(rec - z[i]);
        r.mass_abs = std::max(r.mass_abs, err);
        if (z[i] > 0) {
            if (!(r.x[i] > 0 && r.y[i] > 0)) { return; }
            r.mass_rel = std::max(r.mass_rel, err / z[i]);
        } else if (rec != 0) { return; }
    }
    (void)check_composition(r.x, n);
    (void)check_composition(r.y, n);
    r.status = RR_INTERIOR;
}

// ---------------------------------------------------------------- two-phase split
struct SplitOptions {
    double fug_tol = 1e-11, mass_abs = 1e-12, mass_rel = 1e-10, min_pf = 1e-10, lk_sep = 1e-7,
           z_sep = 1e-8, max_log_step = 2, res_dec = 1e-4, armijo = 1e-4;
    int max_iterations = 512, max_backtracks = 24;
    uint64_t max_evaluations = 20000;
};
enum AttemptStatus { A_CONVERGED = 0, A_NO_INTERIOR, A_RR_FAILURE, A_INDISTINGUISHABLE, A_DISAPPEARANCE,
                     A_PROPERTY_FAILURE, A_ITERATION_LIMIT, A_EVALUATION_LIMIT, A_LINE_SEARCH_FAILED,
                     A_UNREPRESENTABLE_SEED, A_BALANCE_FAILURE };
struct SplitState {
    RRResult fr;
    RPhase liq, vap;
    double lk[NCMAX], r[NCMAX], common[NCMAX];
    double norm, gibbs, guard;
};
struct SplitAttempt {
    int status;
    bool has_point;
    SplitState pt;
    uint64_t evaluations;
};

// iterate_pt_split, pt_split.hpp:274-408. Returns the attempt.
void iterate_split(const Mixture& m, double p, double t, const double* z, const double* lk0,
                   const SplitOptions& o, SplitAttempt& A, Counters& cnt) {
    const int n = m.n;
    (void)check_composition(z, n);
    A.status = A_UNREPRESENTABLE_SEED; A.has_point = false; A.evaluations = 0;
    // returns -1 on success, else an attempt status (rr failure / evaluation limit) or
    // 100 for a property failure
    auto evaluate = [&](const double* lk, SplitState& s) -> int {
        rachford_rice(n, z, lk, s.fr, cnt);
        if (s.fr.status != RR_INTERIOR) {
            return s.fr.status == RR_DEGENERATE ? A_INDISTINGUISHABLE
                 : (s.fr.status == RR_NO_INTERIOR ? A_NO_INTERIOR : A_RR_FAILURE);
        }
        for (int role = 0; role < 2; ++role) {
            if (A.evaluations >= o.max_evaluations) { return A_EVALUATION_LIMIT; }
            ++A.evaluations; ++cnt.prop;
            RPhase& ph = role == 0 ? s.liq : s.vap;
            if (side_provider(m, p, t, role == 0 ? s.fr.x : s.fr.y, role, ph) != NONE) { return 100; }
            if (!std::isfinite(ph.z) || ph.z <= 0 || !ph.act.smooth) { return 100; }
        }
        double corr = 0, mag = 1, mc = 0;
        const double beta = s.fr.beta;
        s.norm = 0; s.gibbs = 0;
        for (int i = 0; i < n; ++i) {
            s.lk[i] = 0; s.r[i] = 0; s.common[i] = 0;
            if (z[i] == 0) { continue; }
            const double lx = std::log(s.fr.x[i]), ly = std::log(s.fr.y[i]);
            const double mul = lx + s.liq.act.lnphi[i], muv = ly + s.vap.act.lnphi[i];
            s.lk[i] = ly - lx;
            s.r[i] = (lx - ly) + (s.liq.act.lnphi[i] - s.vap.act.lnphi[i]);
            s.common[i] = midpoint(mul, muv);
            s.norm = std::max(s.norm, std::fabs(s.r[i]));
            kadd((1 - beta) * s.fr.x[i] * mul + beta * s.fr.y[i] * muv, s.gibbs, corr);
            kadd((1 - beta) * s.fr.x[i] * (std::fabs(lx) + std::fabs(s.liq.act.lnphi[i])) +
                 beta * s.fr.y[i] * (std::fabs(ly) + std::fabs(s.vap.act.lnphi[i])), mag, mc);
            if (!std::isfinite(s.r[i]) || !std::isfinite(s.common[i])) { return 100; }
        }
        s.guard = 256 * EPS * mag;
        if (!std::isfinite(s.gibbs) || !std::isfinite(s.guard)) { return 100; }
        return -1;
    };
    {
        const int e = evaluate(lk0, A.pt);
        if (e != -1) { A.status = e == 100 ? A_PROPERTY_FAILURE : e; return; }
        A.has_point = true;
    }
    int iterations = 0;
    for (;;) {
        const SplitState& pt = A.pt;
        if (pt.norm <= o.fug_tol) {
            const double beta = pt.fr.beta;
            double contrast = 0;
            for (int i = 0; i < n; ++i) { contrast = std::max(contrast, std::fabs(pt.lk[i])); }
            if (!(pt.fr.mass_abs <= o.mass_abs && pt.fr.mass_rel <= o.mass_rel)) { A.status = A_BALANCE_FAILURE; }
            else if (beta <= o.min_pf || 1 - beta <= o.min_pf) { A.status = A_DISAPPEARANCE; }
            else if (contrast <= o.lk_sep ||
                     pt.vap.z - pt.liq.z <= o.z_sep * std::max(pt.vap.z, pt.liq.z)) { A.status = A_INDISTINGUISHABLE; }
            else { A.status = A_CONVERGED; }
            return;
        }
        if (iterations >= o.max_iterations) { A.status = A_ITERATION_LIMIT; return; }
        const double scale = std::max(1.0, pt.norm / o.max_log_step);
        // split_gibbs_slope, pt_split.hpp:213-231
        double slope;
        {
            double h = 0, hc = 0, f = 0, fc = 0, s2 = 0, s2c = 0;
            for (int i = 0; i < n; ++i) {
                if (z[i] == 0) { continue; }
                const double x = pt.fr.x[i], y = pt.fr.y[i];
                const double c = (x / z[i]) * y;
                const double d = y - x;
                const double r = pt.r[i];
                kadd(d * (d / z[i]), h, hc);
                kadd(c * r, f, fc);
                kadd((c * r) * r, s2, s2c);
            }
            if (!(h > 0) || !std::isfinite(h) || !std::isfinite(f) || !std::isfinite(s2) || !std::isfinite(scale) ||
                !(scale >= 1)) {
                slope = NaN;
            } else {
                const double beta = pt.fr.beta;
                slope = -(beta * (1 - beta) * s2 + (f / h) * f) / scale;
            }
        }
        if (!std::isfinite(slope) || !(slope < 0)) { A.status = A_LINE_SEARCH_FAILED; return; }
        bool accepted = false;
        double alpha = 1;
        A.status = A_LINE_SEARCH_FAILED;
        for (int bt = 0; bt < o.max_backtracks; ++bt, alpha *= 0.5) {
            double nlk[NCMAX];
            for (int i = 0; i < n; ++i) { nlk[i] = pt.lk[i] + (alpha / scale) * pt.r[i]; }
            SplitState nx;
            const int e = evaluate(nlk, nx);
            if (e == -1) {
                const double pred = -alpha * slope;
                // split_accept_step, pt_split.hpp:232-244
                const double g = pt.guard + nx.guard;
                bool ok;
                if (pred > g) { ok = nx.gibbs <= pt.gibbs - o.armijo * pred; }
                else {
                    ok = nx.gibbs <= pt.gibbs + g &&
                         (nx.norm <= o.fug_tol || pt.norm - nx.norm > o.res_dec * (alpha / scale) * pt.norm);
                }
                if (ok) { A.pt = nx; accepted = true; break; }
            } else if (e == A_EVALUATION_LIMIT) {
                A.status = A_EVALUATION_LIMIT; return;
            }
        }
        if (!accepted) { return; }
        ++iterations;
    }
}

// split_seed, pt_split.hpp:245-267
bool split_seed(int n, const double* z, const double* w, bool as_vapor, double* lk) {
    double amount = 0.1;
    for (int i = 0; i < n; ++i) { if (z[i] > 0) { amount = std::min(amount, 0.5 * (z[i] / w[i])); } }
    if (!(amount > 0 && amount < 1)) { return false; }
    double other[NCMAX];
    for (int i = 0; i < n; ++i) {
        other[i] = 0; lk[i] = 0;
        if (z[i] == 0) { continue; }
        other[i] = std::fma(-amount, w[i], z[i]) / (1 - amount);
        if (!(other[i] > 0)) { return false; }
    }
    const double s = ksum(other, n);
    if (!std::isfinite(s) || std::fabs(s - 1) > 64 * EPS) { return false; }
    for (int i = 0; i < n; ++i) {
        if (z[i] == 0) { continue; }
        other[i] /= s;
        lk[i] = (as_vapor ? 1.0 : -1.0) * (std::log(w[i]) - std::log(other[i]));
        if (!std::isfinite(lk[i])) { return false; }
    }
    return true;
}

enum VleStatus { V_SINGLE = 0, V_TWO, V_UNSTABLE, V_INDETERMINATE };
struct VleResult {
    int status;
    StabResult init;
    bool has_candidate;
    SplitAttempt cand;
    bool has_final;
    StabResult fin;
    bool any_disappearance;
};

// solve_pt_vle, pt_split.hpp:414-525 (with Pr76VleEvaluator as both providers).
// starts: extra starts of the initial stability test; final_starts (nfinal >= 0): extra
// starts of the common-tangent review (pt_split.hpp:414-512 initial_starts / final_starts),
// nfinal < 0: the initial starts again.
void solve_vle(const Mixture& m, double p, double t, const double* feed, const SplitOptions& so,
               const double* const* starts, int nstarts, VleResult& V, Counters& cnt,
               const double* const* final_starts = nullptr, int nfinal = -1) {
    const int n = m.n;
    if (!std::isfinite(p) || p <= 0 || !std::isfinite(t) || t <= 0) { throw Rejected{}; }
    (void)check_composition(feed, n);
    // final starts are validated before any property call (pt_split.hpp:440-446)
    for (int e = 0; e < nfinal; ++e) {
        (void)check_composition(final_starts[e], n);
        for (int i = 0; i < n; ++i) { if ((feed[i] > 0) != (final_starts[e][i] > 0)) { throw Rejected{}; } }
    }
    for (int e = 0; e < nstarts; ++e) {
        (void)check_composition(starts[e], n);
        for (int i = 0; i < n; ++i) { if ((feed[i] > 0) != (starts[e][i] > 0)) { throw Rejected{}; } }
    }
    V.status = V_INDETERMINATE; V.has_candidate = false; V.has_final = false; V.any_disappearance = false;
    StabOptions so_init;
    stability(m, p, t, feed, nullptr, starts, nstarts, so_init, V.init, cnt);
    if (V.init.status == S_NO_INSTABILITY) { V.status = V_SINGLE; return; }
    if (V.init.status != S_UNSTABLE) { return; }
    const double* z = V.init.feed;
    int attempts = 0;
    uint64_t split_evals = 0;
    bool limit = false;
    SplitAttempt A;
    for (int k = 0; k < V.init.ntrials && !limit; ++k) {
        const Trial& w = V.init.trials[k];
        if (w.status != T_NEGATIVE || !w.has_point || !negative(w.pt, V.init.opt)) { continue; }
        for (int side = 0; side < 2; ++side) {
            const bool as_vapor = side == 0;
            if (attempts >= 16 || split_evals >= so.max_evaluations) { limit = true; break; }
            double lk[NCMAX];
            if (split_seed(n, z, w.pt.w, as_vapor, lk)) {
                SplitOptions it = so;
                it.max_evaluations -= split_evals;
                iterate_split(m, p, t, z, lk, it, A, cnt);
            } else {
                A.status = A_UNREPRESENTABLE_SEED; A.has_point = false; A.evaluations = 0;
            }
            split_evals += A.evaluations;
            ++attempts;
            if (A.status == A_DISAPPEARANCE) { V.any_disappearance = true; }
            if (A.status == A_CONVERGED && (!V.has_candidate || A.pt.gibbs < V.cand.pt.gibbs)) {
                V.cand = A; V.has_candidate = true;
            }
        }
    }
    if (!V.has_candidate) { return; }
    const SplitState& c = V.cand.pt;
    double feed_g = 0, corr = 0, mag = 1;
    for (int i = 0; i < n; ++i) {
        if (z[i] == 0) { continue; }
        const double lz = std::log(z[i]);
        const double ph = V.init.ref.lnphi[i];
        kadd(z[i] * (lz + ph), feed_g, corr);
        mag += z[i] * (std::fabs(lz) + std::fabs(ph));
    }
    const double dg = c.gibbs - feed_g;
    const double guard = c.guard + 256 * EPS * mag;
    StabOptions fo;
    fo.tpd_tol += 0.5 * c.norm;
    if (!std::isfinite(dg) || !std::isfinite(guard) || !std::isfinite(fo.tpd_tol)) { return; }
    const double* fs[6];
    int nf = 0;
    if (nfinal < 0) {
        for (int e = 0; e < nstarts; ++e) { fs[nf++] = starts[e]; }
    } else {
        for (int e = 0; e < nfinal; ++e) { fs[nf++] = final_starts[e]; }
    }
    fs[nf++] = c.fr.x;
    fs[nf++] = c.fr.y;
    stability(m, p, t, z, c.common, fs, nf, fo, V.fin, cnt);
    V.has_final = true;
    if (V.fin.status == S_UNSTABLE) { V.status = V_UNSTABLE; }
    else if (dg > guard) { V.status = V_INDETERMINATE; }
    else if (V.fin.status == S_NO_INSTABILITY) { V.status = V_TWO; }
    else { V.status = V_INDETERMINATE; }
}

// ---------------------------------------------------------------- three phases
struct P3Options {
    double mu_tol = 1e-11, mass_abs = 1e-12, mass_rel = 1e-10, bal_tol = 2e-13, min_pf = 1e-10,
           lsep = 1e-7, max_log_step = 2.0, res_dec = 1e-4;
    int max_iterations = 512, max_backtracks = 32, max_bal_it = 192, max_bal_bt = 48;
    uint64_t max_evaluations = 30000;
};
enum P3Status { P_CONVERGED = 0, P_NO_BALANCE, P_DEGENERATE, P_BAL_ITER, P_BAL_LS, P_BAL_FAILURE,
                P_DISAPPEARANCE, P_INDISTINGUISHABLE, P_PROPERTY, P_ITERATION_LIMIT, P_EVALUATION_LIMIT,
                P_LINE_SEARCH_FAILED };
struct P3State {
    double frac[3];
    double comp[3][NCMAX];
    SPhase act[3];
    double z[3];
    double lk1[NCMAX], lk2[NCMAX], r1[NCMAX], r2[NCMAX], common[NCMAX];
    double norm, gibbs, guard, rr_res, mass_abs, mass_rel;
};
struct P3Result {
    int status;
    bool has_point;
    P3State pt;
    int disappearing;
};

struct RR3Eval {
    double f[2], j[3];
    double comp[3][NCMAX];
};
inline bool rr3_feasible(const double* b) {
    return std::isfinite(b[0]) && std::isfinite(b[1]) && b[0] >= 0 && b[1] >= 0 && b[0] + b[1] <= 1;
}
// rr3_evaluate, pt_three_phase.hpp:187-236
bool rr3_eval(int n, const double* z, const double* l1, const double* l2, const double* beta, RR3Eval& v) {
    if (!rr3_feasible(beta)) { return false; }
    const double b0 = 1.0 - beta[0] - beta[1];
    double f0 = 0, f0c = 0, f1 = 0, f1c = 0, j00 = 0, j00c = 0, j01 = 0, j01c = 0, j11 = 0, j11c = 0;
    for (int i = 0; i < n; ++i) {
        v.comp[0][i] = v.comp[1][i] = v.comp[2][i] = 0;
        if (z[i] == 0) { continue; }
        const double sh = std::max({0.0, l1[i], l2[i]});
        const double e0 = std::exp(-sh), e1 = std::exp(l1[i] - sh), e2 = std::exp(l2[i] - sh);
        const double den = b0 * e0 + beta[0] * e1 + beta[1] * e2;
        if (!std::isfinite(den) || !(den > 0)) { return false; }
        const double x0 = z[i] * e0 / den, x1 = z[i] * e1 / den, x2 = z[i] * e2 / den;
        if (!std::isfinite(x0) || !std::isfinite(x1) || !std::isfinite(x2) || !(x0 > 0) || !(x1 > 0) || !(x2 > 0)) {
            return false;
        }
        v.comp[0][i] = x0; v.comp[1][i] = x1; v.comp[2][i] = x2;
        const double d1 = x1 - x0, d2 = x2 - x0;
        kadd(d1, f0, f0c);
        kadd(d2, f1, f1c);
        kadd(-(d1 * d1) / z[i], j00, j00c);
        kadd(-(d1 * d2) / z[i], j01, j01c);
        kadd(-(d2 * d2) / z[i], j11, j11c);
    }
    v.f[0] = f0; v.f[1] = f1;
    v.j[0] = j00; v.j[1] = j01; v.j[2] = j11;
    return true;
}

enum RR3Status { G_SUCCESS = 0, G_DEGENERATE, G_ITERATION_LIMIT, G_LINE_SEARCH, G_UNREPRESENTABLE };
struct RR3Result {
    int status;
    double beta[2];
    double comp[3][NCMAX];
    double residual, mass_abs, mass_rel;
};
// solve_generalized_rr3, pt_three_phase.hpp:238-355
void rr3_solve(int n, const double* z, const double* l1, const double* l2, const double* b_init,
               const P3Options& o, RR3Result& r, Counters& cnt) {
    r.status = G_UNREPRESENTABLE; r.residual = INF; r.mass_abs = 0; r.mass_rel = 0;
    if (!rr3_feasible(b_init)) { return; }
    for (int i = 0; i < n; ++i) { if (!std::isfinite(l1[i]) || !std::isfinite(l2[i])) { return; } }
    double beta[2] = {b_init[0], b_init[1]};
    RR3Eval ev;
    for (int it = 0;; ++it) {
        if (!rr3_eval(n, z, l1, l2, beta, ev)) { return; }
        const double norm = std::max(std::fabs(ev.f[0]), std::fabs(ev.f[1]));
        if (!std::isfinite(norm)) { return; }
        if (norm <= o.bal_tol) { r.status = G_SUCCESS; r.residual = norm; break; }
        if (it >= o.max_bal_it) { r.status = G_ITERATION_LIMIT; return; }
        ++cnt.rr3;
        const double j00 = ev.j[0], j01 = ev.j[1], j11 = ev.j[2];
        const double det = j00 * j11 - j01 * j01;
        const double ds = std::fabs(j00 * j11) + std::fabs(j01 * j01) + 1.0;
        if (!std::isfinite(det) || std::fabs(det) <= 256.0 * EPS * ds) { r.status = G_DEGENERATE; return; }
        const double f0 = ev.f[0], f1 = ev.f[1];
        const double step[2] = {(-f0 * j11 + j01 * f1) / det, (j01 * f0 - j00 * f1) / det};
        if (!std::isfinite(step[0]) || !std::isfinite(step[1])) { return; }
        bool acc = false;
        double alpha = 1.0;
        for (int bt = 0; bt < o.max_bal_bt; ++bt, alpha *= 0.5) {
            const double cand[2] = {beta[0] + alpha * step[0], beta[1] + alpha * step[1]};
            if (!rr3_feasible(cand)) { continue; }
            RR3Eval nx;
            if (!rr3_eval(n, z, l1, l2, cand, nx)) { continue; }
            const double nn = std::max(std::fabs(nx.f[0]), std::fabs(nx.f[1]));
            if (!std::isfinite(nn)) { continue; }
            if (nn < norm || nn <= o.bal_tol) { beta[0] = cand[0]; beta[1] = cand[1]; acc = true; break; }
        }
        if (!acc) { r.status = G_LINE_SEARCH; return; }
    }
    if (!rr3_eval(n, z, l1, l2, beta, ev)) { r.status = G_UNREPRESENTABLE; return; }
    r.beta[0] = beta[0]; r.beta[1] = beta[1];
    double raw[3];
    for (int ph = 0; ph < 3; ++ph) {
        raw[ph] = ksum(ev.comp[ph], n);
        if (!std::isfinite(raw[ph]) || !(raw[ph] > 0)) { r.status = G_UNREPRESENTABLE; return; }
        for (int i = 0; i < n; ++i) { r.comp[ph][i] = ev.comp[ph][i] / raw[ph]; }
    }
    const double fr[3] = {1.0 - beta[0] - beta[1], beta[0], beta[1]};
    double ae = 0, re = 0;
    for (int i = 0; i < n; ++i) {
        const double rec = fr[0] * r.comp[0][i] + fr[1] * r.comp[1][i] + fr[2] * r.comp[2][i];
        const double err = std::fabs(rec - z[i]);
        ae = std::max(ae, err);
        if (z[i] > 0) { re = std::max(re, err / z[i]); } else if (rec != 0) { re = INF; }
    }
    r.mass_abs = ae; r.mass_rel = re;
}

double log_distance(int n, const double* a, const double* b, const double* z) { // pt_three_phase.hpp:357-380
    double d = 0;
    for (int i = 0; i < n; ++i) {
        if (z[i] == 0) { if (a[i] != 0 || b[i] != 0) { return INF; } continue; }
        if (!(a[i] > 0) || !(b[i] > 0)) { return INF; }
        const double v = std::fabs(std::log(a[i]) - std::log(b[i]));
        if (!std::isfinite(v)) { return INF; }
        d = std::max(d, v);
    }
    return d;
}

// iterate_pt_three_phase, pt_three_phase.hpp:402-674, with Pr76ThreePhaseEvaluator sides.
void iterate_three(const Mixture& m, double p, double t, const double* feed_in, const double* lk1_0,
                   const double* lk2_0, const double* b0, const int* sides, const P3Options& o, P3Result& R,
                   Counters& cnt) {
    const int n = m.n;
    if (!rr3_feasible(b0)) { throw Rejected{}; }
    for (int i = 0; i < n; ++i) { if (!std::isfinite(lk1_0[i]) || !std::isfinite(lk2_0[i])) { throw Rejected{}; } }
    R.has_point = false; R.disappearing = -1; R.status = P_NO_BALANCE;
    const double isum = check_composition(feed_in, n);
    double feed[NCMAX];
    for (int i = 0; i < n; ++i) { feed[i] = feed_in[i] / isum; }
    for (int i = 0; i < n; ++i) { if (feed_in[i] > 0 && feed[i] == 0) { R.status = P_BAL_FAILURE; return; } }
    uint64_t evaluations = 0;
    // returns -1 on success, else a P3Status
    auto evaluate = [&](const double* l1, const double* l2, const double* seed, P3State& s) -> int {
        RR3Result bal;
        rr3_solve(n, feed, l1, l2, seed, o, bal, cnt);
        switch (bal.status) {
        case G_SUCCESS: break;
        case G_DEGENERATE: return P_DEGENERATE;
        case G_ITERATION_LIMIT: return P_BAL_ITER;
        case G_LINE_SEARCH: return P_BAL_LS;
        default: return P_NO_BALANCE;
        }
        s.frac[0] = 1.0 - bal.beta[0] - bal.beta[1]; s.frac[1] = bal.beta[0]; s.frac[2] = bal.beta[1];
        for (int ph = 0; ph < 3; ++ph) {
            for (int i = 0; i < n; ++i) { s.comp[ph][i] = bal.comp[ph][i]; }
            if (evaluations >= o.max_evaluations) { return P_EVALUATION_LIMIT; }
            ++evaluations; ++cnt.prop;
            RPhase prop;
            if (side_provider(m, p, t, s.comp[ph], sides[ph], prop) != NONE) { return P_PROPERTY; }
            if (!prop.act.smooth || !std::isfinite(prop.z) || !(prop.z > 0)) { return P_PROPERTY; }
            s.act[ph] = prop.act;
            s.z[ph] = prop.z;
        }
        s.rr_res = bal.residual; s.mass_abs = bal.mass_abs; s.mass_rel = bal.mass_rel;
        double gc = 0, mag = 1, mc = 0;
        s.norm = 0; s.gibbs = 0;
        for (int i = 0; i < n; ++i) {
            s.lk1[i] = s.lk2[i] = s.r1[i] = s.r2[i] = s.common[i] = 0;
            if (feed[i] == 0) { continue; }
            const double l0 = std::log(s.comp[0][i]), ll1 = std::log(s.comp[1][i]), ll2 = std::log(s.comp[2][i]);
            const double mu0 = l0 + s.act[0].lnphi[i], mu1 = ll1 + s.act[1].lnphi[i], mu2 = ll2 + s.act[2].lnphi[i];
            s.lk1[i] = ll1 - l0;
            s.lk2[i] = ll2 - l0;
            s.r1[i] = mu0 - mu1;
            s.r2[i] = mu0 - mu2;
            s.common[i] = (mu0 + mu1 + mu2) / 3.0;
            s.norm = std::max({s.norm, std::fabs(s.r1[i]), std::fabs(s.r2[i])});
            for (int ph = 0; ph < 3; ++ph) {
                const double lx = std::log(s.comp[ph][i]);
                const double mu = lx + s.act[ph].lnphi[i];
                kadd(s.frac[ph] * s.comp[ph][i] * mu, s.gibbs, gc);
                kadd(s.frac[ph] * s.comp[ph][i] * (std::fabs(lx) + std::fabs(s.act[ph].lnphi[i])), mag, mc);
            }
        }
        s.guard = 256.0 * EPS * mag;
        if (!std::isfinite(s.norm) || !std::isfinite(s.gibbs) || !std::isfinite(s.guard)) { return P_PROPERTY; }
        return -1;
    };
    {
        const int e = evaluate(lk1_0, lk2_0, b0, R.pt);
        if (e != -1) { R.status = e; return; }
        R.has_point = true;
    }
    int iterations = 0;
    for (;;) {
        P3State& pt = R.pt;
        if (pt.norm <= o.mu_tol) {
            if (pt.rr_res > o.bal_tol || pt.mass_abs > o.mass_abs || pt.mass_rel > o.mass_rel) {
                R.status = P_BAL_FAILURE; return;
            }
            for (int ph = 0; ph < 3; ++ph) {
                if (pt.frac[ph] <= o.min_pf) { R.status = P_DISAPPEARANCE; R.disappearing = ph; return; }
            }
            const double d01 = log_distance(n, pt.comp[0], pt.comp[1], feed);
            const double d02 = log_distance(n, pt.comp[0], pt.comp[2], feed);
            const double d12 = log_distance(n, pt.comp[1], pt.comp[2], feed);
            if (!(d01 > o.lsep) || !(d02 > o.lsep) || !(d12 > o.lsep)) { R.status = P_INDISTINGUISHABLE; return; }
            R.status = P_CONVERGED;
            return;
        }
        if (iterations >= o.max_iterations) { R.status = P_ITERATION_LIMIT; return; }
        const double scale = std::max(1.0, pt.norm / o.max_log_step);
        bool acc = false;
        double alpha = 1.0;
        for (int bt = 0; bt < o.max_backtracks; ++bt, alpha *= 0.5) {
            double n1[NCMAX], n2[NCMAX];
            for (int i = 0; i < n; ++i) {
                n1[i] = pt.lk1[i] + (alpha / scale) * pt.r1[i];
                n2[i] = pt.lk2[i] + (alpha / scale) * pt.r2[i];
            }
            const double seed[2] = {pt.frac[1], pt.frac[2]};
            P3State nx;
            const int e = evaluate(n1, n2, seed, nx);
            if (e == -1) {
                const double req = o.res_dec * (alpha / scale) * pt.norm;
                const bool prog = nx.norm <= o.mu_tol || pt.norm - nx.norm > req;
                const bool gok = nx.gibbs <= pt.gibbs + pt.guard + nx.guard;
                if (prog && gok) { R.pt = nx; acc = true; break; }
            } else if (e == P_EVALUATION_LIMIT) {
                R.status = P_EVALUATION_LIMIT; return;
            }
        }
        if (!acc) { R.status = P_LINE_SEARCH_FAILED; return; }
        ++iterations;
    }
}

// ---------------------------------------------------------------- max3
enum Max3Status { M_SINGLE = 0, M_TWO, M_THREE, M_BOUNDARY, M_HIGHER, M_INDETERMINATE };
struct Max3Result {
    int status;
    int attempts;     // three-phase attempts run (Pr76PtMax3Result::attempts.size())
    bool start_consumed; // an attempt ran from the supplied start (pr76_pt_continuation.hpp:183-190)
    VleResult base;
    bool use_neighbor;
    VleResult neighbor;
    P3State three;
};

// pr76_witness_root_sides, pr76_three_phase.hpp:273-305. Bits: 1 = lower, 2 = upper.
int witness_sides(const Mixture& m, double p, double t, const double* w, int branch) {
    PhaseCtx ctx;
    if (prepare(m, p, t, w, ctx) != NONE) { return 0; }
    int adm[3], na = 0;
    for (int k = 0; k < ctx.rs.count; ++k) { if (ctx.rs.r[k].slope_sign > 0) { adm[na++] = k; } }
    bool found = false;
    for (int k = 0; k < na; ++k) { found = found || adm[k] == branch; }
    if (na == 0 || !found) { return 0; }
    if (na == 1) { return 3; }
    if (branch == adm[0]) { return 1; }
    if (branch == adm[na - 1]) { return 2; }
    return 0;
}

bool log_ratio(int n, const double* num, const double* den, const double* z, double* out) { // :248-271
    for (int i = 0; i < n; ++i) {
        out[i] = 0;
        if (z[i] == 0) { if (num[i] != 0 || den[i] != 0) { return false; } continue; }
        if (!(num[i] > 0) || !(den[i] > 0)) { return false; }
        out[i] = std::log(num[i]) - std::log(den[i]);
        if (!std::isfinite(out[i])) { return false; }
    }
    return true;
}

// solve_pr76_pt_max3, pr76_three_phase.hpp:328-561; start / start_frac: one caller-supplied
// three-phase start (three_phase_starts), or nullptr.
// initial / final: extra starts of the initial stability test and of the base split's
// common-tangent review (solve_pr76_pt_max3 initial_starts / final_starts,
// pr76_three_phase.hpp:328-341).
void solve_max3(const Mixture& m, double p, double t, const double* feed, const SplitOptions& so,
                Max3Result& M, Counters& cnt, const double* const* initial = nullptr, int ninitial = 0,
                const double* const* final_starts = nullptr, int nfinal = 0, const double* start = nullptr,
                const double* start_frac = nullptr, int nstart = 1) {
    const int n = m.n;
    (void)check_composition(feed, n);
    M.use_neighbor = false;
    M.attempts = 0;
    M.start_consumed = false;
    // pr76_pt_max3_check_starts (pr76_three_phase.hpp:211-242): the supplied starts
    // (start[s*3*n ...], seeds start_frac[2*s ...], s < nstart) are validated before any solve;
    // a violation is the reference's domain_error
    for (int sl = 0; start && sl < nstart; ++sl) {
        const double* st = start + sl * 3 * n;
        const double* sf = start_frac + 2 * sl;
        if (!(std::isfinite(sf[0]) && std::isfinite(sf[1]) && sf[0] >= 0.0 && sf[1] >= 0.0 && sf[0] + sf[1] <= 1.0)) {
            throw Rejected{};
        }
        for (int ph = 0; ph < 3; ++ph) {
            (void)check_composition(st + ph * n, n);
            for (int i = 0; i < n; ++i) { if ((feed[i] > 0) != (st[ph * n + i] > 0)) { throw Rejected{}; } }
        }
    }
    solve_vle(m, p, t, feed, so, initial, ninitial, M.base, cnt, final_starts, nfinal);
    switch (M.base.status) {
    case V_SINGLE: M.status = M_SINGLE; return;
    case V_TWO: M.status = M_TWO; return;
    case V_INDETERMINATE: M.status = M_INDETERMINATE; return;
    default: break;
    }
    if (!M.base.has_candidate || !M.base.has_final) { M.status = M_INDETERMINATE; return; }
    const StabResult& fin = M.base.fin;
    const SplitState& pair = M.base.cand.pt;
    const double* z = M.base.init.feed;
    bool saw_higher = false, saw_boundary = false;
    int automatic = 0;
    bool auto_limit = false;
    P3Options po;
    int continuation = 0;
    bool cont_limit = false;
    // returns true when an attempt closes the solve; supplied: the caller's start (its own
    // attempt budget, pr76_three_phase.hpp:385-394)
    auto run_attempt = [&](const double* l1, const double* l2, const double* fr, const int* sides,
                           bool supplied = false, int index = -1) -> bool {
        int& count = supplied ? continuation : automatic;
        bool& limit = supplied ? cont_limit : auto_limit;
        if (count >= 16) { limit = true; return false; }
        ++count;
        ++M.attempts;
        if (supplied && index == 0) { M.start_consumed = true; }
        P3Result E;
        iterate_three(m, p, t, z, l1, l2, fr, sides, po, E, cnt);
        if (E.status == P_DISAPPEARANCE) {
            saw_boundary = true;
            if (E.has_point && E.disappearing >= 0) {
                const double* st[2];
                int ns = 0;
                for (int ph = 0; ph < 3; ++ph) { if (ph != E.disappearing) { st[ns++] = E.pt.comp[ph]; } }
                solve_vle(m, p, t, z, so, st, 2, M.neighbor, cnt);
                if (M.neighbor.status == V_TWO) { M.status = M_TWO; M.use_neighbor = true; return true; }
            }
            return false;
        }
        if (E.status != P_CONVERGED) { return false; }
        const P3State& c = E.pt;
        if (!(c.norm <= po.mu_tol && c.rr_res <= po.bal_tol && c.mass_abs <= po.mass_abs &&
              c.mass_rel <= po.mass_rel)) { return false; }
        StabOptions fo;
        fo.tpd_tol += c.norm;
        const double* st[3] = {c.comp[0], c.comp[1], c.comp[2]};
        StabResult* F = new StabResult;
        stability(m, p, t, z, c.common, st, 3, fo, *F, cnt);
        const int fs = F->status;
        delete F;
        if (fs == S_UNSTABLE) { saw_higher = true; }
        else if (fs == S_NO_INSTABILITY) { M.status = M_THREE; M.three = c; return true; }
        return false;
    };
    // the caller-supplied starts first, in order, each over the eight root-side combinations,
    // lower before upper, phase 0 outermost (pr76_three_phase.hpp:472-498); M.start_consumed
    // records an attempt from start 0 (pr76_pt_continuation_supplied_start_consumed)
    for (int sl = 0; start && sl < nstart && !cont_limit; ++sl) {
        const double* st = start + sl * 3 * n;
        double sl1[NCMAX], sl2[NCMAX];
        if (!(log_ratio(n, st + n, st, z, sl1) && log_ratio(n, st + 2 * n, st, z, sl2))) { continue; }
        for (int code = 0; code < 8 && !cont_limit; ++code) {
            const int sd[3] = {code >> 2, (code >> 1) & 1, code & 1};
            if (run_attempt(sl1, sl2, start_frac + 2 * sl, sd, true, sl)) { return; }
        }
    }
    double src1[NCMAX];
    if (log_ratio(n, pair.fr.y, pair.fr.x, z, src1)) {
        for (int k = 0; k < fin.ntrials && !auto_limit; ++k) {
            const Trial& w = fin.trials[k];
            if (!(w.status == T_NEGATIVE && w.has_point && negative(w.pt, fin.opt))) { continue; }
            double l2[NCMAX];
            if (!log_ratio(n, w.pt.w, pair.fr.x, z, l2)) { continue; }
            const int sides = witness_sides(m, p, t, w.pt.w, w.pt.branch);
            for (int s = 0; s < 2; ++s) {
                if (!(sides & (1 << s))) { continue; }
                const double bv = pair.fr.beta;
                const double fr[2] = {(1.0 - 0.1) * bv, 0.1};
                const int sd[3] = {0, 1, s};
                if (run_attempt(src1, l2, fr, sd)) { return; }
                if (auto_limit) { break; }
            }
        }
    }
    M.status = saw_higher ? M_HIGHER : (saw_boundary ? M_BOUNDARY : M_INDETERMINATE);
}

// ---------------------------------------------------------------- publication
struct Out {
    int32_t* status;
    int32_t* phase_count;
    double* beta;
    double* zfactor;
    int32_t* branch;
    double* x;
    double* ln_phi;
    uint32_t* counters;
};

struct Phase {
    double beta;
    const double* x;
    const SPhase* act;
    double z;
};

void write_cell(const Out& o, int n, int64_t N, int64_t c, int set, int max3, int split, int np, const Phase* ph,
                const Counters& cnt, int flags) {
    o.status[c] = set | (max3 << 4) | (split << 8) | flags;
    o.phase_count[c] = np;
    for (int k = 0; k < 3; ++k) {
        const bool have = k < np;
        o.beta[k * N + c] = have ? ph[k].beta : NaN;
        o.zfactor[k * N + c] = have ? ph[k].z : NaN;
        o.branch[k * N + c] = have ? (ph[k].act->branch | (ph[k].act->smooth ? 0x100 : 0)) : -1;
        for (int i = 0; i < n; ++i) {
            o.x[(k * n + i) * N + c] = have ? ph[k].x[i] : NaN;
            o.ln_phi[(k * n + i) * N + c] = have ? ph[k].act->lnphi[i] : NaN;
        }
    }
    if (o.counters) {
        o.counters[0 * N + c] = static_cast<uint32_t>(cnt.prop);
        o.counters[1 * N + c] = static_cast<uint32_t>(cnt.tpd);
        o.counters[2 * N + c] = static_cast<uint32_t>(cnt.rr);
        o.counters[3 * N + c] = static_cast<uint32_t>(cnt.rr3);
    }
}

// pt_vle_phase_set.hpp:169-258: pair = [liquid (1-beta_V), vapor (beta_V)].
int pair_phases(const VleResult& V, Phase* ph) {
    const SplitState& s = V.cand.pt;
    ph[0] = {1.0 - s.fr.beta, s.fr.x, &s.liq.act, s.liq.z};
    ph[1] = {s.fr.beta, s.fr.y, &s.vap.act, s.vap.z};
    return 2;
}

// Status-word flags of the base split (mpmc_hnu_amd.h MPMC_FLAG_SPLIT_*): the facts of
// PtSplitResult that the transition evidence reads (pr76_pt_flash_backend.hpp:42-95).
int split_flags(const VleResult& V) {
    return (V.any_disappearance ? 0x4000 : 0) | (V.has_candidate ? 0x8000 : 0);
}

void publish_vle(const Out& o, int n, int64_t N, int64_t c, const VleResult& V, int max3_code, const Counters& cnt,
                 int extra_flags = 0) {
    Phase ph[3];
    const int fl = split_flags(V) | extra_flags;
    switch (V.status) {
    case V_SINGLE:
        if (!V.init.has_ref) { write_cell(o, n, N, c, 2, max3_code, V.status, 0, ph, cnt, fl); return; }
        ph[0] = {1.0, V.init.feed, &V.init.ref, NaN};
        write_cell(o, n, N, c, 0, max3_code, V.status, 1, ph, cnt, fl);
        return;
    case V_TWO:
        write_cell(o, n, N, c, 0, max3_code, V.status, pair_phases(V, ph), ph, cnt, fl);
        return;
    case V_UNSTABLE:
        write_cell(o, n, N, c, 1, max3_code, V.status, V.has_candidate ? pair_phases(V, ph) : 0, ph, cnt, fl);
        return;
    default:
        write_cell(o, n, N, c, 2, max3_code, V.status, V.has_candidate ? pair_phases(V, ph) : 0, ph, cnt, fl);
        return;
    }
}

// pt_phase_set.hpp:90-150 (project_max3_solution).
void publish_max3(const Out& o, int n, int64_t N, int64_t c, const Max3Result& M, const Counters& cnt) {
    Phase ph[3];
    const int split = M.base.status;
    const int fl = split_flags(M.base) | (M.attempts > 0 ? 0x10000 : 0) | (M.start_consumed ? 0x20000 : 0);
    switch (M.status) {
    case M_SINGLE:
        publish_vle(o, n, N, c, M.base, M.status, cnt, (M.attempts > 0 ? 0x10000 : 0) | (M.start_consumed ? 0x20000 : 0));
        return;
    case M_TWO: {
        const VleResult& V = M.use_neighbor ? M.neighbor : M.base;
        write_cell(o, n, N, c, 0, M.status, split, pair_phases(V, ph), ph, cnt, fl);
        return;
    }
    case M_THREE:
        for (int k = 0; k < 3; ++k) { ph[k] = {M.three.frac[k], M.three.comp[k], &M.three.act[k], M.three.z[k]}; }
        write_cell(o, n, N, c, 0, M.status, split, 3, ph, cnt, fl);
        return;
    case M_HIGHER:
        write_cell(o, n, N, c, 1, M.status, split, M.base.has_candidate ? pair_phases(M.base, ph) : 0, ph, cnt, fl);
        return;
    default:
        write_cell(o, n, N, c, 2, M.status, split, 0, ph, cnt, fl);
        return;
    }
}

// ---------------------------------------------------------------- forward-mode AD
// Dual number of ad/dual.hpp:78-345 with the derivative rules of ad/math.hpp
// (log :200-214, log1p :258-272, sqrt :277-291): value plus N directional derivatives.
template <int N>
struct Dual {
    double v = 0;
    double d[N] = {};
    Dual() = default;
    Dual(double x) : v(x) {}
};
struct AdDomain {};   // std::domain_error (AD math domain, division by zero, mixing domain)
struct RangeFail {};  // std::range_error
struct PhaseFail {};  // Pr76PhaseError
struct IndexFail {};  // std::out_of_range (escapes the sensitivity's handlers)

inline double val(double x) { return x; }
template <int N> inline double val(const Dual<N>& x) { return x.v; }
inline bool fin(double x) { return std::isfinite(x); }
template <int N> inline bool fin(const Dual<N>& x) {
    if (!std::isfinite(x.v)) { return false; }
    for (int k = 0; k < N; ++k) { if (!std::isfinite(x.d[k])) { return false; } }
    return true;
}

template <int N> Dual<N> operator+(Dual<N> a, const Dual<N>& b) { a.v += b.v; for (int k = 0; k < N; ++k) { a.d[k] += b.d[k]; } return a; }
template <int N> Dual<N> operator-(Dual<N> a, const Dual<N>& b) { a.v -= b.v; for (int k = 0; k < N; ++k) { a.d[k] -= b.d[k]; } return a; }
template <int N> Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r;
    for (int k = 0; k < N; ++k) { r.d[k] = a.d[k] * b.v + a.v * b.d[k]; }
    r.v = a.v * b.v;
    return r;
}
template <int N> Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    if (b.v == 0) { throw AdDomain{}; }
    Dual<N> r;
    const double q = a.v / b.v;
    for (int k = 0; k < N; ++k) { r.d[k] = (a.d[k] - q * b.d[k]) / b.v; }
    r.v = q;
    return r;
}
template <int N> Dual<N> operator+(Dual<N> a, double s) { a.v += s; return a; }
template <int N> Dual<N> operator+(double s, const Dual<N>& a) { return Dual<N>(s) + a; }
template <int N> Dual<N> operator-(Dual<N> a, double s) { a.v -= s; return a; }
template <int N> Dual<N> operator-(double s, const Dual<N>& a) { return Dual<N>(s) - a; }
template <int N> Dual<N> operator*(Dual<N> a, double s) { a.v *= s; for (int k = 0; k < N; ++k) { a.d[k] *= s; } return a; }
template <int N> Dual<N> operator*(double s, Dual<N> a) { a.v = s * a.v; for (int k = 0; k < N; ++k) { a.d[k] = s * a.d[k]; } return a; }
template <int N> Dual<N> operator/(Dual<N> a, double s) {
    if (s == 0) { throw AdDomain{}; }
    a.v /= s; for (int k = 0; k < N; ++k) { a.d[k] /= s; } return a;
}
template <int N> Dual<N> operator/(double s, const Dual<N>& a) { return Dual<N>(s) / a; }
template <int N> Dual<N>& operator+=(Dual<N>& a, const Dual<N>& b) { a = a + b; return a; }
template <int N> Dual<N>& operator-=(Dual<N>& a, const Dual<N>& b) { a = a - b; return a; }
template <int N> Dual<N> dlog(const Dual<N>& x) {
    if (!std::isfinite(x.v) || !(x.v > 0)) { throw AdDomain{}; }
    Dual<N> r(std::log(x.v));
    for (int k = 0; k < N; ++k) { r.d[k] = x.d[k] / x.v; }
    return r;
}
template <int N> Dual<N> dlog1p(const Dual<N>& x) {
    if (!std::isfinite(x.v) || !(x.v > -1)) { throw AdDomain{}; }
    Dual<N> r(std::log1p(x.v));
    for (int k = 0; k < N; ++k) { r.d[k] = x.d[k] / (1.0 + x.v); }
    return r;
}
template <int N> Dual<N> dsqrt(const Dual<N>& x) {
    if (!std::isfinite(x.v) || !(x.v > 0)) { throw AdDomain{}; }
    const double s = std::sqrt(x.v);
    Dual<N> r(s);
    for (int k = 0; k < N; ++k) { r.d[k] = (x.d[k] / s) / 2.0; }
    return r;
}
inline double dlog(double x) { return std::log(x); }
inline double dlog1p(double x) { return std::log1p(x); }
inline double dsqrt(double x) { return std::sqrt(x); }

// log1p(u)/u in Number arithmetic, pr76_phase.hpp:67-79.
template <typename Num>
Num log1p_over_x_n(const Num& u) {
    if (std::fabs(val(u)) <= 0.125) {
        constexpr int degree = 53 / 3 + 2;
        Num v((degree % 2 == 0 ? 1.0 : -1.0) / static_cast<double>(degree + 1));
        for (int k = degree - 1; k >= 0; --k) { v = (k % 2 == 0 ? 1.0 : -1.0) / static_cast<double>(k + 1) + u * v; }
        return v;
    }
    return dlog1p(u) / u;
}

// Pr76Phase::evaluate_reduced (pr76_phase.hpp:193-258) through Pr76Mixture::evaluate_impl
// <Reduced=true, WithSums=true> (pr76_mixture.hpp:124-219) and Pr76Pure::evaluate
// (pr76_pure.hpp:144-165): n-1 independent fractions, x_last = 1 - sum in Number
// arithmetic; roots on the primal and the first-order IFT for the selected root.
// Reduced = false is Pr76Phase::evaluate_full (pr76_phase.hpp:135-140): xr holds all n
// fractions, each checked to lie in [0, 1] and their compensated primal sum within 64 eps
// of one (pr76_mixture.hpp:146-158).
template <typename Num, bool Reduced = true>
void phase_eval(const Mixture& m, const Num& p, const Num& t, const Num* xr, int root_index, Num& zout,
                Num* lnphi) {
    const int n = m.n;
    if (!fin(p) || !(val(p) > 0)) { throw AdDomain{}; }
    Num last(1.0), isum(0.0);
    if constexpr (Reduced) {
        for (int i = 0; i + 1 < n; ++i) {
            if (!fin(xr[i]) || val(xr[i]) < 0 || val(xr[i]) > 1) { throw AdDomain{}; }
            isum += xr[i];
        }
        last -= isum;
        if (!fin(last) || val(last) < 0 || val(last) > 1) { throw AdDomain{}; }
    } else {
        long double sum = 0, comp = 0;
        for (int i = 0; i < n; ++i) {
            if (!fin(xr[i]) || val(xr[i]) < 0 || val(xr[i]) > 1) { throw AdDomain{}; }
            const long double inc = val(xr[i]) - comp;
            const long double next = sum + inc;
            comp = (next - sum) - inc;
            sum = next;
        }
        if (std::fabs(sum - 1.0L) > static_cast<long double>(64.0 * EPS)) { throw AdDomain{}; }
    }
    auto x = [&](int i) -> const Num& { return (Reduced && i == n - 1) ? last : xr[i]; };
    if (!fin(t) || !(val(t) > 0)) { throw AdDomain{}; }
    Num roots_a[NCMAX], sums[NCMAX];
    Num att(0.0), cov(0.0);
    for (int i = 0; i < n; ++i) {
        const Num rr = dsqrt(t) / m.sqrt_tc[i];
        const Num br = 1.0 + m.kappa[i] * (1.0 - rr);
        if (!fin(rr) || !fin(br)) { throw RangeFail{}; }
        const Num a = (br * m.ac[i]) * br;
        if (!fin(a) || (val(a) == 0 && val(br) != 0)) { throw RangeFail{}; }
        att += (x(i) * a) * x(i);
        cov += x(i) * Num(m.b[i]);
        sums[i] = x(i) * a;
        if (n > 1) {
            if (!(val(a) > 0)) { throw AdDomain{}; }
            roots_a[i] = dsqrt(a);
            if (!fin(roots_a[i])) { throw RangeFail{}; }
        }
    }
    for (int i = 0; i < n; ++i) {
        for (int j = i + 1; j < n; ++j) {
            const double f = m.factor[i * n + j];
            const Num cross = ((x(i) * roots_a[i]) * (x(j) * roots_a[j])) * f;
            att += cross;
            att += cross;
            const Num aij = (roots_a[i] * f) * roots_a[j];
            sums[i] += x(j) * aij;
            sums[j] += x(i) * aij;
        }
    }
    if (!fin(att) || !fin(cov) || !(val(cov) > 0)) { throw RangeFail{}; }
    for (int i = 0; i < n; ++i) { if (!fin(sums[i])) { throw RangeFail{}; } }
    // Pr76Phase::coefficients, pr76_phase.hpp:176-191
    const Num rt = R_GAS * t;
    if (!fin(rt) || !(val(rt) > 0)) { throw RangeFail{}; }
    const Num prt = p / rt;
    const Num A = (att / rt) * prt;
    const Num B = cov * prt;
    if (!fin(A) || !fin(B) || !(val(B) > 0) || (val(A) == 0 && val(att) != 0)) { throw RangeFail{}; }
    const RootSet rs = roots(val(A), val(B));
    if (rs.status != RS_SUCCESS) { throw PhaseFail{}; }
    if (root_index < 0 || root_index >= rs.count) { throw IndexFail{}; }
    const Root& r = rs.r[root_index];
    Num fv(r.x);
    if constexpr (!std::is_same_v<Num, double>) {
        if (!r.dvalid) { throw PhaseFail{}; }
        const double y0 = r.x / rs.scale;
        // pr76_cubic_coefficients in Number arithmetic (pr76_roots.hpp:50-58)
        const Num bs = B / rs.scale;
        const double inv = 1.0 / rs.scale;
        const Num c2 = 4.0 * bs - inv;
        const Num c1 = (A / rs.scale) / rs.scale + 2.0 * bs * bs - 4.0 * bs * inv;
        const Num c0 = ((-2.0 * bs) * bs) * inv;
        const Num resid = ((Num(y0) + c2) * y0 + c1) * y0 + c0;
        const double dh = (3.0 * y0 + 2.0 * val(c2)) * y0 + val(c1);
        fv -= ((resid - Num(val(resid))) / dh) * rs.scale;
    }
    const Num z = fv + B;
    const double s2 = std::sqrt(2.0);
    const Num den = fv + (2.0 - s2) * B;
    const Num u = (2.0 * s2) * (B / den);
    const Num lf = log1p_over_x_n(u) / den;
    Num zm1 = z - 1.0;
    Num lfree = dlog(fv);
    if (std::fabs(val(zm1)) < 0.25 && val(B) < 0.125) {
        const Num bz = B / z;
        zm1 = B / fv - (A / z) / (1.0 + 2.0 * bz - bz * bz);
        lfree = dlog1p(zm1 - B);
    }
    const Num rt2 = R_GAS * t;
    const Num prt2 = p / rt2;
    for (int i = 0; i < n; ++i) {
        const Num ratio = Num(m.b[i]) / cov;
        const Num coef = ((2.0 * sums[i] - att * ratio) / rt2) * prt2;
        lnphi[i] = ratio * zm1 - lfree - coef * lf;
        if (!fin(lnphi[i])) { throw RangeFail{}; }
    }
    if (!fin(z)) { throw RangeFail{}; }
    zout = z;
}

template <typename Num>
void phase_reduced(const Mixture& m, const Num& p, const Num& t, const Num* xr, int root_index, Num& zout,
                   Num* lnphi) {
    phase_eval<Num, true>(m, p, t, xr, root_index, zout, lnphi);
}

// ---------------------------------------------------------------- closure
// Statuses: ThermodynamicClosurePrimalStatus / LinearizationStatus / LinearizationReason
// (thermodynamic_closure.hpp:18-35) and PtSensitivityStatus (pt_sensitivity.hpp).
enum ClosurePrimal { CP_VALID = 0, CP_NOT_IMPLEMENTED = 1, CP_INDETERMINATE = 2 };
enum ClosureReason { CR_NONE = 0, CR_NOT_IMPLEMENTED, CR_PHASE_BOUNDARY, CR_ILL_CONDITIONED, CR_FEED_SUPPORT,
                     CR_PROPERTY_FAILURE, CR_ARITHMETIC_FAILURE, CR_NOT_ACCEPTED };
struct SensFail { int reason; };

struct ClosureResult {
    int primal = CP_INDETERMINATE, lin = 1, reason = CR_NONE;
    bool thrown = false;
    double zl = NaN, zv = NaN, rhol = NaN, rhov = NaN;
    double jac[(2 * NCMAX + 5) * (NCMAX + 1)];
    double dlk[NCMAX * (NCMAX + 1)];
    bool have_jac = false, have_dlk = false;
    double rcond = NaN, backward = NaN, resid = NaN;
};

// pt_sensitivity_rr_linearization, pt_sensitivity_detail.hpp:45-143
template <int K>
void rr_linearization(int n, const Dual<K>* feed, const Dual<K>* lk, const double* bx, const double* by,
                      double bbeta, Dual<K>& beta, Dual<K>* liq, Dual<K>* vap) {
    double x[NCMAX], y[NCMAX], xs = 0, ys = 0;
    for (int i = 0; i + 1 < n; ++i) { x[i] = bx[i]; y[i] = by[i]; xs += x[i]; ys += y[i]; }
    x[n - 1] = 1.0 - xs;
    y[n - 1] = 1.0 - ys;
    for (int i = 0; i < n; ++i) {
        if (!(feed[i].v > 0) || !(x[i] > 0) || !(y[i] > 0) || !std::isfinite(x[i]) || !std::isfinite(y[i])) {
            throw SensFail{CR_PHASE_BOUNDARY};
        }
    }
    double h = 0;
    for (int i = 0; i < n; ++i) { const double d = y[i] - x[i]; h += d * d / feed[i].v; }
    if (!std::isfinite(h) || !(h > 0)) { throw SensFail{CR_ILL_CONDITIONED}; }
    beta = Dual<K>(bbeta);
    for (int l = 0; l < K; ++l) {
        double num = 0;
        for (int i = 0; i < n; ++i) {
            const double z = feed[i].v;
            num += (x[i] * y[i] / z) * lk[i].d[l] + ((y[i] - x[i]) / z) * feed[i].d[l];
        }
        beta.d[l] = num / h;
        if (!std::isfinite(beta.d[l])) { throw SensFail{CR_ARITHMETIC_FAILURE}; }
    }
    Dual<K> ls(0.0), vs(0.0);
    for (int i = 0; i + 1 < n; ++i) {
        Dual<K> dx(x[i]), dy(y[i]);
        const double z = feed[i].v;
        for (int l = 0; l < K; ++l) {
            const double dlx = feed[i].d[l] / z - ((y[i] - x[i]) / z) * beta.d[l] - bbeta * (y[i] / z) * lk[i].d[l];
            dx.d[l] = x[i] * dlx;
            dy.d[l] = y[i] * (dlx + lk[i].d[l]);
            if (!std::isfinite(dx.d[l]) || !std::isfinite(dy.d[l])) { throw SensFail{CR_ARITHMETIC_FAILURE}; }
        }
        liq[i] = dx; vap[i] = dy;
        ls += dx; vs += dy;
    }
    liq[n - 1] = 1.0 - ls;
    vap[n - 1] = 1.0 - vs;
    if (!(liq[n - 1].v > 0) || !(vap[n - 1].v > 0)) { throw SensFail{CR_PHASE_BOUNDARY}; }
}

// Blocked forward mode of value_and_jacobian_runtime (ad/runtime_differentiate.hpp:87-168),
// direction width 4: ceil(inputs/4) passes, each re-running the whole primal; the
// primal outputs must be identical in every pass.
template <typename Fn>
void jacobian_blocked(int nin, int nout, const double* inputs, Fn fn, double* values, double* J) {
    constexpr int K = 4;
    Dual<K> in[2 * NCMAX + 1], out[3 * NCMAX + 1];
    for (int b0 = 0; b0 < nin; b0 += K) {
        for (int i = 0; i < nin; ++i) {
            in[i] = Dual<K>(inputs[i]);
            if (i >= b0 && i < b0 + K) { in[i].d[i - b0] = 1.0; }
        }
        fn(in, out);
        for (int o = 0; o < nout; ++o) {
            if (b0 == 0) { values[o] = out[o].v; }
            else if (!(values[o] == out[o].v) && !(std::isnan(values[o]) && std::isnan(out[o].v))) {
                throw SensFail{CR_ARITHMETIC_FAILURE};
            }
            for (int k = 0; k < K && b0 + k < nin; ++k) { J[o * nin + b0 + k] = out[o].d[k]; }
        }
    }
}

// pt_sensitivity_factor / _solve / _rcond, pt_sensitivity_detail.hpp:152-259
struct Lu { int n; double a[NCMAX * NCMAX]; int piv[NCMAX]; double norm; };
bool lu_factor(const double* A, int n, Lu& f) {
    f.n = n;
    std::memcpy(f.a, A, sizeof(double) * n * n);
    f.norm = 0;
    for (int r = 0; r < n; ++r) {
        double s = 0;
        for (int c = 0; c < n; ++c) { s += std::fabs(A[r * n + c]); }
        f.norm = std::max(f.norm, s);
    }
    if (!(f.norm > 0) || !std::isfinite(f.norm)) { return false; }
    const double guard = 64.0 * EPS * n * f.norm;
    for (int c = 0; c < n; ++c) {
        int p = c;
        double mag = std::fabs(f.a[c * n + c]);
        for (int r = c + 1; r < n; ++r) {
            const double cand = std::fabs(f.a[r * n + c]);
            if (cand > mag) { mag = cand; p = r; }
        }
        if (!std::isfinite(mag) || !(mag > guard)) { return false; }
        f.piv[c] = p;
        if (p != c) { for (int j = 0; j < n; ++j) { std::swap(f.a[c * n + j], f.a[p * n + j]); } }
        const double dg = f.a[c * n + c];
        for (int r = c + 1; r < n; ++r) {
            const double mlt = f.a[r * n + c] / dg;
            f.a[r * n + c] = mlt;
            for (int j = c + 1; j < n; ++j) { f.a[r * n + j] -= mlt * f.a[c * n + j]; }
        }
    }
    return true;
}
bool lu_solve(const Lu& f, double* rhs) {
    const int n = f.n;
    for (int c = 0; c < n; ++c) { if (f.piv[c] != c) { std::swap(rhs[c], rhs[f.piv[c]]); } }
    for (int r = 0; r < n; ++r) { for (int c = 0; c < r; ++c) { rhs[r] -= f.a[r * n + c] * rhs[c]; } }
    for (int r = n; r-- > 0;) {
        for (int c = r + 1; c < n; ++c) { rhs[r] -= f.a[r * n + c] * rhs[c]; }
        rhs[r] /= f.a[r * n + r];
        if (!std::isfinite(rhs[r])) { return false; }
    }
    return true;
}
double lu_rcond(const Lu& f) {
    const int n = f.n;
    double rows[NCMAX] = {};
    for (int c = 0; c < n; ++c) {
        double rhs[NCMAX] = {};
        rhs[c] = 1.0;
        if (!lu_solve(f, rhs)) { return 0.0; }
        for (int r = 0; r < n; ++r) { rows[r] += std::fabs(rhs[r]); }
    }
    double inv = rows[0];
    for (int r = 1; r < n; ++r) { inv = std::max(inv, rows[r]); }
    const double prod = f.norm * inv;
    if (!(prod > 0) || !std::isfinite(prod)) { return 0.0; }
    return 1.0 / prod;
}

// differentiate_pr76_pt_vle, pr76_sensitivity.hpp:51-390, for an accepted two-phase
// candidate. Writes dlogK (n x q), d beta (q), dx, dy (n x q each). Throws SensFail.
void differentiate(const Mixture& m, double p, double t, const double* feed, const SplitState& c,
                   const SplitOptions& so, double min_pf, double min_rcond, ClosureResult& R,
                   double* dbeta, double* dx, double* dy) {
    const int n = m.n;
    const int q = n + 1;
    for (int i = 0; i < n; ++i) { if (!(feed[i] > 0)) { throw SensFail{CR_FEED_SUPPORT}; } }
    double contrast = 0;
    for (int i = 0; i < n; ++i) { contrast = std::max(contrast, std::fabs(c.lk[i])); }
    if (!(contrast > so.lk_sep) || !(c.vap.z - c.liq.z > so.z_sep * std::max(c.vap.z, c.liq.z))) {
        throw SensFail{CR_PHASE_BOUNDARY};
    }
    const double beta = c.fr.beta;
    const double guard = std::max(min_pf, so.min_pf);
    if (!(beta > guard) || !(1.0 - beta > guard)) { throw SensFail{CR_PHASE_BOUNDARY}; }
    if (!(c.norm <= so.fug_tol)) { throw SensFail{CR_NOT_ACCEPTED}; }
    for (int i = 0; i < n; ++i) {
        if (!(c.fr.x[i] > 0) || !(c.fr.y[i] > 0) || !std::isfinite(c.lk[i])) { throw SensFail{CR_FEED_SUPPORT}; }
        const double rec = std::fma(beta, c.fr.y[i], (1.0 - beta) * c.fr.x[i]);
        const double err = std::fabs(rec - feed[i]);
        if (err > so.mass_abs || err / feed[i] > so.mass_rel) { throw IndexFail{}; } // invalid_argument
    }
    // pr76_sensitivity_local_system, detail/pr76_sensitivity_local.hpp:29-109
    const int nv = 2 * n + 1, no = 3 * n + 1;
    double inputs[2 * NCMAX + 1];
    for (int i = 0; i < n; ++i) { inputs[i] = c.lk[i]; }
    inputs[n] = p;
    inputs[n + 1] = t;
    for (int i = 0; i + 1 < n; ++i) { inputs[n + 2 + i] = feed[i]; }
    const int bl = c.liq.act.branch, bv = c.vap.act.branch;
    auto eqs = [&](const Dual<4>* v, Dual<4>* out) {
        Dual<4> lk[NCMAX], z[NCMAX], fs(0.0);
        for (int i = 0; i < n; ++i) { lk[i] = v[i]; }
        for (int i = 0; i + 1 < n; ++i) { z[i] = v[n + 2 + i]; fs += z[i]; }
        z[n - 1] = 1.0 - fs;
        Dual<4> b, xl[NCMAX], yv[NCMAX];
        rr_linearization<4>(n, z, lk, c.fr.x, c.fr.y, c.fr.beta, b, xl, yv);
        Dual<4> zl, zv, ll[NCMAX], lv[NCMAX];
        phase_reduced<Dual<4>>(m, v[n], v[n + 1], xl, bl, zl, ll);
        phase_reduced<Dual<4>>(m, v[n], v[n + 1], yv, bv, zv, lv);
        for (int i = 0; i < n; ++i) { out[i] = dlog(xl[i]) - dlog(yv[i]) + ll[i] - lv[i]; }
        out[n] = b;
        for (int i = 0; i < n; ++i) { out[n + 1 + i] = xl[i]; out[2 * n + 1 + i] = yv[i]; }
    };
    double vals[3 * NCMAX + 1], J[(3 * NCMAX + 1) * (2 * NCMAX + 1)];
    try {
        jacobian_blocked(nv, no, inputs, eqs, vals, J);
    } catch (const PhaseFail&) {
        throw SensFail{CR_PROPERTY_FAILURE};
    } catch (const RangeFail&) {
        throw SensFail{CR_PROPERTY_FAILURE};
    } catch (const AdDomain&) {
        throw SensFail{CR_ARITHMETIC_FAILURE};
    }
    double rn = 0;
    for (int i = 0; i < n; ++i) { rn = std::max(rn, std::fabs(vals[i])); }
    R.resid = rn;
    const double ascale = 8192.0 * EPS * n;
    if (rn > so.fug_tol + ascale) { throw SensFail{CR_NOT_ACCEPTED}; }
    double A[NCMAX * NCMAX], Bm[NCMAX * (NCMAX + 1)];
    for (int r = 0; r < n; ++r) {
        for (int col = 0; col < n; ++col) { A[r * n + col] = J[r * nv + col]; }
        for (int col = 0; col < q; ++col) { Bm[r * q + col] = J[r * nv + n + col]; }
    }
    Lu f;
    if (!lu_factor(A, n, f)) { throw SensFail{CR_ILL_CONDITIONED}; }
    R.rcond = lu_rcond(f);
    const double need = std::max(256.0 * EPS * n, min_rcond);
    if (!std::isfinite(R.rcond) || !(R.rcond > need)) { throw SensFail{CR_ILL_CONDITIONED}; }
    double* D = R.dlk;
    for (int col = 0; col < q; ++col) {
        double rhs[NCMAX];
        for (int r = 0; r < n; ++r) { rhs[r] = -Bm[r * q + col]; }
        if (!lu_solve(f, rhs)) { throw SensFail{CR_ILL_CONDITIONED}; }
        for (int r = 0; r < n; ++r) { D[r * q + col] = rhs[r]; }
    }
    double be = 0;
    for (int col = 0; col < q; ++col) {
        for (int r = 0; r < n; ++r) {
            double res = Bm[r * q + col];
            double sc = std::fabs(res);
            for (int j = 0; j < n; ++j) {
                const double term = A[r * n + j] * D[j * q + col];
                res += term;
                sc += std::fabs(term);
            }
            be = std::max(be, std::fabs(res) / std::max(1.0, sc));
        }
    }
    R.backward = be;
    if (!std::isfinite(be) || be > ascale) { throw SensFail{CR_ARITHMETIC_FAILURE}; }
    auto total = [&](int o, int col) {
        double v = J[o * nv + n + col];
        for (int j = 0; j < n; ++j) { v += J[o * nv + j] * D[j * q + col]; }
        return v;
    };
    bool finite_all = true;
    for (int col = 0; col < q; ++col) {
        dbeta[col] = total(n, col);
        finite_all = finite_all && std::isfinite(dbeta[col]);
        for (int i = 0; i < n; ++i) {
            dx[i * q + col] = total(n + 1 + i, col);
            dy[i * q + col] = total(2 * n + 1 + i, col);
            finite_all = finite_all && std::isfinite(dx[i * q + col]) && std::isfinite(dy[i * q + col]);
        }
    }
    for (int k = 0; k < n * q; ++k) { finite_all = finite_all && std::isfinite(D[k]); }
    if (!finite_all) { throw SensFail{CR_ARITHMETIC_FAILURE}; }
    const double bb = vals[n];
    for (int col = 0; col < q; ++col) {
        double ls = 0, vs = 0;
        for (int i = 0; i < n; ++i) { ls += dx[i * q + col]; vs += dy[i * q + col]; }
        if (std::fabs(ls) > ascale || std::fabs(vs) > ascale) { throw SensFail{CR_ARITHMETIC_FAILURE}; }
        for (int i = 0; i < n; ++i) {
            const double dz = col < 2 ? 0.0 : (i == col - 2 ? 1.0 : (i + 1 == n ? -1.0 : 0.0));
            const double xv = vals[n + 1 + i], yv = vals[2 * n + 1 + i];
            const double der = (1.0 - bb) * dx[i * q + col] + bb * dy[i * q + col] + (yv - xv) * dbeta[col] - dz;
            const double sc = std::max(1.0, std::fabs(dx[i * q + col]) + std::fabs(dy[i * q + col]) +
                                                std::fabs(dbeta[col]) + std::fabs(dz));
            if (std::fabs(der) > ascale * sc) { throw SensFail{CR_ARITHMETIC_FAILURE}; }
        }
    }
    R.have_dlk = true;
}

// pr76_phase_z_partials, pr76_thermodynamic_closure.hpp:81-118: Z over (p, T, w_0..w_{n-2}).
void z_partials(const Mixture& m, double p, double t, const double* w, int branch, double& zval, double* grad) {
    const int n = m.n;
    double inputs[NCMAX + 1];
    inputs[0] = p;
    inputs[1] = t;
    for (int i = 0; i + 1 < n; ++i) { inputs[2 + i] = w[i]; }
    auto fn = [&](const Dual<4>* v, Dual<4>* out) {
        Dual<4> lp[NCMAX];
        phase_reduced<Dual<4>>(m, v[0], v[1], v + 2, branch, out[0], lp);
    };
    jacobian_blocked(n + 1, 1, inputs, fn, &zval, grad);
}

inline bool same_z(double v, double acc) { // pr76_thermodynamic_closure.hpp:70-79
    if (!std::isfinite(v) || !std::isfinite(acc)) { return false; }
    return std::fabs(v - acc) <= 1024.0 * EPS * std::max(1.0, std::fabs(acc));
}

// build_pr76_pt_vle_closure, pr76_thermodynamic_closure.hpp:148-381, on the oracle's
// own VLE result. Row order of R.jac: beta_V, x_L[n], y_V[n], Z_L, Z_V, rho_L, rho_V.
void build_closure(const Mixture& m, double p, double t, const VleResult& V, const SplitOptions& so, double min_pf,
                   double min_rcond, ClosureResult& R) {
    const int n = m.n;
    const int q = n + 1;
    if (V.status == V_SINGLE) { R.primal = CP_NOT_IMPLEMENTED; return; }
    if (V.status == V_UNSTABLE || V.status == V_INDETERMINATE) { R.primal = CP_INDETERMINATE; return; }
    if (!V.has_candidate || !V.has_final || V.fin.status != S_NO_INSTABILITY) { R.primal = CP_INDETERMINATE; return; }
    const SplitState& c = V.cand.pt;
    const double beta = c.fr.beta;
    if (!std::isfinite(beta) || !(beta > 0) || !(beta < 1) || !std::isfinite(c.liq.z) || !(c.liq.z > 0) ||
        !std::isfinite(c.vap.z) || !(c.vap.z > 0)) {
        R.primal = CP_INDETERMINATE;
        return;
    }
    // primal reproduction at the full stored compositions
    try {
        PhaseCtx cl, cv;
        double lp[NCMAX], zl2 = NaN, zv2 = NaN;
        if (prepare(m, p, t, c.fr.x, cl) != NONE || c.liq.act.branch >= cl.rs.count ||
            lnphi(m, p, t, cl, c.liq.act.branch, lp, &zl2) != NONE) { throw RangeFail{}; }
        if (prepare(m, p, t, c.fr.y, cv) != NONE || c.vap.act.branch >= cv.rs.count ||
            lnphi(m, p, t, cv, c.vap.act.branch, lp, &zv2) != NONE) { throw RangeFail{}; }
        if (!same_z(zl2, c.liq.z) || !same_z(zv2, c.vap.z)) { throw RangeFail{}; }
        const double rl = p / (c.liq.z * R_GAS * t), rv = p / (c.vap.z * R_GAS * t);
        if (!std::isfinite(rl) || !(rl > 0) || !std::isfinite(rv) || !(rv > 0)) { throw RangeFail{}; }
        R.zl = c.liq.z; R.zv = c.vap.z; R.rhol = rl; R.rhov = rv;
        R.primal = CP_VALID;
    } catch (...) {
        R.primal = CP_INDETERMINATE;
        return;
    }
    double dbeta[NCMAX + 1], dx[NCMAX * (NCMAX + 1)], dy[NCMAX * (NCMAX + 1)];
    try {
        differentiate(m, p, t, V.init.feed, c, so, min_pf, min_rcond, R, dbeta, dx, dy);
    } catch (const SensFail& e) {
        R.reason = e.reason;
        if (e.reason == CR_NOT_ACCEPTED) {
            R.primal = CP_INDETERMINATE;
            R.zl = R.zv = R.rhol = R.rhov = NaN;
        }
        return;
    }
    try {
        double zlv, zvv, gl[NCMAX + 1], gv[NCMAX + 1];
        z_partials(m, p, t, c.fr.x, c.liq.act.branch, zlv, gl);
        z_partials(m, p, t, c.fr.y, c.vap.act.branch, zvv, gv);
        if (!same_z(zlv, R.zl) || !same_z(zvv, R.zv)) { throw AdDomain{}; }
        double* J = R.jac;
        bool ok = true;
        for (int col = 0; col < q; ++col) {
            double dzl = col == 0 ? gl[0] : (col == 1 ? gl[1] : 0.0);
            double dzv = col == 0 ? gv[0] : (col == 1 ? gv[1] : 0.0);
            for (int k = 0; k + 1 < n; ++k) {
                dzl += gl[2 + k] * dx[k * q + col];
                dzv += gv[2 + k] * dy[k * q + col];
            }
            const double dp = col == 0 ? 1.0 / p : 0.0;
            const double dt = col == 1 ? 1.0 / t : 0.0;
            J[0 * q + col] = dbeta[col];
            for (int i = 0; i < n; ++i) {
                J[(1 + i) * q + col] = dx[i * q + col];
                J[(1 + n + i) * q + col] = dy[i * q + col];
            }
            J[(2 * n + 1) * q + col] = dzl;
            J[(2 * n + 2) * q + col] = dzv;
            J[(2 * n + 3) * q + col] = R.rhol * (dp - dzl / R.zl - dt);
            J[(2 * n + 4) * q + col] = R.rhov * (dp - dzv / R.zv - dt);
        }
        for (int k = 0; k < (2 * n + 5) * q; ++k) { ok = ok && std::isfinite(J[k]); }
        if (!ok) { throw AdDomain{}; }
        R.have_jac = true;
        R.lin = 0;
        R.reason = CR_NONE;
    } catch (const PhaseFail&) {
        R.reason = CR_PROPERTY_FAILURE;
    } catch (const RangeFail&) {
        R.reason = CR_PROPERTY_FAILURE;
    } catch (...) {
        R.reason = CR_ARITHMETIC_FAILURE;
    }
}

// ---------------------------------------------------------------- two-phase assembly
// Build-defined transport/caloric/rock models. The reference leaves viscosity, enthalpy,
// rock storage and relative permeability to caller-owned providers
// (pr76_selected_phase_property_closure.hpp:32-45, selected_phase_production_cell_evaluator.hpp:24-64);
// these are this build's, evaluated in Number arithmetic so that AD flows through them:
//   mu  = a0 + a1 * rho_mass^2            [Pa s]
//   h   = cp * (T - T0)                   [J/kg]
//   U_r = cv_rock * (T - T0)              [J per rock m3]
//   kr_p = S_p^2 (Corey, exponent 2, no residual saturations)
struct AsmModel {
    double mu_a0, mu_a1, cp, t0, cv_rock;
    double molar_mass[NCMAX];
};

// Natural-variable chart of one two-phase cell (two_phase_cell_state.hpp:120-231):
// q = (p, T, S_0, x_0 independent [n-1], x_1 independent [n-1]), the last component of
// each phase dependent.
inline int q_count(int n) { return 2 * n + 1; }
inline int x_col(int n, int phase, int k) { return 3 + phase * (n - 1) + k; }

struct CellLin {
    // per phase: primal and full-q gradients (pr76_selected_phase_property_closure.hpp:497-530)
    double c[2], rho[2], mu[2], h[2], u[2], kr[2], lam[2];
    std::vector<double> dc[2], drho[2], dmu[2], dh[2], du[2], dkr[2], dlam[2];
    double x[2][NCMAX];
    double sat[2];
    double p, t;
    // equilibrium rows (n) and their Jacobian [n][q]
    double eq[NCMAX];
    std::vector<double> deq;
    // accumulation: N_i, total, E and gradients
    double acc[NCMAX], acc_total, energy;
    std::vector<double> dacc, dacc_total, denergy;
};

// reconstruct_composition (pr76_selected_phase_property_closure.hpp:115-151)
template <typename Num>
void reconstruct(int n, int phase, const Num* q, Num* x) {
    Num sum(0.0);
    for (int i = 0; i + 1 < n; ++i) {
        x[i] = q[x_col(n, phase, i)];
        if (!fin(x[i]) || !(val(x[i]) > 0)) { throw RangeFail{}; }
        sum += x[i];
    }
    x[n - 1] = 1.0 - sum;
    if (!fin(x[n - 1]) || !(val(x[n - 1]) > 0)) { throw RangeFail{}; }
}

// Pr76SelectedPhasePropertyClosure::evaluate (:336-465) for one phase: molar mass, molar
// density (evaluate_selected_phase_molar_density), ln(phi) (evaluate_selected_phase_fugacity),
// mass density, provider mu and h, u = h - p / rho_mass. Outputs c, rho, mu, h, u, lnphi[n].
template <typename Num>
void selected_phase(const Mixture& m, const AsmModel& md, int branch, const Num& p, const Num& t, const Num* x,
                    Num* out) {
    const int n = m.n;
    if (!fin(p) || !(val(p) > 0) || !fin(t) || !(val(t) > 0)) { throw RangeFail{}; }
    Num mm(0.0);
    for (int i = 0; i < n; ++i) {
        if (!fin(x[i]) || !(val(x[i]) > 0)) { throw RangeFail{}; }
        mm += x[i] * md.molar_mass[i];
    }
    if (!fin(mm) || !(val(mm) > 0)) { throw RangeFail{}; }
    Num z, lp[NCMAX];
    phase_eval<Num, false>(m, p, t, x, branch, z, lp); // density evaluation
    const Num den = (z * R_GAS) * t;
    const Num c = p / den;
    Num z2, lp2[NCMAX];
    phase_eval<Num, false>(m, p, t, x, branch, z2, lp2); // fugacity evaluation
    const Num rho = c * mm;
    if (!fin(c) || !(val(c) > 0) || !fin(rho) || !(val(rho) > 0)) { throw RangeFail{}; }
    const Num mu = md.mu_a0 + md.mu_a1 * (rho * rho);
    const Num h = md.cp * (t - md.t0);
    if (!fin(mu) || !(val(mu) > 0) || !fin(h)) { throw RangeFail{}; }
    const Num u = h - p / rho;
    if (!fin(u)) { throw RangeFail{}; }
    out[0] = c;
    out[1] = rho;
    out[2] = mu;
    out[3] = h;
    out[4] = u;
    for (int i = 0; i < n; ++i) {
        if (!fin(lp2[i])) { throw RangeFail{}; }
        out[5 + i] = lp2[i];
    }
}

// evaluate_pr76_selected_phase_property_chart<4> (:533-820) + the 2P bridge (:952-1098),
// relative permeability, mobility (two_phase_transport.hpp:38-137), component and energy
// accumulation (two_phase_accumulation.hpp:18-338) for one cell.
void cell_linearization(const Mixture& m, const AsmModel& md, const double* qv, const int* branch, double porosity,
                        CellLin& L) {
    const int n = m.n;
    const int Q = q_count(n);
    const int stride = 5 + n;
    const int nout = 2 * stride + n;
    std::vector<double> vals(nout), J(static_cast<size_t>(nout) * Q);
    auto fn = [&](const Dual<4>* v, Dual<4>* out) {
        Dual<4> x[2][NCMAX], props[2][5 + NCMAX];
        for (int ph = 0; ph < 2; ++ph) {
            reconstruct(n, ph, v, x[ph]);
            selected_phase(m, md, branch[ph], v[0], v[1], x[ph], props[ph]);
            for (int k = 0; k < stride; ++k) { out[ph * stride + k] = props[ph][k]; }
        }
        for (int i = 0; i < n; ++i) {
            out[2 * stride + i] = dlog(x[0][i] / x[1][i]) + props[0][5 + i] - props[1][5 + i] + dlog(v[0] / v[0]);
        }
    };
    // blocked forward mode; the chart has up to 25 inputs and 2(5+n)+n outputs
    constexpr int K = 4;
    std::vector<Dual<K>> in(Q), out(nout);
    for (int b0 = 0; b0 < Q; b0 += K) {
        for (int i = 0; i < Q; ++i) {
            in[i] = Dual<K>(qv[i]);
            if (i >= b0 && i < b0 + K) { in[i].d[i - b0] = 1.0; }
        }
        fn(in.data(), out.data());
        for (int o = 0; o < nout; ++o) {
            if (b0 == 0) { vals[o] = out[o].v; }
            for (int k = 0; k < K && b0 + k < Q; ++k) { J[static_cast<size_t>(o) * Q + b0 + k] = out[o].d[k]; }
        }
    }
    auto grad = [&](int o) { return std::vector<double>(J.begin() + static_cast<long>(o) * Q, J.begin() + static_cast<long>(o + 1) * Q); };
    L.p = qv[0];
    L.t = qv[1];
    L.sat[0] = qv[2];
    L.sat[1] = 1.0 - qv[2];
    for (int ph = 0; ph < 2; ++ph) {
        double sum = 0;
        for (int i = 0; i + 1 < n; ++i) { L.x[ph][i] = qv[x_col(n, ph, i)]; sum += L.x[ph][i]; }
        L.x[ph][n - 1] = 1.0 - sum;
        const int b = ph * stride;
        L.c[ph] = vals[b]; L.rho[ph] = vals[b + 1]; L.mu[ph] = vals[b + 2]; L.h[ph] = vals[b + 3]; L.u[ph] = vals[b + 4];
        L.dc[ph] = grad(b); L.drho[ph] = grad(b + 1); L.dmu[ph] = grad(b + 2); L.dh[ph] = grad(b + 3); L.du[ph] = grad(b + 4);
        // Corey kr = S^2 and its gradient (S_0 = q[2], S_1 = 1 - q[2])
        L.kr[ph] = L.sat[ph] * L.sat[ph];
        L.dkr[ph].assign(Q, 0.0);
        L.dkr[ph][2] = ph == 0 ? 2.0 * L.sat[0] : -2.0 * L.sat[1];
        // mobility, two_phase_transport.hpp:92-125
        const double mu = L.mu[ph], kr = L.kr[ph];
        if (!std::isfinite(mu) || !(mu > 0) || !std::isfinite(kr) || kr < 0) { throw RangeFail{}; }
        L.lam[ph] = kr / mu;
        L.dlam[ph].assign(Q, 0.0);
        const double mu2 = mu * mu;
        for (int col = 0; col < Q; ++col) {
            L.dlam[ph][col] = (L.dkr[ph][col] * mu - kr * L.dmu[ph][col]) / mu2;
            if (!std::isfinite(L.dlam[ph][col])) { throw RangeFail{}; }
        }
    }
    L.deq.assign(static_cast<size_t>(n) * Q, 0.0);
    for (int i = 0; i < n; ++i) {
        L.eq[i] = vals[2 * stride + i];
        for (int col = 0; col < Q; ++col) { L.deq[static_cast<size_t>(i) * Q + col] = J[static_cast<size_t>(2 * stride + i) * Q + col]; }
    }
    // component accumulation (two_phase_accumulation.hpp:18-160)
    auto dsat = [](int ph, int col) { return col != 2 ? 0.0 : (ph == 0 ? 1.0 : -1.0); };
    auto dcomp = [&](int ph, int i, int col) {
        const int first = x_col(n, ph, 0);
        if (col < first || col >= first + n - 1) { return 0.0; }
        if (i == n - 1) { return -1.0; }
        return col - first == i ? 1.0 : 0.0;
    };
    L.acc_total = 0;
    for (int i = 0; i < n; ++i) { L.acc[i] = 0; }
    for (int ph = 0; ph < 2; ++ph) {
        L.acc_total += porosity * L.sat[ph] * L.c[ph];
        for (int i = 0; i < n; ++i) { L.acc[i] += porosity * L.sat[ph] * L.c[ph] * L.x[ph][i]; }
    }
    L.dacc.assign(static_cast<size_t>(n) * Q, 0.0);
    L.dacc_total.assign(Q, 0.0);
    for (int col = 0; col < Q; ++col) {
        for (int ph = 0; ph < 2; ++ph) {
            const double s = L.sat[ph], ds = dsat(ph, col), c = L.c[ph], dc = L.dc[ph][col];
            L.dacc_total[col] += porosity * (ds * c + s * dc);
            for (int i = 0; i < n; ++i) {
                const double dx = dcomp(ph, i, col);
                L.dacc[static_cast<size_t>(i) * Q + col] += porosity * (ds * c * L.x[ph][i] + s * dc * L.x[ph][i] + s * c * dx);
            }
        }
    }
    // energy accumulation (two_phase_accumulation.hpp:184-338) with the rock model
    const double urock = md.cv_rock * (L.t - md.t0);
    double fluid = 0;
    for (int ph = 0; ph < 2; ++ph) { fluid += porosity * L.sat[ph] * L.rho[ph] * L.u[ph]; }
    const double rock_bulk = (1.0 - porosity) * urock;
    L.energy = fluid + rock_bulk;
    L.denergy.assign(Q, 0.0);
    for (int col = 0; col < Q; ++col) {
        const double durock = col == 1 ? md.cv_rock : 0.0;
        double g = (1.0 - porosity) * durock;
        for (int ph = 0; ph < 2; ++ph) {
            const double s = L.sat[ph], ds = dsat(ph, col), rho = L.rho[ph], drho = L.drho[ph][col];
            const double u = L.u[ph], du = L.du[ph][col];
            g += porosity * (ds * rho * u + s * drho * u + s * rho * du);
        }
        L.denergy[col] = g;
    }
}

struct AsmOut {
    double* residual; // [Q][ncell]
    double* diag;     // [Q*Q][ncell]
    double* offdiag;  // [2][(n+1)*Q][nface]
};

// Face: build_two_phase_potential_upwind_linearization (two_phase_transport.hpp:173-321)
// and build_two_phase_tpfa_face_linearization (two_phase_tpfa.hpp:129-667). Adds the
// owner and neighbour rows' contributions (rates / bulk volume) to res/diag and writes the
// two off-diagonal blocks.
void face_contribution(int n, const CellLin& o, const CellLin& nb, double tf, double gdot, double gth, double vo,
                       double vn, double* res_o, double* res_n, double* diag_o, double* diag_n, double* off_on,
                       double* off_no) {
    const int Q = q_count(n);
    auto dcomp = [&](int ph, int i, int col) {
        const int first = x_col(n, ph, 0);
        if (col < first || col >= first + n - 1) { return 0.0; }
        if (i == n - 1) { return -1.0; }
        return col - first == i ? 1.0 : 0.0;
    };
    std::vector<double> rate(n, 0.0), jo(static_cast<size_t>(n) * Q, 0.0), jn(static_cast<size_t>(n) * Q, 0.0);
    double erate = 0;
    std::vector<double> eo(Q, 0.0), en(Q, 0.0);
    for (int ph = 0; ph < 2; ++ph) {
        const double rf = 0.5 * (o.rho[ph] + nb.rho[ph]);
        const double delta = (nb.p - o.p) - rf * gdot;
        if (!std::isfinite(rf) || !(rf > 0) || !std::isfinite(delta)) { throw RangeFail{}; }
        std::vector<double> dphi_o(Q), dphi_n(Q);
        for (int col = 0; col < Q; ++col) {
            dphi_o[col] = -(col == 0 ? 1.0 : 0.0) - 0.5 * o.drho[ph][col] * gdot;
            dphi_n[col] = (col == 0 ? 1.0 : 0.0) - 0.5 * nb.drho[ph][col] * gdot;
        }
        const bool own = !(delta > 0.0); // owner on a negative potential and on the exact tie
        const double lam = own ? o.lam[ph] : nb.lam[ph];
        const double flux = -tf * lam * delta;
        if (!std::isfinite(flux)) { throw RangeFail{}; }
        std::vector<double> fo(Q), fn(Q);
        for (int col = 0; col < Q; ++col) {
            fo[col] = -tf * (lam * dphi_o[col] + delta * (own ? o.dlam[ph][col] : 0.0));
            fn[col] = -tf * (lam * dphi_n[col] + delta * (own ? 0.0 : nb.dlam[ph][col]));
        }
        const CellLin& up = own ? o : nb;
        const double c = up.c[ph];
        for (int i = 0; i < n; ++i) {
            const double content = c * up.x[ph][i];
            rate[i] += content * flux;
            for (int col = 0; col < Q; ++col) {
                double dco = 0.0, dcn = 0.0;
                if (own) { dco = o.dc[ph][col] * up.x[ph][i] + c * dcomp(ph, i, col); }
                else { dcn = nb.dc[ph][col] * up.x[ph][i] + c * dcomp(ph, i, col); }
                jo[static_cast<size_t>(i) * Q + col] += dco * flux + content * fo[col];
                jn[static_cast<size_t>(i) * Q + col] += dcn * flux + content * fn[col];
            }
        }
        const double rho = up.rho[ph], h = up.h[ph];
        const double ed = rho * h;
        erate += ed * flux;
        for (int col = 0; col < Q; ++col) {
            const double dedo = own ? o.drho[ph][col] * h + rho * o.dh[ph][col] : 0.0;
            const double dedn = own ? 0.0 : nb.drho[ph][col] * h + rho * nb.dh[ph][col];
            eo[col] += dedo * flux + ed * fo[col];
            en[col] += dedn * flux + ed * fn[col];
        }
    }
    const double cond = gth * (o.t - nb.t);
    const double etot = erate + cond;
    eo[1] += gth;
    en[1] -= gth;
    // owner row: +rate / V_o ; neighbour row: -rate / V_n (two_phase_tpfa.hpp:526-640)
    for (int i = 0; i < n; ++i) {
        res_o[i] += rate[i] / vo;
        res_n[i] += -rate[i] / vn;
        for (int col = 0; col < Q; ++col) {
            diag_o[i * Q + col] += jo[static_cast<size_t>(i) * Q + col] / vo;
            off_on[i * Q + col] = jn[static_cast<size_t>(i) * Q + col] / vo;
            off_no[i * Q + col] = -jo[static_cast<size_t>(i) * Q + col] / vn;
            diag_n[i * Q + col] += -jn[static_cast<size_t>(i) * Q + col] / vn;
        }
    }
    res_o[n] += etot / vo;
    res_n[n] += -etot / vn;
    for (int col = 0; col < Q; ++col) {
        diag_o[n * Q + col] += eo[col] / vo;
        off_on[n * Q + col] = en[col] / vo;
        off_no[n * Q + col] = eo[col] * (-1.0 / vn);
        diag_n[n * Q + col] += en[col] * (-1.0 / vn);
    }
}

AsmModel make_model(int n, const double* model) {
    AsmModel md{};
    md.mu_a0 = model[0];
    md.mu_a1 = model[1];
    md.cp = model[2];
    md.t0 = model[3];
    md.cv_rock = model[4];
    for (int i = 0; i < n; ++i) { md.molar_mass[i] = model[5 + i]; }
    return md;
}

template <typename Body>
void parallel_cells(int64_t ncell, int threads, Body body) {
    if (threads <= 0) { threads = static_cast<int>(std::max(1U, std::thread::hardware_concurrency())); }
    std::atomic<int64_t> next{0};
    constexpr int64_t chunk = 16;
    auto worker = [&]() {
        for (;;) {
            const int64_t begin = next.fetch_add(chunk);
            if (begin >= ncell) { return; }
            const int64_t end = std::min(ncell, begin + chunk);
            for (int64_t c = begin; c < end; ++c) { body(c); }
        }
    };
    std::vector<std::thread> pool;
    for (int k = 1; k < threads; ++k) { pool.emplace_back(worker); }
    worker();
    for (auto& th : pool) { th.join(); }
}

} // namespace
} // namespace cpuref

using namespace cpuref;

extern "C" int cpuref_flash_batch_hinted(int nc, const double* tc, const double* pc, const double* omega,
                                         const double* kij, int64_t ncell, const double* p, const double* t,
                                         const double* z, int nhint, const double* hints,
                                         double fugacity_tolerance, int max_phases, int threads, int32_t* status,
                                         int32_t* phase_count, double* beta, double* zfactor, int32_t* branch,
                                         double* x, double* ln_phi, uint32_t* counters);

extern "C" int cpuref_flash_batch(int nc, const double* tc, const double* pc, const double* omega,
                                  const double* kij, int64_t ncell, const double* p, const double* t,
                                  const double* z, double fugacity_tolerance, int max_phases, int threads,
                                  int32_t* status, int32_t* phase_count, double* beta, double* zfactor,
                                  int32_t* branch, double* x, double* ln_phi, uint32_t* counters) {
    return cpuref_flash_batch_hinted(nc, tc, pc, omega, kij, ncell, p, t, z, 0, nullptr, fugacity_tolerance,
                                     max_phases, threads, status, phase_count, beta, zfactor, branch, x, ln_phi,
                                     counters);
}

// Warm-start hints: hints[(h*nc + i)*ncell + c], h < nhint, appended to the initial stability
// starts (Pr76PtFlashBackendOptions::initial_starts); final_hints likewise, the extra starts
// of the common-tangent review (final_starts, PtSolveHints::final_two_phase_stability_starts).
// NaN in component 0: no hint h for cell c.
extern "C" int cpuref_flash_batch_warm(int nc, const double* tc, const double* pc, const double* omega,
                                       const double* kij, int64_t ncell, const double* p, const double* t,
                                       const double* z, int nhint, const double* hints, int nfinal,
                                       const double* final_hints, double fugacity_tolerance, int max_phases,
                                       int threads, int32_t* status, int32_t* phase_count, double* beta,
                                       double* zfactor, int32_t* branch, double* x, double* ln_phi,
                                       uint32_t* counters) {
    if (nhint < 0 || nhint > 6 || (nhint > 0 && !hints)) { return 1; }
    if (nfinal < 0 || nfinal > 6 || (nfinal > 0 && !final_hints)) { return 1; }
    if (nc < 1 || nc > NCMAX || ncell < 0 || (max_phases != 2 && max_phases != 3)) { return 1; }
    const Mixture m = make_mixture(nc, tc, pc, omega, kij);
    const Out o{status, phase_count, beta, zfactor, branch, x, ln_phi, counters};
    SplitOptions so;
    so.fug_tol = fugacity_tolerance;
    parallel_cells(ncell, threads, [&](int64_t c) {
        double feed[NCMAX];
        for (int i = 0; i < nc; ++i) { feed[i] = z[i * ncell + c]; }
        double hw[6][NCMAX], fw[6][NCMAX];
        const double* hs[6];
        const double* fs[6];
        int nh = 0, nf = 0;
        for (int h = 0; h < nhint; ++h) {
            if (std::isnan(hints[(h * nc + 0) * ncell + c])) { continue; }
            for (int i = 0; i < nc; ++i) { hw[nh][i] = hints[(h * nc + i) * ncell + c]; }
            hs[nh] = hw[nh];
            ++nh;
        }
        for (int h = 0; h < nfinal; ++h) {
            if (std::isnan(final_hints[(h * nc + 0) * ncell + c])) { continue; }
            for (int i = 0; i < nc; ++i) { fw[nf][i] = final_hints[(h * nc + i) * ncell + c]; }
            fs[nf] = fw[nf];
            ++nf;
        }
        Counters cnt;
        auto M = std::make_unique<Max3Result>();
        try {
            if (max_phases == 2) {
                solve_vle(m, p[c], t[c], feed, so, hs, nh, M->base, cnt, fs, nf);
                publish_vle(o, nc, ncell, c, M->base, 0xF, cnt);
            } else {
                solve_max3(m, p[c], t[c], feed, so, *M, cnt, hs, nh, fs, nf);
                publish_max3(o, nc, ncell, c, *M, cnt);
            }
        } catch (const Rejected&) {
            Phase ph[3];
            write_cell(o, nc, ncell, c, 2, 0xF, 0xF, 0, ph, cnt, 0x2000);
        }
    });
    return 0;
}

// solve_pr76_pt_continuation (pr76_pt_continuation.hpp:194-310) for npath paths at once: path j
// has feed z[i*npath + j] and ordered states p, t [npoint][npath]. Point k of every path is a
// max3 flash of the path feed normalised once. Its initial stability starts are the previous
// accepted point's phase compositions (carry bit 0) then the nfi fallback initial starts
// fi[h*nc + i], its final-review starts the same compositions then the nff fallback final
// starts, each list without repeats; its three-phase starts are the previous point's
// three-phase state (carry bit 1) then, when has_start, the fixed start start[ph*nc + i]
// with seed[2]. Records are npoint consecutive flash records of npath cells; info[k*npath +
// j] = incoming hint kind (0 none, else the previous accepted phase count) | carried start
// count << 4 | carried three-phase start << 8; status flag 0x20000 marks a consumed carried
// start. A path whose states fail the path check, or one of whose points the flash rejects,
// is rejected at every point (the reference throws for the whole path).
extern "C" int cpuref_pt_continuation(int nc, const double* tc, const double* pc, const double* omega,
                                      const double* kij, int64_t npath, int npoint, const double* p, const double* t,
                                      const double* z, int carry, int nfi, const double* fi, int nff,
                                      const double* ff, int has_start, const double* start, const double* seed,
                                      double fugacity_tolerance, int threads, int32_t* status, int32_t* phase_count,
                                      double* beta, double* zfactor, int32_t* branch, double* x, double* ln_phi,
                                      uint32_t* counters, int32_t* info) {
    if (nc < 1 || nc > NCMAX || npath < 0 || npoint < 1 || npoint > 4096) { return 1; }
    if (nfi < 0 || nfi > 3 || nff < 0 || nff > 3 || (nfi > 0 && !fi) || (nff > 0 && !ff) ||
        (has_start && (!start || !seed))) {
        return 1;
    }
    const Mixture m = make_mixture(nc, tc, pc, omega, kij);
    SplitOptions so;
    so.fug_tol = fugacity_tolerance;
    const int64_t N = npath;
    auto out_at = [&](int k) {
        return Out{status + k * N, phase_count + k * N, beta + k * 3 * N, zfactor + k * 3 * N, branch + k * 3 * N,
                   x + k * 3 * nc * N, ln_phi + k * 3 * nc * N, counters ? counters + k * 4 * N : nullptr};
    };
    parallel_cells(npath, threads, [&](int64_t j) {
        double feed[NCMAX];
        for (int i = 0; i < nc; ++i) { feed[i] = z[i * npath + j]; }
        bool path_ok = true;
        for (int k = 0; k < npoint; ++k) {
            const double pk = p[k * N + j], tk = t[k * N + j];
            if (!(std::isfinite(pk) && pk > 0 && std::isfinite(tk) && tk > 0)) { path_ok = false; }
        }
        // the path's feed is normalised once here and again by each point's stability test
        // (pr76_pt_continuation.hpp:218-219)
        try {
            const double sum = check_composition(feed, nc);
            for (int i = 0; i < nc; ++i) { feed[i] /= sum; }
        } catch (const Rejected&) { path_ok = false; }
        int prev_count = 0;           // accepted phase count of the previous point (0: none)
        double prev[3][NCMAX];
        double prev_frac[2] = {0, 0};
        bool prev_three = false;
        // pr76_pt_continuation_append_unique (pr76_pt_continuation.hpp:95-107)
        auto append = [&](const double** list, int& count, const double* v) {
            for (int e = 0; e < count; ++e) {
                bool same = true;
                for (int i = 0; i < nc; ++i) { same = same && list[e][i] == v[i]; }
                if (same) { return; }
            }
            list[count++] = v;
        };
        for (int k = 0; k < npoint && path_ok; ++k) {
            const Out o = out_at(k);
            const double* initial[6];
            const double* finals[6];
            int ni = 0, nf = 0;
            if (carry & 1) {
                for (int h = 0; h < prev_count; ++h) {
                    append(initial, ni, prev[h]);
                    append(finals, nf, prev[h]);
                }
            }
            for (int h = 0; h < nfi; ++h) { append(initial, ni, fi + h * nc); }
            for (int h = 0; h < nff; ++h) { append(finals, nf, ff + h * nc); }
            const bool three = (carry & 2) && prev_three;
            double st3[2 * 3 * NCMAX], sf3[4];
            int ns = 0;
            if (three) {
                for (int ph = 0; ph < 3; ++ph) { for (int i = 0; i < nc; ++i) { st3[ph * nc + i] = prev[ph][i]; } }
                sf3[0] = prev_frac[0];
                sf3[1] = prev_frac[1];
                ++ns;
            }
            if (has_start) {
                for (int q = 0; q < 3 * nc; ++q) { st3[ns * 3 * nc + q] = start[q]; }
                sf3[2 * ns] = seed[0];
                sf3[2 * ns + 1] = seed[1];
                ++ns;
            }
            info[k * N + j] = prev_count | (((carry & 1) ? prev_count : 0) << 4) | (three ? 0x100 : 0);
            Counters cnt;
            auto M = std::make_unique<Max3Result>();
            try {
                solve_max3(m, p[k * N + j], t[k * N + j], feed, so, *M, cnt, initial, ni, finals, nf,
                           ns ? st3 : nullptr, ns ? sf3 : nullptr, ns);
            } catch (const Rejected&) {
                path_ok = false;
                break;
            }
            // the consumed flag reports the carried start only
            // (carried_three_phase_start_consumed, pr76_pt_continuation.hpp:256-259)
            M->start_consumed = M->start_consumed && three;
            publish_max3(o, nc, N, j, *M, cnt);
            prev_three = false;
            switch (M->status) {
            case M_SINGLE:
                prev_count = 1;
                for (int i = 0; i < nc; ++i) { prev[0][i] = M->base.init.feed[i]; }
                break;
            case M_TWO: {
                const SplitState& pr = (M->use_neighbor ? M->neighbor : M->base).cand.pt;
                prev_count = 2;
                for (int i = 0; i < nc; ++i) { prev[0][i] = pr.fr.x[i]; prev[1][i] = pr.fr.y[i]; }
                break;
            }
            case M_THREE:
                prev_count = 3;
                prev_three = true;
                for (int ph = 0; ph < 3; ++ph) { for (int i = 0; i < nc; ++i) { prev[ph][i] = M->three.comp[ph][i]; } }
                prev_frac[0] = M->three.frac[1];
                prev_frac[1] = M->three.frac[2];
                break;
            default:
                prev_count = 0;
                break;
            }
        }
        if (!path_ok) {
            for (int k = 0; k < npoint; ++k) {
                Phase ph[3];
                Counters cnt;
                write_cell(out_at(k), nc, N, j, 2, 0xF, 0xF, 0, ph, cnt, 0x2000);
                info[k * N + j] = 0;
            }
        }
    });
    return 0;
}

extern "C" int cpuref_flash_batch_hinted(int nc, const double* tc, const double* pc, const double* omega,
                                         const double* kij, int64_t ncell, const double* p, const double* t,
                                         const double* z, int nhint, const double* hints,
                                         double fugacity_tolerance, int max_phases, int threads, int32_t* status,
                                         int32_t* phase_count, double* beta, double* zfactor, int32_t* branch,
                                         double* x, double* ln_phi, uint32_t* counters) {
    return cpuref_flash_batch_warm(nc, tc, pc, omega, kij, ncell, p, t, z, nhint, hints, 0, nullptr,
                                   fugacity_tolerance, max_phases, threads, status, phase_count, beta, zfactor,
                                   branch, x, ln_phi, counters);
}

extern "C" int cpuref_phase_batch(int nc, const double* tc, const double* pc, const double* omega,
                                  const double* kij, int64_t ncell, const double* p, const double* t,
                                  const double* x, int32_t* root_status, int32_t* root_count, double* zroot,
                                  int32_t* root_flags, double* ln_phi) {
    if (nc < 1 || nc > NCMAX || ncell < 0) { return 1; }
    const Mixture m = make_mixture(nc, tc, pc, omega, kij);
    for (int64_t c = 0; c < ncell; ++c) {
        double w[NCMAX];
        for (int i = 0; i < nc; ++i) { w[i] = x[i * ncell + c]; }
        root_count[c] = 0;
        for (int k = 0; k < 3; ++k) {
            zroot[k * ncell + c] = NaN;
            root_flags[k * ncell + c] = 0;
            for (int i = 0; i < nc; ++i) { ln_phi[(k * nc + i) * ncell + c] = NaN; }
        }
        PhaseCtx ctx;
        try {
            const int e = prepare(m, p[c], t[c], w, ctx);
            if (e == ROOT_RANGE && ctx.rs.status == RS_SUCCESS) { root_status[c] = RS_UNREPRESENTABLE; continue; }
            root_status[c] = ctx.rs.status;
            if (e != NONE) { continue; }
        } catch (const Rejected&) {
            return 2;
        }
        root_count[c] = ctx.rs.count;
        for (int k = 0; k < ctx.rs.count; ++k) {
            const Root& r = ctx.rs.r[k];
            zroot[k * ncell + c] = r.z;
            root_flags[k * ncell + c] = (r.slope_sign + 1) | (r.dvalid ? 0x10 : 0);
            double lp[NCMAX];
            if (lnphi(m, p[c], t[c], ctx, k, lp, nullptr) == NONE) {
                for (int i = 0; i < nc; ++i) { ln_phi[(k * nc + i) * ncell + c] = lp[i]; }
            }
        }
    }
    return 0;
}

extern "C" int cpuref_rachford_rice(int nc, const double* z, const double* log_k, double* beta, double* x,
                                    double* y, int32_t* iterations) {
    if (nc < 1 || nc > NCMAX) { return -1; }
    RRResult r;
    Counters cnt;
    try {
        rachford_rice(nc, z, log_k, r, cnt);
    } catch (const Rejected&) {
        return -2;
    }
    *beta = r.beta;
    for (int i = 0; i < nc; ++i) { x[i] = r.x[i]; y[i] = r.y[i]; }
    *iterations = r.iterations;
    return r.status;
}

extern "C" int cpuref_closure_batch(int nc, const double* tc, const double* pc, const double* omega,
                                    const double* kij, int64_t ncell, const double* p, const double* t,
                                    const double* z, double fugacity_tolerance, double min_fraction,
                                    double min_rcond, int threads, int32_t* status, double* primal, double* jac,
                                    double* log_k_jac, double* diag) {
    if (nc < 2 || nc > NCMAX || ncell < 0) { return 1; }
    if (!std::isfinite(min_fraction) || !(min_fraction > 0) || !(min_fraction < 0.5) ||
        !std::isfinite(min_rcond) || min_rcond < 0 || min_rcond >= 1) { return 1; }
    const Mixture m = make_mixture(nc, tc, pc, omega, kij);
    SplitOptions so;
    so.fug_tol = fugacity_tolerance;
    const int q = nc + 1;
    const int rows = 2 * nc + 5;
    parallel_cells(ncell, threads, [&](int64_t c) {
        double feed[NCMAX];
        for (int i = 0; i < nc; ++i) { feed[i] = z[i * ncell + c]; }
        Counters cnt;
        auto V = std::make_unique<VleResult>();
        auto R = std::make_unique<ClosureResult>();
        try {
            solve_vle(m, p[c], t[c], feed, so, nullptr, 0, *V, cnt);
            build_closure(m, p[c], t[c], *V, so, min_fraction, min_rcond, *R);
        } catch (...) {
            R = std::make_unique<ClosureResult>();
            R->thrown = true;
        }
        status[c] = R->thrown ? (CP_INDETERMINATE | (1 << 4) | 0x2000) : (R->primal | (R->lin << 4) | (R->reason << 8));
        primal[0 * ncell + c] = R->zl;
        primal[1 * ncell + c] = R->zv;
        primal[2 * ncell + c] = R->rhol;
        primal[3 * ncell + c] = R->rhov;
        for (int k = 0; k < rows * q; ++k) { jac[k * ncell + c] = R->have_jac ? R->jac[k] : NaN; }
        if (log_k_jac) {
            for (int k = 0; k < nc * q; ++k) { log_k_jac[k * ncell + c] = R->have_dlk ? R->dlk[k] : NaN; }
        }
        if (diag) {
            diag[0 * ncell + c] = R->rcond;
            diag[1 * ncell + c] = R->backward;
            diag[2 * ncell + c] = R->resid;
        }
    });
    return 0;
}

// TwoPhaseSnesAssemblyContext3D::evaluate_complete_assembly
// (two_phase_snes_assembly.hpp:550-1090) on one rank: every cell's linearization, the
// backward-Euler accumulation residuals, every face's TPFA contributions, then the
// fugacity-equilibrium rows. Row order per cell: components, energy, equilibrium
// (two_phase_cell_state.hpp:198-222). Returns 0, or 2 when any cell raised a domain error
// (the reference's global domain_error status; nothing is assembled then).
extern "C" int cpuref_assembly(int nc, const double* tc, const double* pc, const double* omega, const double* kij,
                               const double* model, int64_t ncell, const double* q, const int32_t* branch,
                               const double* bulk, const double* poro, const double* prev, double dt, int64_t nface,
                               const int32_t* owner, const int32_t* neighbour, const double* trans, const double* gdot,
                               const double* gth, int threads, int32_t* cell_status, double* residual, double* diag,
                               double* offdiag) {
    if (nc < 2 || nc > NCMAX || ncell < 0 || nface < 0 || !(dt > 0)) { return 1; }
    const Mixture m = make_mixture(nc, tc, pc, omega, kij);
    const AsmModel md = make_model(nc, model);
    const int Q = q_count(nc);
    std::vector<CellLin> cells(static_cast<size_t>(ncell));
    std::atomic<int> domain{0};
    parallel_cells(ncell, threads, [&](int64_t c) {
        double qv[2 * NCMAX + 1];
        for (int k = 0; k < Q; ++k) { qv[k] = q[k * ncell + c]; }
        const int br[2] = {branch[c], branch[ncell + c]};
        cell_status[c] = 0;
        try {
            cell_linearization(m, md, qv, br, poro[c], cells[c]);
        } catch (...) {
            cell_status[c] = 1;
            domain = 1;
        }
    });
    if (domain) { return 2; }
    const double inv_dt = 1.0 / dt;
    for (int64_t c = 0; c < ncell; ++c) {
        const CellLin& L = cells[c];
        for (int i = 0; i < nc; ++i) {
            residual[i * ncell + c] = (L.acc[i] - prev[i * ncell + c]) * inv_dt;
            for (int col = 0; col < Q; ++col) { diag[(i * Q + col) * ncell + c] = L.dacc[static_cast<size_t>(i) * Q + col] * inv_dt; }
        }
        residual[nc * ncell + c] = (L.energy - prev[nc * ncell + c]) / dt;
        for (int col = 0; col < Q; ++col) { diag[(nc * Q + col) * ncell + c] = L.denergy[col] / dt; }
        for (int i = 0; i < nc; ++i) {
            residual[(nc + 1 + i) * ncell + c] = L.eq[i];
            for (int col = 0; col < Q; ++col) {
                diag[((nc + 1 + i) * Q + col) * ncell + c] = L.deq[static_cast<size_t>(i) * Q + col];
            }
        }
    }
    const int rows = nc + 1;
    std::vector<double> ro(Q), rn(Q), dgo(static_cast<size_t>(Q) * Q), dgn(static_cast<size_t>(Q) * Q);
    std::vector<double> on(static_cast<size_t>(rows) * Q), no(static_cast<size_t>(rows) * Q);
    for (int64_t f = 0; f < nface; ++f) {
        const int64_t a = owner[f], b = neighbour[f];
        std::fill(ro.begin(), ro.end(), 0.0);
        std::fill(rn.begin(), rn.end(), 0.0);
        std::fill(dgo.begin(), dgo.end(), 0.0);
        std::fill(dgn.begin(), dgn.end(), 0.0);
        try {
            face_contribution(nc, cells[a], cells[b], trans[f], gdot[f], gth[f], bulk[a], bulk[b], ro.data(), rn.data(),
                              dgo.data(), dgn.data(), on.data(), no.data());
        } catch (...) {
            return 3;
        }
        for (int r = 0; r < rows; ++r) {
            residual[r * ncell + a] += ro[r];
            residual[r * ncell + b] += rn[r];
            for (int col = 0; col < Q; ++col) {
                diag[(r * Q + col) * ncell + a] += dgo[r * Q + col];
                diag[(r * Q + col) * ncell + b] += dgn[r * Q + col];
                offdiag[(static_cast<int64_t>(0 * rows * Q) + r * Q + col) * nface + f] = on[r * Q + col];
                offdiag[(static_cast<int64_t>(1 * rows * Q) + r * Q + col) * nface + f] = no[r * Q + col];
            }
        }
    }
    return 0;
}

// Pore-volume component and energy inventories N_i, E at a state (the previous-time
// snapshots of build_backward_euler_*_accumulation_residual): acc[(nc+1)][ncell].
extern "C" int cpuref_accumulation(int nc, const double* tc, const double* pc, const double* omega,
                                   const double* kij, const double* model, int64_t ncell, const double* q,
                                   const int32_t* branch, const double* poro, int threads, double* acc) {
    if (nc < 2 || nc > NCMAX || ncell < 0) { return 1; }
    const Mixture m = make_mixture(nc, tc, pc, omega, kij);
    const AsmModel md = make_model(nc, model);
    const int Q = q_count(nc);
    std::atomic<int> domain{0};
    parallel_cells(ncell, threads, [&](int64_t c) {
        double qv[2 * NCMAX + 1];
        for (int k = 0; k < Q; ++k) { qv[k] = q[k * ncell + c]; }
        const int br[2] = {branch[c], branch[ncell + c]};
        CellLin L;
        try {
            cell_linearization(m, md, qv, br, poro[c], L);
            for (int i = 0; i < nc; ++i) { acc[i * ncell + c] = L.acc[i]; }
            acc[nc * ncell + c] = L.energy;
        } catch (...) {
            domain = 1;
        }
    });
    return domain ? 2 : 0;
}
