// PR76 PT flash on gfx950, one cell per lane: multistart TPD stability, Rachford-Rice,
// logK-SSI two-phase split with common-tangent review, three-phase SSI with the
// generalized RR Newton, max3 orchestration and phase-set publication.
//
// Decision rules, thresholds and arithmetic order follow the reference so that phase
// counts and labels agree cell by cell:
//   stability  modules/flash/include/mpmc/flash/pt_stability.hpp:178-616
//   RR         modules/flash/include/mpmc/flash/rachford_rice.hpp:42-147
//   split      modules/flash/include/mpmc/flash/pt_split.hpp:204-525
//   3-phase    modules/flash/include/mpmc/flash/pt_three_phase.hpp:180-674
//   max3       modules/flash/include/mpmc/flash/pr76_three_phase.hpp:248-561
//   publish    modules/flash/include/mpmc/flash/pt_phase_set.hpp:90-150,
//              pt_vle_phase_set.hpp:169-258
// Fixed capacity replaces the reference's vectors: NC components per lane and at most
// NC+2+4 stability trials (automatic starts plus the review's extra starts).
#pragma once

#include "pr76_device.hpp"

namespace mpmc::dev {

struct Counters {
    uint32_t prop, tpd, rr, rr3;
};

template <int NC>
__device__ __forceinline__ double ksum(const double* v) {
    double s = 0, c = 0;
#pragma unroll
    for (int i = 0; i < NC; ++i) { kadd(v[i], s, c); }
    return s;
}

// stability_check_composition (pt_stability.hpp:265-277); returns the compensated sum or
// sets `rejected` for the reference's domain_error.
template <int NC>
__device__ __forceinline__ double check_composition(const double* w, bool& rejected) {
#pragma unroll
    for (int i = 0; i < NC; ++i) { if (!finite(w[i]) || w[i] < 0 || w[i] > 1) { rejected = true; } }
    const double s = ksum<NC>(w);
    if (__builtin_fabs(s - 1.0) > 64.0 * kEps) { rejected = true; }
    return s;
}

// ------------------------------------------------------------------ TPD / stability
enum TrialStatus : int { T_NEGATIVE = 0, T_STATIONARY, T_ITERATION_LIMIT, T_EVALUATION_LIMIT,
                         T_LINE_SEARCH_FAILED, T_PROPERTY_FAILURE, T_NONSMOOTH, T_UNREPRESENTABLE };
enum StabStatus : int { S_UNSTABLE = 0, S_NO_INSTABILITY, S_INDETERMINATE };

template <int NC>
struct TpdPoint {
    double w[NC], res[NC];
    double value, guard, stat;
    int branch;
    bool smooth;
};

template <int NC>
struct Witness {
    double w[NC];
    int branch;
};

template <int NC>
struct StabOut {
    static constexpr int kMaxWitness = NC + 6;
    int status;
    bool has_ref;
    Activity<NC> ref;
    double feed[NC];
    double tpd_tol;
    int nwit;
    Witness<NC> wit[kMaxWitness];
};

// tangent_plane_distance (pt_stability.hpp:333-375)
template <int NC>
MPMC_HOT int tpd_point(const double* w, const double* z, const Activity<NC>& trial,
                                      const Activity<NC>& ref, TpdPoint<NC>& r, bool& rejected) {
    const double sum = check_composition<NC>(w, rejected);
    (void)check_composition<NC>(z, rejected);
    if (rejected) { return REJECTED; }
    bool fin = true;
#pragma unroll
    for (int i = 0; i < NC; ++i) { fin = fin && finite(trial.lnphi[i]) && finite(ref.lnphi[i]); }
    if (!fin) { return NONFINITE; }
    r.value = 0; r.stat = 0; r.branch = trial.branch; r.smooth = trial.smooth;
    double corr = 0, mag = 1, mc = 0;
    bool boundary = false;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        r.w[i] = w[i];
        r.res[i] = 0;
        if (z[i] == 0) { if (w[i] != 0) { rejected = true; } continue; }
        if (w[i] == 0) { boundary = true; continue; }
        const double lw = log(w[i]), lz = log(z[i]);
        const double lr = __builtin_fabs(w[i] - z[i]) < 0.5 * z[i] ? log1p((w[i] - z[i]) / z[i]) : lw - lz;
        const double q = lr + (trial.lnphi[i] - ref.lnphi[i]);
        r.res[i] = q;
        kadd(w[i] * q, r.value, corr);
        kadd(w[i] * (__builtin_fabs(lw) + __builtin_fabs(lz) + __builtin_fabs(trial.lnphi[i]) +
                     __builtin_fabs(ref.lnphi[i])), mag, mc);
    }
    if (rejected) { return REJECTED; }
    const double mean = r.value / sum;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        if (z[i] == 0) { continue; }
        r.res[i] = w[i] == 0 ? -kInf : r.res[i] - mean;
        r.stat = smax(r.stat, __builtin_fabs(r.res[i]));
    }
    r.guard = 256.0 * kEps * mag;
    if (!finite(r.value) || !finite(r.guard) || (!boundary && !finite(r.stat))) { return NONFINITE; }
    return ISSUE_NONE;
}

__device__ __forceinline__ bool tpd_negative(double value, double guard, double tol) { return value < -tol - guard; }

// Multistart TPD search (detail::test_pt_stability_impl, pt_stability.hpp:383-616). With
// imposed == nullptr the reference phase is the least-Gibbs feed phase; otherwise the
// imposed common log-activities define it (test_pt_stability_against, :637-644).
// Negative-TPD witnesses are kept, in trial order, for the split and max3 seeds.
template <int NC, bool kN = false>
__device__ __noinline__ void stability(const Mix<NC>& m, double p, double t, const double* feed_in,
                                       const double* imposed, const double* const* extra, int nextra,
                                       double tpd_tol, StabOut<NC>& R, Counters& cnt, bool& rejected) {
    constexpr double kStatTol = 1e-8, kArmijo = 1e-4, kLogStep = 4.0;
    constexpr int kMaxIt = 512, kMaxBt = 32;
    constexpr uint32_t kMaxEval = 100000;
    R.status = S_INDETERMINATE;
    R.has_ref = false;
    R.nwit = 0;
    R.tpd_tol = tpd_tol;
    if (!finite(p) || p <= 0 || !finite(t) || t <= 0) { rejected = true; return; }
    const double isum = check_composition<NC>(feed_in, rejected);
    if (imposed) {
#pragma unroll
        for (int i = 0; i < NC; ++i) { if (!finite(imposed[i])) { rejected = true; } }
    }
    int active = 0;
#pragma unroll
    for (int i = 0; i < NC; ++i) { if (feed_in[i] > 0) { ++active; } }
    for (int e = 0; e < nextra; ++e) {
        (void)check_composition<NC>(extra[e], rejected);
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            if ((feed_in[i] == 0 && extra[e][i] != 0) || (feed_in[i] > 0 && extra[e][i] == 0)) { rejected = true; }
        }
    }
    if (rejected) { return; }
    bool lost = false;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        R.feed[i] = feed_in[i] / isum;
        lost = lost || (feed_in[i] > 0 && R.feed[i] == 0);
    }
    if (lost) { return; }
    const int ntrials = (active > 1 ? active + 2 : 1) + nextra;
    uint32_t evaluations = 0;
    Activity<NC> imp;
    if (!imposed) {
        ++evaluations; ++cnt.prop;
        const int e = stability_provider<NC, kN>(m, p, t, R.feed, R.ref);
        if (e == REJECTED) { rejected = true; return; }
        if (e != ISSUE_NONE) { return; } // every trial: property_failure
        R.has_ref = true;
    } else {
        bool fin = true;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            imp.lnphi[i] = R.feed[i] > 0 ? imposed[i] - log(R.feed[i]) : 0.0;
            fin = fin && finite(imp.lnphi[i]);
        }
        imp.branch = 0;
        imp.smooth = true;
        if (!fin) { return; }
    }
    const Activity<NC>& dref = imposed ? imp : R.ref;
    bool neg_found = false, all_stat = true;
    for (int k = 0; k < ntrials; ++k) {
        // start k, normalised as add_start does (pt_stability.hpp:442-466)
        double w0[NC];
        if (k == 0) {
#pragma unroll
            for (int i = 0; i < NC; ++i) { w0[i] = R.feed[i]; }
        } else if (active > 1 && k == 1) {
#pragma unroll
            for (int i = 0; i < NC; ++i) { w0[i] = R.feed[i] > 0 ? 1.0 / static_cast<double>(active) : 0.0; }
        } else if (active > 1 && k < active + 2) {
            int target = k - 2, seen = 0, vert = -1;
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                if (R.feed[i] != 0) { if (seen == target) { vert = i; } ++seen; }
            }
#pragma unroll
            for (int i = 0; i < NC; ++i) { w0[i] = R.feed[i] * 0.1 + (i == vert ? 0.9 : 0.0); }
        } else {
            const double* src = extra[k - (active > 1 ? active + 2 : 1)];
#pragma unroll
            for (int i = 0; i < NC; ++i) { w0[i] = src[i]; }
        }
        {
            const double s = ksum<NC>(w0);
#pragma unroll
            for (int i = 0; i < NC; ++i) { w0[i] /= s; }
        }
        bool support_lost = false;
#pragma unroll
        for (int i = 0; i < NC; ++i) { support_lost = support_lost || (R.feed[i] > 0 && w0[i] == 0); }
        if (support_lost) { all_stat = false; continue; }
        if (evaluations >= kMaxEval) { all_stat = false; continue; }
        TpdPoint<NC> pt;
        int status;
        {
            ++evaluations; ++cnt.prop; ++cnt.tpd;
            Activity<NC> ph;
            int e = stability_provider<NC, kN>(m, p, t, w0, ph);
            if (e == ISSUE_NONE) { e = tpd_point<NC>(w0, R.feed, ph, dref, pt, rejected); }
            if (e == REJECTED || rejected) { rejected = true; return; }
            if (e != ISSUE_NONE) { all_stat = false; continue; }
        }
        int iterations = 0;
        for (;;) {
            if (tpd_negative(pt.value, pt.guard, tpd_tol)) { status = T_NEGATIVE; neg_found = true; break; }
            if (!pt.smooth) { status = T_NONSMOOTH; break; }
            if (pt.stat <= kStatTol) { status = T_STATIONARY; break; }
            if (iterations >= kMaxIt) { status = T_ITERATION_LIMIT; break; }
            const double scale = smax(1.0, pt.stat / kLogStep);
            double dir[NC], lw[NC];
            double slope = 0, sc = 0;
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                dir[i] = 0;
                lw[i] = 0;
                if (R.feed[i] == 0) { continue; }
                dir[i] = -pt.res[i] / scale;
                kadd((pt.w[i] * pt.res[i]) * dir[i], slope, sc);
                lw[i] = log(pt.w[i]);
            }
            if (!finite(slope) || !(slope < 0)) { status = T_LINE_SEARCH_FAILED; break; }
            bool accepted = false;
            double alpha = 1;
            status = T_LINE_SEARCH_FAILED;
            for (int bt = 0; bt < kMaxBt; ++bt, alpha *= 0.5) {
                double cand[NC];
                double largest = -kInf;
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    if (R.feed[i] > 0) {
                        cand[i] = lw[i] + alpha * dir[i];
                        largest = smax(largest, cand[i]);
                    }
                }
#pragma unroll
                for (int i = 0; i < NC; ++i) { cand[i] = R.feed[i] > 0 ? exp(cand[i] - largest) : 0.0; }
                const double s = ksum<NC>(cand);
                bool positive = finite(s) && s > 0;
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    cand[i] /= s;
                    positive = positive && (R.feed[i] == 0 || cand[i] > 0);
                }
                if (!positive) { continue; }
                if (evaluations >= kMaxEval) { status = T_EVALUATION_LIMIT; break; }
                TpdPoint<NC> nx;
                ++evaluations; ++cnt.prop; ++cnt.tpd;
                Activity<NC> ph;
                int e = stability_provider<NC, kN>(m, p, t, cand, ph);
                if (e == ISSUE_NONE) { e = tpd_point<NC>(cand, R.feed, ph, dref, nx, rejected); }
                if (e == REJECTED || rejected) { rejected = true; return; }
                if (e != ISSUE_NONE) { continue; }
                bool ok = tpd_negative(nx.value, nx.guard, tpd_tol);
                if (!ok) {
                    // stability_accept_step, pt_stability.hpp:312-326
                    const double pred = -alpha * slope;
                    const double g = pt.guard + nx.guard;
                    if (pred <= g) { ok = nx.value <= pt.value + g && alpha / scale > 0 && nx.stat < pt.stat; }
                    else { ok = nx.value <= pt.value - kArmijo * pred; }
                }
                if (ok) { pt = nx; accepted = true; break; }
            }
            if (!accepted) { break; }
            ++iterations;
        }
        if (status == T_NEGATIVE && R.nwit < StabOut<NC>::kMaxWitness) {
#pragma unroll
            for (int i = 0; i < NC; ++i) { R.wit[R.nwit].w[i] = pt.w[i]; }
            R.wit[R.nwit].branch = pt.branch;
            ++R.nwit;
        }
        all_stat = all_stat && status == T_STATIONARY;
    }
    R.status = neg_found ? S_UNSTABLE : (all_stat ? S_NO_INSTABILITY : S_INDETERMINATE);
}

// ------------------------------------------------------------------ Rachford-Rice
enum RRStatus : int { RR_INTERIOR = 0, RR_NO_INTERIOR, RR_DEGENERATE, RR_ITERATION_LIMIT, RR_UNREPRESENTABLE };

template <int NC>
struct RRState {
    double beta;
    double x[NC], y[NC];
    double mass_abs, mass_rel;
};

// rr_value in the reference's arithmetic order (rachford_rice.hpp:53-63), with the
// per-component exp(-|lk|) and expm1 numerators precomputed (same values).
template <int NC>
__device__ __forceinline__ double rr_value_exact(double beta, const double* z, const double* lk, const double* e,
                                                 const double* num) {
    double v = 0, c = 0;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        if (z[i] == 0) { continue; }
        const double den = lk[i] >= 0 ? beta + (1 - beta) * e[i] : (1 - beta) + beta * e[i];
        kadd((z[i] / den) * num[i], v, c);
    }
    return v;
}

struct RRBracket {
    bool ok;
    double lo, hi, a, s2;
};

// Enclosure of the exact root of the RR function used to skip bisection evaluations.
template <int NC>
__device__ __forceinline__ RRBracket rr_enclose(const double* z, const double* lk, const double* e,
                                                const double* num, double guess) {
    RRBracket br{false, 0, 1, 0, 0};
    if (!(guess >= 0)) { return br; } // negative or NaN guess: plain bisection
    double a = 0, s2 = 0;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        if (z[i] > 0) {
            a += z[i] * __builtin_fabs(num[i]);
            s2 += z[i] * num[i] * num[i];
        }
    }
    if (!(a > 1e-200) || !(s2 > 1e-200) || !finite(a) || !finite(s2)) { return br; }
    br.a = a * (1.0 + 64 * kEps);  // upper bound on the rounded sum
    br.s2 = s2 * (1.0 - 64 * kEps); // lower bound
    double b = (guess > 0 && guess < 1) ? guess : 0.5, lo = 0, hi = 1;
    for (int it = 0; it < 48; ++it) {
        double f = 0, fp = 0;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            if (z[i] == 0) { continue; }
            const double den = lk[i] >= 0 ? b + (1 - b) * e[i] : (1 - b) + b * e[i];
            const double q = num[i] / den;
            f += z[i] * q;
            fp -= z[i] * q * q;
        }
        if (!finite(f) || !finite(fp) || !(fp < 0)) { return br; }
        if (f > 0) { lo = b; } else if (f < 0) { hi = b; } else { break; }
        double bn = b - f / fp;
        if (!(bn > lo && bn < hi)) { bn = 0.5 * (lo + hi); }
        const double step = __builtin_fabs(bn - b);
        b = bn;
        if (step <= 2 * kEps * smin(b, 1 - b)) { break; }
    }
    const double dm = smin(b, 1 - b);
    if (!(dm > 1e-150)) { return br; }
    const double delta = 4.0 * (32.0 * kEps * br.a / (0.5 * dm)) / br.s2 + 4 * kEps * dm;
    const double blo = b - delta, bhi = b + delta;
    if (!(blo > 0.5 * b) || !(bhi < b + 0.5 * (1 - b))) { return br; }
    const double flo = rr_value_exact<NC>(blo, z, lk, e, num);
    const double fhi = rr_value_exact<NC>(bhi, z, lk, e, num);
    const double elo = 32.0 * kEps * br.a / smin(blo, 1 - blo);
    const double ehi = 32.0 * kEps * br.a / smin(bhi, 1 - bhi);
    if (!(flo > elo) || !(fhi < -ehi)) { return br; }
    br.ok = true;
    br.lo = blo;
    br.hi = bhi;
    return br;
}

template <int NC>
__device__ __forceinline__ double rr_log_sum(const double* z, const double* lk, double sign) {
    double largest = -kInf;
#pragma unroll
    for (int i = 0; i < NC; ++i) { if (z[i] > 0) { largest = smax(largest, log(z[i]) + sign * lk[i]); } }
    double s = 0, c = 0;
#pragma unroll
    for (int i = 0; i < NC; ++i) { if (z[i] > 0) { kadd(exp(log(z[i]) + sign * lk[i] - largest), s, c); } }
    return largest + log(s);
}

// solve_rachford_rice (rachford_rice.hpp:70-147): beta bisection on (0,1) in the
// overflow-safe log-K form, then explicit roundoff normalisation and balance check.
template <int NC>
MPMC_HOT int rachford_rice(const double* z, const double* lk, RRState<NC>& r, Counters& cnt,
                                          bool& rejected, double guess = 0.5) {
    (void)check_composition<NC>(z, rejected);
    double contrast = 0;
    // the log-K split used by rr_value: e_i = exp(-|lk_i|), numerator via expm1
    double e[NC], num[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        if (!finite(lk[i])) { rejected = true; }
        if (z[i] > 0) { contrast = smax(contrast, __builtin_fabs(lk[i])); }
        e[i] = exp(-__builtin_fabs(lk[i]));
        num[i] = lk[i] >= 0 ? -expm1(-lk[i]) : expm1(lk[i]);
    }
    if (rejected) { return RR_UNREPRESENTABLE; }
    r.mass_abs = 0;
    r.mass_rel = 0;
    if (contrast == 0) { return RR_DEGENERATE; }
    {
        // The reference's interior-root test: ls = log(sum z_i K_i) and rs = log(sum z_i / K_i)
        // in the overflow-safe log-sum-exp form, with no interior root when either is within
        // 64 eps of zero. With every |ln K_i| below 700 both are finite, and a plain sum
        // S+ = sum z_i K_i (S- likewise) carries a relative error of a few eps: S > 1 + 1e-9
        // then certifies ls (rs) > 64 eps, and the 2 nc + 2 logarithms and 2 nc exponentials
        // of the exact test are evaluated only when a sum is within that band.
        bool quick = true;
        double sp = 0, sm = 0;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            if (z[i] == 0) { continue; }
            quick = quick && __builtin_fabs(lk[i]) < 700.0;
            const double inv = 1.0 / e[i];
            sp += z[i] * (lk[i] >= 0 ? inv : e[i]);
            sm += z[i] * (lk[i] >= 0 ? e[i] : inv);
        }
        if (!(quick && sp > 1.0 + 1e-9 && sm > 1.0 + 1e-9)) {
            const double ls = rr_log_sum<NC>(z, lk, 1.0), rs = rr_log_sum<NC>(z, lk, -1.0);
            if (!finite(ls) || !finite(rs)) { return RR_UNREPRESENTABLE; }
            if (ls <= 64 * kEps || rs <= 64 * kEps) { return RR_NO_INTERIOR; }
        }
    }
    // Certified signs. With e_i, num_i fixed, the exact F(b) = sum z_i num_i / den_i(b) is
    // strictly decreasing on (0,1) with |F'| >= S2 = sum z_i num_i^2 (den_i <= 1), and the
    // reference's computed value differs from F(b) by at most E(b) = 32 eps A / min(b,1-b)
    // (A = sum z_i |num_i|; each term carries <= 6 roundings, the compensated sum <= 2).
    // A safeguarded Newton estimate b0 is enclosed by [blo, bhi], checked with two exact
    // evaluations; a midpoint b outside the enclosure with S2 dist(b, [blo,bhi]) > E(b)
    // has a computed value of known, nonzero sign and is not evaluated. The visited
    // midpoints, the decisions and the final beta are those of the plain bisection.
    RRBracket cert = rr_enclose<NC>(z, lk, e, num, guess);
    double left = 0, right = 1, beta = 0.5, value = 0;
    bool resolved = false;
    int it0 = 0;
    if (cert.ok) {
        // Every bisection interval is a dyadic interval [j 2^-k, (j+1) 2^-k] (midpoints of
        // dyadic endpoints in [0, 1] are exact), and a midpoint farther than m from the
        // enclosure is skipped by the rule below: every skipped midpoint lies in
        // (lo/2, (1 + hi)/2), where the rule's bound 32 eps A / min(b, 1 - b) is below
        // s2 m / 2. So the first k levels, while [lo - m, hi + m] stays inside one dyadic
        // interval, are taken at once: k = the common leading bits of the two ends (at most
        // 49, where intervals are still wider than the termination width).
        const double m = 128.0 * kEps * cert.a / (cert.s2 * smin(cert.lo, 1 - cert.hi));
        const double lo = cert.lo - m, hi = cert.hi + m;
        if (finite(m) && lo > 0 && hi < 1) {
            const uint64_t x = static_cast<uint64_t>(lo * 4503599627370496.0); // floor(lo 2^52)
            const uint64_t y = static_cast<uint64_t>(hi * 4503599627370496.0);
            int k = x == y ? 52 : __builtin_clzll(x ^ y) - 12;
            k = k > 49 ? 49 : k;
            if (k > 0) {
                const uint64_t keep = ~((uint64_t{1} << (52 - k)) - 1);
                left = __builtin_ldexp(static_cast<double>(x & keep), -52);
                right = left + __builtin_ldexp(1.0, -k);
                it0 = k;
                cnt.rr += static_cast<uint32_t>(k);
            }
        }
    }
    for (int it = it0; it < 192; ++it) {
        ++cnt.rr;
        beta = (left + right) / 2;
        if (!(beta > 0 && beta < 1)) { break; }
        const bool term = (beta == left) | (beta == right) | (right - left <= 8 * kEps * smin(beta, 1 - beta));
        if (cert.ok && !term) {
            const double dm = smin(beta, 1 - beta);
            if (dm > 1e-200) {
                const double eb = 32.0 * kEps * cert.a / dm;
                if (beta < cert.lo && cert.s2 * (cert.lo - beta) > eb) { left = beta; continue; }
                if (beta > cert.hi && cert.s2 * (beta - cert.hi) > eb) { right = beta; continue; }
            }
        }
        value = rr_value_exact<NC>(beta, z, lk, e, num);
        if (!finite(value)) { return RR_UNREPRESENTABLE; }
        if (value == 0 || term) {
            resolved = true;
            break;
        }
        if (value > 0) { left = beta; } else { right = beta; }
    }
    if (!resolved) { return RR_ITERATION_LIMIT; }
    r.beta = beta;
    if (__builtin_fabs(value) > 64 * kEps) { return RR_UNREPRESENTABLE; }
    bool ok = true;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        r.x[i] = 0; r.y[i] = 0;
        if (z[i] == 0) { continue; }
        const double den = lk[i] >= 0 ? beta + (1 - beta) * e[i] : (1 - beta) + beta * e[i];
        const double larger = z[i] / den;
        r.x[i] = lk[i] >= 0 ? larger * e[i] : larger;
        r.y[i] = lk[i] >= 0 ? larger : larger * e[i];
        ok = ok && finite(larger) && r.x[i] > 0 && r.y[i] > 0;
    }
    if (!ok) { return RR_UNREPRESENTABLE; }
    const double sl = ksum<NC>(r.x), sv = ksum<NC>(r.y);
    if (__builtin_fabs(sl - 1) > 64 * kEps || __builtin_fabs(sv - 1) > 64 * kEps) { return RR_UNREPRESENTABLE; }
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        r.x[i] /= sl;
        r.y[i] /= sv;
        const double rec = __builtin_fma(beta, r.y[i], (1 - beta) * r.x[i]);
        const double err = __builtin_fabs(rec - z[i]);
        r.mass_abs = smax(r.mass_abs, err);
        if (z[i] > 0) {
            ok = ok && r.x[i] > 0 && r.y[i] > 0;
            r.mass_rel = smax(r.mass_rel, err / z[i]);
        } else if (rec != 0) { ok = false; }
    }
    if (!ok) { return RR_UNREPRESENTABLE; }
    (void)check_composition<NC>(r.x, rejected);
    (void)check_composition<NC>(r.y, rejected);
    return rejected ? RR_UNREPRESENTABLE : RR_INTERIOR;
}

// ------------------------------------------------------------------ two-phase split
enum AttemptStatus : int { A_CONVERGED = 0, A_NO_INTERIOR, A_RR_FAILURE, A_INDISTINGUISHABLE, A_DISAPPEARANCE,
                           A_PROPERTY_FAILURE, A_ITERATION_LIMIT, A_EVALUATION_LIMIT, A_LINE_SEARCH_FAILED,
                           A_UNREPRESENTABLE_SEED, A_BALANCE_FAILURE };

template <int NC>
struct SplitState {
    RRState<NC> fr;
    Activity<NC> liq, vap;
    double zl, zv;
    double lk[NC], r[NC], common[NC];
    double norm, gibbs, guard;
};

// evaluate() of iterate_pt_split (pt_split.hpp:300-341): RR, liquid and vapor
// candidates, fugacity residual, reduced Gibbs energy. Returns -1 on success, an
// AttemptStatus for RR/budget failures, or A_PROPERTY_FAILURE.
template <int NC, bool kN = false>
MPMC_HOT int split_evaluate(const Mix<NC>& m, double p, double t, const double* z, const double* lk,
                                           uint32_t& evaluations, uint32_t max_eval, SplitState<NC>& s,
                                           Counters& cnt, bool& rejected, double beta_guess = 0.5) {
    const int rs = rachford_rice<NC>(z, lk, s.fr, cnt, rejected, beta_guess);
    if (rejected) { return A_PROPERTY_FAILURE; }
    if (rs != RR_INTERIOR) {
        return rs == RR_DEGENERATE ? A_INDISTINGUISHABLE : (rs == RR_NO_INTERIOR ? A_NO_INTERIOR : A_RR_FAILURE);
    }
    for (int role = 0; role < 2; ++role) {
        if (evaluations >= max_eval) { return A_EVALUATION_LIMIT; }
        ++evaluations; ++cnt.prop;
        const int e = role == 0 ? side_provider<NC, kN>(m, p, t, s.fr.x, 0, s.liq, s.zl)
                                : side_provider<NC, kN>(m, p, t, s.fr.y, 1, s.vap, s.zv);
        if (e == REJECTED) { rejected = true; return A_PROPERTY_FAILURE; }
        if (e != ISSUE_NONE) { return A_PROPERTY_FAILURE; }
    }
    if (!finite(s.zl) || s.zl <= 0 || !finite(s.zv) || s.zv <= 0) { return A_PROPERTY_FAILURE; }
    double corr = 0, mag = 1, mc = 0;
    const double beta = s.fr.beta;
    s.norm = 0;
    s.gibbs = 0;
    bool fin = true;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        s.lk[i] = 0; s.r[i] = 0; s.common[i] = 0;
        if (z[i] == 0) { continue; }
        const double lx = log(s.fr.x[i]), ly = log(s.fr.y[i]);
        const double mul = lx + s.liq.lnphi[i], muv = ly + s.vap.lnphi[i];
        s.lk[i] = ly - lx;
        s.r[i] = (lx - ly) + (s.liq.lnphi[i] - s.vap.lnphi[i]);
        s.common[i] = (mul + muv) / 2;
        s.norm = smax(s.norm, __builtin_fabs(s.r[i]));
        kadd((1 - beta) * s.fr.x[i] * mul + beta * s.fr.y[i] * muv, s.gibbs, corr);
        kadd((1 - beta) * s.fr.x[i] * (__builtin_fabs(lx) + __builtin_fabs(s.liq.lnphi[i])) +
                 beta * s.fr.y[i] * (__builtin_fabs(ly) + __builtin_fabs(s.vap.lnphi[i])), mag, mc);
        fin = fin && finite(s.r[i]) && finite(s.common[i]);
    }
    if (!fin) { return A_PROPERTY_FAILURE; }
    s.guard = 256 * kEps * mag;
    if (!finite(s.gibbs) || !finite(s.guard)) { return A_PROPERTY_FAILURE; }
    return -1;
}

// iterate_pt_split (pt_split.hpp:274-408): logK <- logK + (alpha/scale) r, accepted by the
// Gibbs Armijo rule with the closed-form slope, else by residual progress.
template <int NC, bool kN = false>
__device__ __noinline__ int iterate_split(const Mix<NC>& m, double p, double t, const double* z, const double* lk0,
                                          double fug_tol, uint32_t max_eval, uint32_t& evaluations,
                                          SplitState<NC>& pt, Counters& cnt, bool& rejected) {
    constexpr double kMassAbs = 1e-12, kMassRel = 1e-10, kMinPf = 1e-10, kLkSep = 1e-7, kZSep = 1e-8,
                     kMaxLogStep = 2, kResDec = 1e-4, kArmijo = 1e-4;
    constexpr int kMaxIt = 512, kMaxBt = 24;
    (void)check_composition<NC>(z, rejected);
    if (rejected) { return A_PROPERTY_FAILURE; }
    {
        const int e = split_evaluate<NC, kN>(m, p, t, z, lk0, evaluations, max_eval, pt, cnt, rejected);
        if (e != -1) { return e; }
    }
    int iterations = 0;
    for (;;) {
        if (pt.norm <= fug_tol) {
            const double beta = pt.fr.beta;
            double contrast = 0;
#pragma unroll
            for (int i = 0; i < NC; ++i) { contrast = smax(contrast, __builtin_fabs(pt.lk[i])); }
            if (!(pt.fr.mass_abs <= kMassAbs && pt.fr.mass_rel <= kMassRel)) { return A_BALANCE_FAILURE; }
            if (beta <= kMinPf || 1 - beta <= kMinPf) { return A_DISAPPEARANCE; }
            if (contrast <= kLkSep || pt.zv - pt.zl <= kZSep * smax(pt.zv, pt.zl)) { return A_INDISTINGUISHABLE; }
            return A_CONVERGED;
        }
        if (iterations >= kMaxIt) { return A_ITERATION_LIMIT; }
        const double scale = smax(1.0, pt.norm / kMaxLogStep);
        double slope;
        {
            // split_gibbs_slope (pt_split.hpp:213-231)
            double h = 0, hc = 0, f = 0, fc = 0, s2 = 0, s2c = 0;
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                if (z[i] == 0) { continue; }
                const double x = pt.fr.x[i], y = pt.fr.y[i];
                const double c = (x / z[i]) * y;
                const double d = y - x;
                const double r = pt.r[i];
                kadd(d * (d / z[i]), h, hc);
                kadd(c * r, f, fc);
                kadd((c * r) * r, s2, s2c);
            }
            if (!(h > 0) || !finite(h) || !finite(f) || !finite(s2) || !finite(scale) || !(scale >= 1)) {
                slope = __builtin_nan("");
            } else {
                const double beta = pt.fr.beta;
                slope = -(beta * (1 - beta) * s2 + (f / h) * f) / scale;
            }
        }
        if (!finite(slope) || !(slope < 0)) { return A_LINE_SEARCH_FAILED; }
        bool accepted = false;
        double alpha = 1;
        for (int bt = 0; bt < kMaxBt; ++bt, alpha *= 0.5) {
            double nlk[NC];
#pragma unroll
            for (int i = 0; i < NC; ++i) { nlk[i] = pt.lk[i] + (alpha / scale) * pt.r[i]; }
            SplitState<NC> nx;
            const int e = split_evaluate<NC, kN>(m, p, t, z, nlk, evaluations, max_eval, nx, cnt, rejected, pt.fr.beta);
            if (rejected) { return A_PROPERTY_FAILURE; }
            if (e == -1) {
                const double pred = -alpha * slope;
                const double g = pt.guard + nx.guard;
                bool ok;
                if (pred > g) { ok = nx.gibbs <= pt.gibbs - kArmijo * pred; }
                else {
                    ok = nx.gibbs <= pt.gibbs + g &&
                         (nx.norm <= fug_tol || pt.norm - nx.norm > kResDec * (alpha / scale) * pt.norm);
                }
                if (ok) { pt = nx; accepted = true; break; }
            } else if (e == A_EVALUATION_LIMIT) {
                return A_EVALUATION_LIMIT;
            }
        }
        if (!accepted) { return A_LINE_SEARCH_FAILED; }
        ++iterations;
    }
}

// split_seed (pt_split.hpp:245-267)
template <int NC>
__device__ __forceinline__ bool split_seed(const double* z, const double* w, bool as_vapor, double* lk) {
    double amount = 0.1;
#pragma unroll
    for (int i = 0; i < NC; ++i) { if (z[i] > 0) { amount = smin(amount, 0.5 * (z[i] / w[i])); } }
    if (!(amount > 0 && amount < 1)) { return false; }
    double other[NC];
    bool ok = true;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        other[i] = 0;
        lk[i] = 0;
        if (z[i] == 0) { continue; }
        other[i] = __builtin_fma(-amount, w[i], z[i]) / (1 - amount);
        ok = ok && other[i] > 0;
    }
    if (!ok) { return false; }
    const double s = ksum<NC>(other);
    if (!finite(s) || __builtin_fabs(s - 1) > 64 * kEps) { return false; }
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        if (z[i] == 0) { continue; }
        other[i] /= s;
        lk[i] = (as_vapor ? 1.0 : -1.0) * (log(w[i]) - log(other[i]));
        ok = ok && finite(lk[i]);
    }
    return ok;
}

enum VleStatus : int { V_SINGLE = 0, V_TWO, V_UNSTABLE, V_INDETERMINATE };

template <int NC>
struct VleOut {
    int status;
    StabOut<NC> init;      // initial feed stability (reference phase, feed, witnesses)
    bool has_candidate;
    SplitState<NC> cand;   // selected converged pair
    bool has_final;
    StabOut<NC> fin;       // common-tangent review (witnesses seed max3)
    bool any_disappearance; // an attempt ended at a phase-fraction endpoint
};

// solve_pt_vle (pt_split.hpp:414-525) with the PR76 evaluator as both providers.
template <int NC, bool kN = false>
// starts: extra starts of the initial stability test; final_starts (nfinal >= 0): extra
// starts of the common-tangent review (pt_split.hpp:414-512 initial_starts / final_starts),
// nfinal < 0: the initial starts again.
__device__ __noinline__ void solve_vle(const Mix<NC>& m, double p, double t, const double* feed, double fug_tol,
                                       const double* const* starts, int nstarts, VleOut<NC>& V, Counters& cnt,
                                       bool& rejected, const double* const* final_starts = nullptr,
                                       int nfinal = -1) {
    constexpr uint32_t kMaxSplitEval = 20000;
    constexpr int kMaxAttempts = 16;
    V.status = V_INDETERMINATE;
    V.has_candidate = false;
    V.has_final = false;
    V.any_disappearance = false;
    if (!finite(p) || p <= 0 || !finite(t) || t <= 0) { rejected = true; return; }
    (void)check_composition<NC>(feed, rejected);
    for (int e = 0; e < nstarts; ++e) {
        (void)check_composition<NC>(starts[e], rejected);
#pragma unroll
        for (int i = 0; i < NC; ++i) { if ((feed[i] > 0) != (starts[e][i] > 0)) { rejected = true; } }
    }
    if (rejected) { return; }
    stability<NC, kN>(m, p, t, feed, nullptr, starts, nstarts, 1e-10, V.init, cnt, rejected);
    if (rejected) { return; }
    if (V.init.status == S_NO_INSTABILITY) { V.status = V_SINGLE; return; }
    if (V.init.status != S_UNSTABLE) { return; }
    const double* z = V.init.feed;
    int attempts = 0;
    uint32_t split_evals = 0;
    bool limit = false;
    for (int k = 0; k < V.init.nwit && !limit; ++k) {
        for (int side = 0; side < 2; ++side) {
            if (attempts >= kMaxAttempts || split_evals >= kMaxSplitEval) { limit = true; break; }
            double lk[NC];
            int status = A_UNREPRESENTABLE_SEED;
            uint32_t evals = 0;
            SplitState<NC> st;
            if (split_seed<NC>(z, V.init.wit[k].w, side == 0, lk)) {
                status = iterate_split<NC, kN>(m, p, t, z, lk, fug_tol, kMaxSplitEval - split_evals, evals, st, cnt,
                                           rejected);
                if (rejected) { return; }
            }
            split_evals += evals;
            ++attempts;
            if (status == A_DISAPPEARANCE) { V.any_disappearance = true; }
            if (status == A_CONVERGED && (!V.has_candidate || st.gibbs < V.cand.gibbs)) {
                V.cand = st;
                V.has_candidate = true;
            }
        }
    }
    if (!V.has_candidate) { return; }
    const SplitState<NC>& c = V.cand;
    double feed_g = 0, corr = 0, mag = 1;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        if (z[i] == 0) { continue; }
        const double lz = log(z[i]);
        const double ph = V.init.ref.lnphi[i];
        kadd(z[i] * (lz + ph), feed_g, corr);
        mag += z[i] * (__builtin_fabs(lz) + __builtin_fabs(ph));
    }
    const double dg = c.gibbs - feed_g;
    const double guard = c.guard + 256 * kEps * mag;
    const double ftol = 1e-10 + 0.5 * c.norm;
    if (!finite(dg) || !finite(guard) || !finite(ftol)) { return; }
    const double* fs[6];
    int nf = 0;
    if (nfinal < 0) {
        for (int e = 0; e < nstarts; ++e) { fs[nf++] = starts[e]; }
    } else {
        for (int e = 0; e < nfinal; ++e) { fs[nf++] = final_starts[e]; }
    }
    fs[nf++] = c.fr.x;
    fs[nf++] = c.fr.y;
    stability<NC, kN>(m, p, t, z, c.common, fs, nf, ftol, V.fin, cnt, rejected);
    if (rejected) { return; }
    V.has_final = true;
    if (V.fin.status == S_UNSTABLE) { V.status = V_UNSTABLE; }
    else if (dg > guard) { V.status = V_INDETERMINATE; }
    else if (V.fin.status == S_NO_INSTABILITY) { V.status = V_TWO; }
    else { V.status = V_INDETERMINATE; }
}

// ------------------------------------------------------------------ three phases
enum P3Status : int { P_CONVERGED = 0, P_NO_BALANCE, P_DEGENERATE, P_BAL_ITER, P_BAL_LS, P_BAL_FAILURE,
                      P_DISAPPEARANCE, P_INDISTINGUISHABLE, P_PROPERTY, P_ITERATION_LIMIT, P_EVALUATION_LIMIT,
                      P_LINE_SEARCH_FAILED };

template <int NC>
struct P3State {
    double frac[3];
    double comp[3][NC];
    Activity<NC> act[3];
    double z[3];
    double lk1[NC], lk2[NC], r1[NC], r2[NC], common[NC];
    double norm, gibbs, guard, rr_res, mass_abs, mass_rel;
};

__device__ __forceinline__ bool rr3_feasible(double b0, double b1) {
    return finite(b0) && finite(b1) && b0 >= 0 && b1 >= 0 && b0 + b1 <= 1;
}

// rr3_evaluate (pt_three_phase.hpp:187-236). comp may be nullptr.
template <int NC>
__device__ __forceinline__ bool rr3_eval(const double* z, const double* l1, const double* l2, double b1, double b2,
                                         double* f, double* j, double (*comp)[NC]) {
    if (!rr3_feasible(b1, b2)) { return false; }
    const double b0 = 1.0 - b1 - b2;
    double f0 = 0, f0c = 0, f1 = 0, f1c = 0, j00 = 0, j00c = 0, j01 = 0, j01c = 0, j11 = 0, j11c = 0;
    bool ok = true;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        if (comp) { comp[0][i] = comp[1][i] = comp[2][i] = 0; }
        if (z[i] == 0) { continue; }
        const double sh = smax(smax(0.0, l1[i]), l2[i]);
        const double e0 = exp(-sh), e1 = exp(l1[i] - sh), e2 = exp(l2[i] - sh);
        const double den = b0 * e0 + b1 * e1 + b2 * e2;
        if (!finite(den) || !(den > 0)) { ok = false; continue; }
        const double x0 = z[i] * e0 / den, x1 = z[i] * e1 / den, x2 = z[i] * e2 / den;
        if (!finite(x0) || !finite(x1) || !finite(x2) || !(x0 > 0) || !(x1 > 0) || !(x2 > 0)) { ok = false; continue; }
        if (comp) { comp[0][i] = x0; comp[1][i] = x1; comp[2][i] = x2; }
        const double d1 = x1 - x0, d2 = x2 - x0;
        kadd(d1, f0, f0c);
        kadd(d2, f1, f1c);
        kadd(-(d1 * d1) / z[i], j00, j00c);
        kadd(-(d1 * d2) / z[i], j01, j01c);
        kadd(-(d2 * d2) / z[i], j11, j11c);
    }
    f[0] = f0; f[1] = f1;
    j[0] = j00; j[1] = j01; j[2] = j11;
    return ok;
}

enum RR3Status : int { G_SUCCESS = 0, G_DEGENERATE, G_ITERATION_LIMIT, G_LINE_SEARCH, G_UNREPRESENTABLE };

// solve_generalized_rr3 (pt_three_phase.hpp:238-355): 2x2 Newton with feasibility
// backtracking, then normalisation and material-balance errors.
template <int NC>
__device__ __noinline__ int rr3_solve(const double* z, const double* l1, const double* l2, double bi1, double bi2,
                                      double* beta_out, double (*comp)[NC], double& residual, double& mass_abs,
                                      double& mass_rel, Counters& cnt) {
    constexpr double kBalTol = 2e-13;
    constexpr int kMaxIt = 192, kMaxBt = 48;
    residual = kInf;
    mass_abs = 0;
    mass_rel = 0;
    if (!rr3_feasible(bi1, bi2)) { return G_UNREPRESENTABLE; }
    bool fin = true;
#pragma unroll
    for (int i = 0; i < NC; ++i) { fin = fin && finite(l1[i]) && finite(l2[i]); }
    if (!fin) { return G_UNREPRESENTABLE; }
    double b1 = bi1, b2 = bi2;
    double f[2], j[3];
    for (int it = 0;; ++it) {
        if (!rr3_eval<NC>(z, l1, l2, b1, b2, f, j, nullptr)) { return G_UNREPRESENTABLE; }
        const double norm = smax(__builtin_fabs(f[0]), __builtin_fabs(f[1]));
        if (!finite(norm)) { return G_UNREPRESENTABLE; }
        if (norm <= kBalTol) { residual = norm; break; }
        if (it >= kMaxIt) { return G_ITERATION_LIMIT; }
        ++cnt.rr3;
        const double det = j[0] * j[2] - j[1] * j[1];
        const double ds = __builtin_fabs(j[0] * j[2]) + __builtin_fabs(j[1] * j[1]) + 1.0;
        if (!finite(det) || __builtin_fabs(det) <= 256.0 * kEps * ds) { return G_DEGENERATE; }
        const double s0 = (-f[0] * j[2] + j[1] * f[1]) / det, s1 = (j[1] * f[0] - j[0] * f[1]) / det;
        if (!finite(s0) || !finite(s1)) { return G_UNREPRESENTABLE; }
        bool acc = false;
        double alpha = 1.0;
        for (int bt = 0; bt < kMaxBt; ++bt, alpha *= 0.5) {
            const double c1 = b1 + alpha * s0, c2 = b2 + alpha * s1;
            if (!rr3_feasible(c1, c2)) { continue; }
            double nf[2], nj[3];
            if (!rr3_eval<NC>(z, l1, l2, c1, c2, nf, nj, nullptr)) { continue; }
            const double nn = smax(__builtin_fabs(nf[0]), __builtin_fabs(nf[1]));
            if (!finite(nn)) { continue; }
            if (nn < norm || nn <= kBalTol) { b1 = c1; b2 = c2; acc = true; break; }
        }
        if (!acc) { return G_LINE_SEARCH; }
    }
    if (!rr3_eval<NC>(z, l1, l2, b1, b2, f, j, comp)) { return G_UNREPRESENTABLE; }
    beta_out[0] = b1;
    beta_out[1] = b2;
#pragma unroll
    for (int ph = 0; ph < 3; ++ph) {
        const double raw = ksum<NC>(comp[ph]);
        if (!finite(raw) || !(raw > 0)) { return G_UNREPRESENTABLE; }
#pragma unroll
        for (int i = 0; i < NC; ++i) { comp[ph][i] /= raw; }
    }
    const double fr0 = 1.0 - b1 - b2;
    double ae = 0, re = 0;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        const double rec = fr0 * comp[0][i] + b1 * comp[1][i] + b2 * comp[2][i];
        const double err = __builtin_fabs(rec - z[i]);
        ae = smax(ae, err);
        if (z[i] > 0) { re = smax(re, err / z[i]); } else if (rec != 0) { re = kInf; }
    }
    mass_abs = ae;
    mass_rel = re;
    return G_SUCCESS;
}

template <int NC>
__device__ __forceinline__ double log_distance(const double* a, const double* b, const double* z) {
    double d = 0;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        if (z[i] == 0) { if (a[i] != 0 || b[i] != 0) { return kInf; } continue; }
        if (!(a[i] > 0) || !(b[i] > 0)) { return kInf; }
        const double v = __builtin_fabs(log(a[i]) - log(b[i]));
        if (!finite(v)) { return kInf; }
        d = smax(d, v);
    }
    return d;
}

// evaluate_state of iterate_pt_three_phase (pt_three_phase.hpp:463-555)
template <int NC, bool kN = false>
__device__ __noinline__ int p3_evaluate(const Mix<NC>& m, double p, double t, const double* feed, const double* l1,
                                        const double* l2, double s1, double s2, const int* sides,
                                        uint32_t& evaluations, P3State<NC>& s, Counters& cnt, bool& rejected) {
    constexpr uint32_t kMaxEval = 30000;
    double beta[2];
    const int bs = rr3_solve<NC>(feed, l1, l2, s1, s2, beta, s.comp, s.rr_res, s.mass_abs, s.mass_rel, cnt);
    switch (bs) {
    case G_SUCCESS: break;
    case G_DEGENERATE: return P_DEGENERATE;
    case G_ITERATION_LIMIT: return P_BAL_ITER;
    case G_LINE_SEARCH: return P_BAL_LS;
    default: return P_NO_BALANCE;
    }
    s.frac[0] = 1.0 - beta[0] - beta[1];
    s.frac[1] = beta[0];
    s.frac[2] = beta[1];
    for (int ph = 0; ph < 3; ++ph) {
        if (evaluations >= kMaxEval) { return P_EVALUATION_LIMIT; }
        ++evaluations; ++cnt.prop;
        const int e = side_provider<NC, kN>(m, p, t, s.comp[ph], sides[ph], s.act[ph], s.z[ph]);
        if (e == REJECTED) { rejected = true; return P_PROPERTY; }
        if (e != ISSUE_NONE) { return P_PROPERTY; }
        if (!finite(s.z[ph]) || !(s.z[ph] > 0)) { return P_PROPERTY; }
    }
    double gc = 0, mag = 1, mc = 0;
    s.norm = 0;
    s.gibbs = 0;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        s.lk1[i] = s.lk2[i] = s.r1[i] = s.r2[i] = s.common[i] = 0;
        if (feed[i] == 0) { continue; }
        const double lx[3] = {log(s.comp[0][i]), log(s.comp[1][i]), log(s.comp[2][i])};
        const double mu0 = lx[0] + s.act[0].lnphi[i], mu1 = lx[1] + s.act[1].lnphi[i], mu2 = lx[2] + s.act[2].lnphi[i];
        s.lk1[i] = lx[1] - lx[0];
        s.lk2[i] = lx[2] - lx[0];
        s.r1[i] = mu0 - mu1;
        s.r2[i] = mu0 - mu2;
        s.common[i] = (mu0 + mu1 + mu2) / 3.0;
        s.norm = smax(smax(s.norm, __builtin_fabs(s.r1[i])), __builtin_fabs(s.r2[i]));
#pragma unroll
        for (int ph = 0; ph < 3; ++ph) {
            const double mu = lx[ph] + s.act[ph].lnphi[i];
            kadd(s.frac[ph] * s.comp[ph][i] * mu, s.gibbs, gc);
            kadd(s.frac[ph] * s.comp[ph][i] * (__builtin_fabs(lx[ph]) + __builtin_fabs(s.act[ph].lnphi[i])), mag, mc);
        }
    }
    s.guard = 256.0 * kEps * mag;
    if (!finite(s.norm) || !finite(s.gibbs) || !finite(s.guard)) { return P_PROPERTY; }
    return -1;
}

// iterate_pt_three_phase (pt_three_phase.hpp:402-674)
template <int NC, bool kN = false>
__device__ __noinline__ int iterate_three(const Mix<NC>& m, double p, double t, const double* feed_in,
                                          const double* l1_0, const double* l2_0, double b1, double b2,
                                          const int* sides, P3State<NC>& pt, int& disappearing, Counters& cnt,
                                          bool& rejected) {
    constexpr double kMuTol = 1e-11, kMassAbs = 1e-12, kMassRel = 1e-10, kBalTol = 2e-13, kMinPf = 1e-10,
                     kLsep = 1e-7, kMaxLogStep = 2.0, kResDec = 1e-4;
    constexpr int kMaxIt = 512, kMaxBt = 32;
    disappearing = -1;
    if (!rr3_feasible(b1, b2)) { rejected = true; return P_NO_BALANCE; }
    const double isum = check_composition<NC>(feed_in, rejected);
    if (rejected) { return P_NO_BALANCE; }
    double feed[NC];
    bool lost = false;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        feed[i] = feed_in[i] / isum;
        lost = lost || (feed_in[i] > 0 && feed[i] == 0);
    }
    if (lost) { return P_BAL_FAILURE; }
    uint32_t evaluations = 0;
    {
        const int e = p3_evaluate<NC, kN>(m, p, t, feed, l1_0, l2_0, b1, b2, sides, evaluations, pt, cnt, rejected);
        if (e != -1) { return e; }
    }
    int iterations = 0;
    for (;;) {
        if (pt.norm <= kMuTol) {
            if (pt.rr_res > kBalTol || pt.mass_abs > kMassAbs || pt.mass_rel > kMassRel) { return P_BAL_FAILURE; }
            for (int ph = 0; ph < 3; ++ph) {
                if (pt.frac[ph] <= kMinPf) { disappearing = ph; return P_DISAPPEARANCE; }
            }
            const double d01 = log_distance<NC>(pt.comp[0], pt.comp[1], feed);
            const double d02 = log_distance<NC>(pt.comp[0], pt.comp[2], feed);
            const double d12 = log_distance<NC>(pt.comp[1], pt.comp[2], feed);
            if (!(d01 > kLsep) || !(d02 > kLsep) || !(d12 > kLsep)) { return P_INDISTINGUISHABLE; }
            return P_CONVERGED;
        }
        if (iterations >= kMaxIt) { return P_ITERATION_LIMIT; }
        const double scale = smax(1.0, pt.norm / kMaxLogStep);
        bool acc = false;
        double alpha = 1.0;
        for (int bt = 0; bt < kMaxBt; ++bt, alpha *= 0.5) {
            double n1[NC], n2[NC];
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                n1[i] = pt.lk1[i] + (alpha / scale) * pt.r1[i];
                n2[i] = pt.lk2[i] + (alpha / scale) * pt.r2[i];
            }
            P3State<NC> nx;
            const int e = p3_evaluate<NC, kN>(m, p, t, feed, n1, n2, pt.frac[1], pt.frac[2], sides, evaluations, nx, cnt,
                                          rejected);
            if (rejected) { return P_PROPERTY; }
            if (e == -1) {
                const double req = kResDec * (alpha / scale) * pt.norm;
                const bool prog = nx.norm <= kMuTol || pt.norm - nx.norm > req;
                const bool gok = nx.gibbs <= pt.gibbs + pt.guard + nx.guard;
                if (prog && gok) { pt = nx; acc = true; break; }
            } else if (e == P_EVALUATION_LIMIT) {
                return P_EVALUATION_LIMIT;
            }
        }
        if (!acc) { return P_LINE_SEARCH_FAILED; }
        ++iterations;
    }
}

// pr76_witness_root_sides (pr76_three_phase.hpp:273-305); bit 0 = lower, bit 1 = upper.
template <int NC, bool kN = false>
__device__ __noinline__ int witness_sides(const Mix<NC>& m, double p, double t, const double* w, int branch) {
    PhaseCtx<NC> ctx;
    if (phase_prepare<NC, kN>(m, p, t, w, ctx) != ISSUE_NONE) { return 0; }
    // first and last admissible root, in root order
    int na = 0, first = -1, last = -1;
    bool found = false;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k < ctx.rs.count && root_at(ctx.rs, k).slope_sign > 0) {
            if (na == 0) { first = k; }
            last = k;
            ++na;
            found = found || k == branch;
        }
    }
    if (na == 0 || !found) { return 0; }
    if (na == 1) { return 3; }
    if (branch == first) { return 1; }
    if (branch == last) { return 2; }
    return 0;
}

template <int NC>
__device__ __forceinline__ bool log_ratio(const double* num, const double* den, const double* z, double* out) {
    bool ok = true;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        out[i] = 0;
        if (z[i] == 0) { ok = ok && num[i] == 0 && den[i] == 0; continue; }
        ok = ok && num[i] > 0 && den[i] > 0;
        out[i] = log(num[i]) - log(den[i]);
        ok = ok && finite(out[i]);
    }
    return ok;
}

enum Max3Status : int { M_SINGLE = 0, M_TWO, M_THREE, M_BOUNDARY, M_HIGHER, M_INDETERMINATE };

// Output of the max3 orchestration after an unstable two-phase review: the accepted
// three-phase state, or the fresh two-phase neighbour of a disappearance boundary.
template <int NC>
struct Max3Tail {
    int status;
    int attempts;          // three-phase attempts run (Pr76PtMax3Result::attempts.size())
    bool start_consumed;   // an attempt ran from the caller-supplied start
    bool use_neighbor;
    SplitState<NC> neighbor;
    P3State<NC> three;
};

// The part of solve_pr76_pt_max3 (pr76_three_phase.hpp:360-561) that runs once the
// two-phase common-tangent review found lower-Gibbs evidence. Caller-supplied three-phase
// starts (Pr76PtMax3Options::three_phase_starts) are tried first, in slot order, each over
// the eight root-side combinations (:472-498): slot 0 is the PR76 PT continuation's carried
// state (inserted first, pr76_pt_continuation.hpp:240-251), slot 1 a fixed start; then
// negative final-review witnesses seed attempts (witness sides from
// pr76_witness_root_sides), each followed by the three-phase final review or, on
// disappearance, a fresh two-phase neighbour solve. `witness(i, w, branch)` yields the i-th
// negative witness of the review in trial order. start0 / start1: three compositions [3][NC]
// or nullptr; frac0 / frac1: their phase-fraction seeds of phases 1 and 2. Two slots make at
// most 16 supplied attempts, the reference's continuation quota (max_three_phase_attempts).
template <int NC, bool kN, typename WitnessFn>
__device__ __noinline__ void max3_tail(const Mix<NC>& m, double p, double t, const double* z, const double* px,
                                       const double* py, double pbeta, int nwit, WitnessFn witness, double fug_tol,
                                       Max3Tail<NC>& M, Counters& cnt, bool& rejected, const double* start0 = nullptr,
                                       const double* frac0 = nullptr, const double* start1 = nullptr,
                                       const double* frac1 = nullptr) {
    M.use_neighbor = false;
    M.attempts = 0;
    M.start_consumed = false;
    bool saw_higher = false, saw_boundary = false;
    // one attempt (run_attempt, pr76_three_phase.hpp:377-467); true when it closes the solve
    // (or the cell is rejected)
    auto attempt = [&](const double* l1, const double* l2, double b1, double b2, const int* sd) -> bool {
        ++M.attempts;
        int disappearing;
        P3State<NC> E;
        const int st = iterate_three<NC, kN>(m, p, t, z, l1, l2, b1, b2, sd, E, disappearing, cnt, rejected);
        if (rejected) { return true; }
        if (st == P_DISAPPEARANCE) {
            saw_boundary = true;
            if (disappearing >= 0) {
                const double* stv[2];
                int ns = 0;
                for (int ph = 0; ph < 3; ++ph) { if (ph != disappearing) { stv[ns++] = E.comp[ph]; } }
                VleOut<NC> nb;
                solve_vle<NC, kN>(m, p, t, z, fug_tol, stv, 2, nb, cnt, rejected);
                if (rejected) { return true; }
                if (nb.status == V_TWO) {
                    M.status = M_TWO;
                    M.use_neighbor = true;
                    M.neighbor = nb.cand;
                    return true;
                }
            }
            return false;
        }
        if (st != P_CONVERGED) { return false; }
        if (!(E.norm <= 1e-11 && E.rr_res <= 2e-13 && E.mass_abs <= 1e-12 && E.mass_rel <= 1e-10)) { return false; }
        const double* stv[3] = {E.comp[0], E.comp[1], E.comp[2]};
        StabOut<NC> F;
        stability<NC, kN>(m, p, t, z, E.common, stv, 3, 1e-10 + E.norm, F, cnt, rejected);
        if (rejected) { return true; }
        if (F.status == S_UNSTABLE) { saw_higher = true; }
        else if (F.status == S_NO_INSTABILITY) { M.status = M_THREE; M.three = E; return true; }
        return false;
    };
    // true when an attempt from this start closes the solve; `carried`: slot 0
    auto supplied = [&](const double* st, const double* sf, bool carried) -> bool {
        double l1[NC], l2[NC];
        if (!(log_ratio<NC>(st + NC, st, z, l1) && log_ratio<NC>(st + 2 * NC, st, z, l2))) { return false; }
        // side of phase 0 outermost, of phase 2 innermost; 0 = lower, 1 = upper admissible
        for (int code = 0; code < 8; ++code) {
            const int sd[3] = {code >> 2, (code >> 1) & 1, code & 1};
            M.start_consumed = M.start_consumed || carried;
            if (attempt(l1, l2, sf[0], sf[1], sd)) { return true; }
        }
        return false;
    };
    if (start0 && supplied(start0, frac0, true)) { return; }
    if (start1 && supplied(start1, frac1, false)) { return; }
    int automatic = 0;
    bool auto_limit = false;
    double src1[NC];
    if (log_ratio<NC>(py, px, z, src1)) {
        for (int k = 0; k < nwit && !auto_limit; ++k) {
            double ww[NC];
            int wbranch;
            witness(k, ww, wbranch);
            double l2[NC];
            if (!log_ratio<NC>(ww, px, z, l2)) { continue; }
            const int sides = witness_sides<NC, kN>(m, p, t, ww, wbranch);
            for (int s = 0; s < 2; ++s) {
                if (!(sides & (1 << s))) { continue; }
                if (automatic >= 16) { auto_limit = true; break; }
                ++automatic;
                const int sd[3] = {0, 1, s};
                if (attempt(src1, l2, (1.0 - 0.1) * pbeta, 0.1, sd)) { return; }
            }
        }
    }
    M.status = saw_higher ? M_HIGHER : (saw_boundary ? M_BOUNDARY : M_INDETERMINATE);
}

// A caller-supplied three-phase start is valid when its fraction seed lies in the simplex
// and each composition is a composition with the feed's support; otherwise the reference
// rejects the request (pr76_pt_max3_check_starts, pr76_three_phase.hpp:211-242).
template <int NC>
__device__ __forceinline__ void check_three_phase_start(const double* feed, const double* start,
                                                        const double* start_frac, bool& rejected) {
    if (!rr3_feasible(start_frac[0], start_frac[1])) { rejected = true; }
    for (int ph = 0; ph < 3; ++ph) {
        (void)check_composition<NC>(start + ph * NC, rejected);
#pragma unroll
        for (int i = 0; i < NC; ++i) { if ((feed[i] > 0) != (start[ph * NC + i] > 0)) { rejected = true; } }
    }
}

template <int NC>
struct Max3Out {
    int status;
    VleOut<NC> base;
    Max3Tail<NC> tail;
};

// solve_pr76_pt_max3 (pr76_three_phase.hpp:328-561), serial in one lane; start0 / frac0
// and start1 / frac1: the supplied three-phase start slots of max3_tail, or nullptr.
template <int NC, bool kN = false>
// initial / final: extra starts of the initial stability test and of the base split's
// common-tangent review (solve_pr76_pt_max3 initial_starts / final_starts,
// pr76_three_phase.hpp:328-341; the warm-start hints of PtSolveHints).
__device__ __noinline__ void solve_max3(const Mix<NC>& m, double p, double t, const double* feed, double fug_tol,
                                        Max3Out<NC>& M, Counters& cnt, bool& rejected,
                                        const double* const* initial = nullptr, int ninitial = 0,
                                        const double* const* final_starts = nullptr, int nfinal = 0,
                                        const double* start0 = nullptr, const double* frac0 = nullptr,
                                        const double* start1 = nullptr, const double* frac1 = nullptr) {
    M.tail.use_neighbor = false;
    M.tail.attempts = 0;
    M.tail.start_consumed = false;
    (void)check_composition<NC>(feed, rejected);
    if (start0) { check_three_phase_start<NC>(feed, start0, frac0, rejected); }
    if (start1) { check_three_phase_start<NC>(feed, start1, frac1, rejected); }
    if (rejected) { return; }
    solve_vle<NC, kN>(m, p, t, feed, fug_tol, initial, ninitial, M.base, cnt, rejected, final_starts, nfinal);
    if (rejected) { return; }
    switch (M.base.status) {
    case V_SINGLE: M.status = M_SINGLE; return;
    case V_TWO: M.status = M_TWO; return;
    case V_INDETERMINATE: M.status = M_INDETERMINATE; return;
    default: break;
    }
    if (!M.base.has_candidate || !M.base.has_final) { M.status = M_INDETERMINATE; return; }
    const StabOut<NC>& fin = M.base.fin;
    max3_tail<NC, kN>(m, p, t, M.base.init.feed, M.base.cand.fr.x, M.base.cand.fr.y, M.base.cand.fr.beta, fin.nwit,
                  [&](int i, double* w, int& branch) {
#pragma unroll
                      for (int c = 0; c < NC; ++c) { w[c] = fin.wit[i].w[c]; }
                      branch = fin.wit[i].branch;
                  },
                  fug_tol, M.tail, cnt, rejected, start0, frac0, start1, frac1);
    M.status = M.tail.status;
}

} // namespace mpmc::dev
