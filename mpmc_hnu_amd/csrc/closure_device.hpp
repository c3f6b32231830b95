// Thermodynamic closure of accepted two-phase flash results on gfx950: the primal
// {Z, molar density} of both phases and the local implicit linearization of
// {beta_V, x_L, y_V, Z_L, Z_V, rho_L, rho_V} over q = (p, T, z_0..z_{nc-2}).
//
// Reference: build_pr76_pt_vle_closure (modules/physics/include/mpmc/physics/
// pr76_thermodynamic_closure.hpp:148-381) on differentiate_pr76_pt_vle
// (modules/flash/include/mpmc/flash/pr76_sensitivity.hpp:51-390), whose local system is
// pr76_sensitivity_local_system (flash/detail/pr76_sensitivity_local.hpp:29-109) with the
// Rachford-Rice manifold linearized by pt_sensitivity_rr_linearization
// (flash/detail/pt_sensitivity_detail.hpp:45-143).
//
// Device formulation. The reference seeds the 2nc+1 variables (logK, p, T, z_red) into
// Dual<double, 4> blocks and re-runs the RR linearization and both phase evaluations per
// block (4 passes at nc = 6), then pr76_phase_z_partials runs two more blocked passes
// per phase. Here each phase is evaluated once in Dual<nc+1> over its own reduced chart
// (p, T, x_0..x_{nc-2}); that single pass yields d ln(phi)/d(chart) and dZ/d(chart).
// The RR linearization is linear in the seeds, so the equilibrium Jacobian
// A = dF/dlogK, B = dF/dq and the total derivatives are assembled from those partials
// by the chain rule. Same mathematics, fewer operations; values agree with the
// reference to rounding.
//
// Two guards of the reference need the split iterate's logK and fugacity norm
// (pr76_sensitivity.hpp:144-186): max|logK| > log_k_separation and norm <= tolerance.
// Both are the convergence conditions every accepted candidate passed at the same
// values (pt_split.hpp:351-361), so they hold for every published two-phase set and
// are not re-tested; the Z-separation guard is.
#pragma once

#include "dual.hpp"
#include "flash_kernels.hpp"

#include <algorithm>

namespace mpmc::dev {

// ThermodynamicClosurePrimalStatus / LinearizationReason (thermodynamic_closure.hpp:18-35)
enum ClosurePrimal : int { CP_VALID = 0, CP_NOT_IMPLEMENTED = 1, CP_INDETERMINATE = 2 };
enum ClosureReason : int { CR_NONE = 0, CR_NOT_IMPLEMENTED, CR_PHASE_BOUNDARY, CR_ILL_CONDITIONED,
                           CR_FEED_SUPPORT, CR_PROPERTY_FAILURE, CR_ARITHMETIC_FAILURE, CR_NOT_ACCEPTED };
constexpr int32_t kClosureThrown = 0x2000;

struct ClosureIn {
    const int32_t* status;
    const int32_t* phase_count;
    const double* beta;
    const double* zfactor;
    const int32_t* branch;
    const double* x;
};
struct ClosureOut {
    int32_t* status;
    double* primal;    // [4][n]
    double* jac;       // [2nc+5][nc+1][n]
    double* log_k_jac; // [nc][nc+1][n] or null
    double* diag;      // [3][n] or null
};
struct ClosureOpt {
    double min_fraction;
    double min_rcond;
    double fug_tol;
};

// Failure classes of one Dual phase evaluation, named by the exception the reference
// raises: Pr76PhaseError, std::range_error, std::domain_error, std::out_of_range.
enum PhaseFailure : int { PF_OK = 0, PF_PHASE, PF_RANGE, PF_DOMAIN, PF_INDEX };

// Pr76Phase::evaluate_reduced<Dual> (pr76_phase.hpp:193-258) through
// Pr76Mixture::evaluate_impl<true, true> (pr76_mixture.hpp:124-219) and
// Pr76Pure::evaluate (pr76_pure.hpp:144-165), seeded with the chart directions
// p -> 0, T -> 1, x_k -> 2 + k; x_last = 1 - sum(x_red) in dual arithmetic. Roots are
// the FP64 bisection on the primal; the selected root's derivative is the first-order
// implicit-function step of pr76_phase.hpp:209-222. Writes Z and its gradient,
// ln(phi) and its Jacobian J[i * (NC+1) + k].
template <int NC>
__device__ __forceinline__ int phase_dual_body(const Mix<NC>& m, double p, double t, const double* xr, int root_index,
                                               double& zval, double* zgrad, double* lnphi, double* J) {
    constexpr int Q = NC + 1;
    using D = Dual<Q>;
    if (!finite(p) || !(p > 0)) { return PF_DOMAIN; }
    const D P = D::seed(p, 0);
    const D T = D::seed(t, 1);
    D isum(0.0);
    for (int k = 0; k + 1 < NC; ++k) {
        if (!finite(xr[k]) || xr[k] < 0 || xr[k] > 1) { return PF_DOMAIN; }
        isum = isum + D::seed(xr[k], 2 + k);
    }
    const D last = D(1.0) - isum;
    if (!last.finite() || last.v < 0 || last.v > 1) { return PF_DOMAIN; }
    if (!finite(t) || !(t > 0)) { return PF_DOMAIN; }
    D roots[NC], sums[NC];
    D att(0.0), cov(0.0);
    const D st = dsqrt(T);
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        const D xi = i == NC - 1 ? last : D::seed(xr[i], 2 + i);
        const D rr = st / m.sqrt_tc[i];
        const D br = 1.0 + m.kappa[i] * (1.0 - rr);
        if (!rr.finite() || !br.finite()) { return PF_RANGE; }
        const D a = (br * m.ac[i]) * br;
        if (!a.finite() || (a.v == 0 && br.v != 0)) { return PF_RANGE; }
        att = att + (xi * a) * xi;
        cov = cov + xi * m.b[i];
        sums[i] = xi * a;
        if (NC > 1) {
            if (!(a.v > 0)) { return PF_DOMAIN; }
            roots[i] = dsqrt(a);
            if (!roots[i].finite()) { return PF_RANGE; }
        }
    }
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        const D xi = i == NC - 1 ? last : D::seed(xr[i], 2 + i);
#pragma unroll
        for (int j = i + 1; j < NC; ++j) {
            const D xj = j == NC - 1 ? last : D::seed(xr[j], 2 + j);
            const double f = m.factor[i * NC + j];
            const D cross = ((xi * roots[i]) * (xj * roots[j])) * f;
            att = att + cross;
            att = att + cross;
            const D aij = (roots[i] * f) * roots[j];
            sums[i] = sums[i] + xj * aij;
            sums[j] = sums[j] + xi * aij;
        }
    }
    if (!att.finite() || !cov.finite() || !(cov.v > 0)) { return PF_RANGE; }
    for (int i = 0; i < NC; ++i) { if (!sums[i].finite()) { return PF_RANGE; } }
    // Pr76Phase::coefficients (pr76_phase.hpp:176-191)
    const D rt = m.r_gas * T;
    if (!rt.finite() || !(rt.v > 0)) { return PF_RANGE; }
    const D prt = P / rt;
    const D A = (att / rt) * prt;
    const D B = cov * prt;
    if (!A.finite() || !B.finite() || !(B.v > 0) || (A.v == 0 && att.v != 0)) { return PF_RANGE; }
    const RootSet rs = pr76_roots<false>(A.v, B.v);
    if (rs.status != RS_SUCCESS) { return PF_PHASE; }
    if (root_index < 0 || root_index >= rs.count) { return PF_INDEX; }
    const Root r = root_at(rs, root_index);
    if (!r.dvalid) { return PF_PHASE; }
    const double scale = smax(smax(1.0, B.v), __builtin_sqrt(__builtin_fabs(A.v)));
    const double y0 = r.x / scale;
    // pr76_cubic_coefficients in dual arithmetic (pr76_roots.hpp:50-58), scale held fixed
    const D bs = B / scale;
    const double inv = 1.0 / scale;
    const D c2 = 4.0 * bs - inv;
    const D c1 = (A / scale) / scale + (2.0 * bs) * bs - (4.0 * bs) * inv;
    const D c0 = ((-2.0 * bs) * bs) * inv;
    const D resid = ((D(y0) + c2) * y0 + c1) * y0 + c0;
    const double dh = (3.0 * y0 + 2.0 * c2.v) * y0 + c1.v;
    const D fv = D(r.x) - ((resid - D(resid.v)) / dh) * scale;
    const D z = fv + B;
    const double s2 = 1.4142135623730951; // sqrt(2), correctly rounded
    const D den = fv + (2.0 - s2) * B;
    const D u = (2.0 * s2) * (B / den);
    D lf;
    if (__builtin_fabs(u.v) <= 0.125) {
        // log1p(u)/u series of degree 19 (pr76_phase.hpp:67-79)
        D v(-1.0 / 20.0);
        for (int k = 18; k >= 0; --k) { v = (k % 2 == 0 ? 1.0 : -1.0) / static_cast<double>(k + 1) + u * v; }
        lf = v / den;
    } else {
        lf = (dlog1p(u) / u) / den;
    }
    D zm1 = z - 1.0;
    D lfree = dlog(fv);
    if (__builtin_fabs(zm1.v) < 0.25 && B.v < 0.125) {
        const D bz = B / z;
        zm1 = B / fv - (A / z) / ((1.0 + 2.0 * bz) - bz * bz);
        lfree = dlog1p(zm1 - B);
    }
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        const D ratio = m.b[i] / cov;
        const D coef = ((2.0 * sums[i] - att * ratio) / rt) * prt;
        const D l = (ratio * zm1 - lfree) - coef * lf;
        if (!l.finite()) { return PF_RANGE; }
        lnphi[i] = l.v;
        for (int k = 0; k < Q; ++k) { J[i * Q + k] = l.d[k]; }
    }
    if (!z.finite()) { return PF_RANGE; }
    zval = z.v;
    for (int k = 0; k < Q; ++k) { zgrad[k] = z.d[k]; }
    return PF_OK;
}

// Out-of-line copy for callers that evaluate several phases per lane.
template <int NC>
__device__ __noinline__ int phase_dual(const Mix<NC>& m, double p, double t, const double* xr, int root_index,
                                       double& zval, double* zgrad, double* lnphi, double* J) {
    return phase_dual_body<NC>(m, p, t, xr, root_index, zval, zgrad, lnphi, J);
}

__device__ __forceinline__ bool closure_same_z(double v, double accepted) { // pr76_thermodynamic_closure.hpp:70-79
    if (!finite(v) || !finite(accepted)) { return false; }
    return __builtin_fabs(v - accepted) <= 1024.0 * kEps * smax(1.0, __builtin_fabs(accepted));
}

// LU with partial pivoting, pivot guard, solve and 1/(|A|inf |A^-1|inf)
// (pt_sensitivity_detail.hpp:152-259), fixed size, in registers.
template <int N>
__device__ __forceinline__ bool lu_factor(double* a, int* piv, double& norm) {
    norm = 0;
    for (int r = 0; r < N; ++r) {
        double s = 0;
        for (int c = 0; c < N; ++c) { s += __builtin_fabs(a[r * N + c]); }
        norm = smax(norm, s);
    }
    if (!(norm > 0) || !finite(norm)) { return false; }
    const double guard = 64.0 * kEps * static_cast<double>(N) * norm;
    #pragma unroll 1
    for (int c = 0; c < N; ++c) {
        int p = c;
        double mag = __builtin_fabs(a[c * N + c]);
        for (int r = c + 1; r < N; ++r) {
            const double cand = __builtin_fabs(a[r * N + c]);
            if (cand > mag) { mag = cand; p = r; }
        }
        if (!finite(mag) || !(mag > guard)) { return false; }
        piv[c] = p;
        if (p != c) {
            for (int r = c + 1; r < N; ++r) {
                if (r == p) {
                    for (int j = 0; j < N; ++j) {
                        const double tmp = a[c * N + j];
                        a[c * N + j] = a[r * N + j];
                        a[r * N + j] = tmp;
                    }
                }
            }
        }
        const double dg = a[c * N + c];
        for (int r = c + 1; r < N; ++r) {
            const double mlt = a[r * N + c] / dg;
            a[r * N + c] = mlt;
            for (int j = c + 1; j < N; ++j) { a[r * N + j] -= mlt * a[c * N + j]; }
        }
    }
    return true;
}

template <int N>
__device__ __forceinline__ bool lu_solve(const double* a, const int* piv, double* rhs) {
    #pragma unroll 1
    for (int c = 0; c < N; ++c) {
        for (int r = c + 1; r < N; ++r) {
            if (piv[c] == r) { const double tmp = rhs[c]; rhs[c] = rhs[r]; rhs[r] = tmp; }
        }
    }
    for (int r = 0; r < N; ++r) {
        for (int c = 0; c < r; ++c) { rhs[r] -= a[r * N + c] * rhs[c]; }
    }
    bool ok = true;
    for (int r = N - 1; r >= 0; --r) {
        for (int c = r + 1; c < N; ++c) { rhs[r] -= a[r * N + c] * rhs[c]; }
        rhs[r] /= a[r * N + r];
        ok = ok && finite(rhs[r]);
        if (!ok) { break; }
    }
    return ok;
}

// One accepted two-phase cell. `word` receives the status word; outputs are written
// for the linearization only when it is available (the caller pre-fills NaN).
template <int NC>
__device__ __noinline__ int32_t closure_cell(const Mix<NC>& m, int64_t n, int64_t c, double P, double T,
                                             const double* __restrict__ z, const ClosureIn& in,
                                             const ClosureOut& out, const ClosureOpt& opt) {
    constexpr int Q = NC + 1;
    const double nan = __builtin_nan("");
    double x[NC], y[NC], feed[NC];
    for (int i = 0; i < NC; ++i) {
        x[i] = in.x[(0 * NC + i) * n + c];
        y[i] = in.x[(1 * NC + i) * n + c];
        feed[i] = z[i * n + c];
    }
    const double beta = in.beta[1 * n + c];
    const double zl = in.zfactor[0 * n + c];
    const double zv = in.zfactor[1 * n + c];
    const int bl = in.branch[0 * n + c] & 0xFF;
    const int bv = in.branch[1 * n + c] & 0xFF;
    auto word = [](int primal, int lin, int reason) { return static_cast<int32_t>(primal | (lin << 4) | (reason << 8)); };

    // ---- primal: reproduce Z at the full stored compositions (:215-269)
    if (!finite(beta) || !(beta > 0) || !(beta < 1) || !finite(zl) || !(zl > 0) || !finite(zv) || !(zv > 0)) {
        return word(CP_INDETERMINATE, 1, CR_NONE);
    }
    {
        PhaseCtx<NC> ctx;
        double lp[NC], z2 = nan;
        if (phase_prepare<NC, false>(m, P, T, x, ctx) != ISSUE_NONE || bl >= ctx.rs.count ||
            phase_lnphi<NC>(m, P, T, ctx, bl, lp, &z2) != ISSUE_NONE || !closure_same_z(z2, zl)) {
            return word(CP_INDETERMINATE, 1, CR_NONE);
        }
        if (phase_prepare<NC, false>(m, P, T, y, ctx) != ISSUE_NONE || bv >= ctx.rs.count ||
            phase_lnphi<NC>(m, P, T, ctx, bv, lp, &z2) != ISSUE_NONE || !closure_same_z(z2, zv)) {
            return word(CP_INDETERMINATE, 1, CR_NONE);
        }
    }
    const double rhol = P / ((zl * m.r_gas) * T);
    const double rhov = P / ((zv * m.r_gas) * T);
    if (!finite(rhol) || !(rhol > 0) || !finite(rhov) || !(rhov > 0)) { return word(CP_INDETERMINATE, 1, CR_NONE); }
    out.primal[0 * n + c] = zl;
    out.primal[1 * n + c] = zv;
    out.primal[2 * n + c] = rhol;
    out.primal[3 * n + c] = rhov;
    auto sens_fail = [&](int reason) {
        if (reason == CR_NOT_ACCEPTED) {
            for (int k = 0; k < 4; ++k) { out.primal[k * n + c] = nan; }
            return word(CP_INDETERMINATE, 1, reason);
        }
        return word(CP_VALID, 1, reason);
    };

    // ---- sensitivity guards (pr76_sensitivity.hpp:120-213)
    for (int i = 0; i < NC; ++i) { if (!(feed[i] > 0)) { return sens_fail(CR_FEED_SUPPORT); } }
    if (!(zv - zl > 1e-8 * smax(zv, zl))) { return sens_fail(CR_PHASE_BOUNDARY); }
    const double guard = smax(opt.min_fraction, 1e-10);
    if (!(beta > guard) || !(1.0 - beta > guard)) { return sens_fail(CR_PHASE_BOUNDARY); }
    for (int i = 0; i < NC; ++i) {
        if (!(x[i] > 0) || !(y[i] > 0)) { return sens_fail(CR_FEED_SUPPORT); }
        const double rec = __builtin_fma(beta, y[i], (1.0 - beta) * x[i]);
        const double err = __builtin_fabs(rec - feed[i]);
        if (err > 1e-12 || err / feed[i] > 1e-10) { return word(CP_INDETERMINATE, 1, CR_NONE) | kClosureThrown; }
    }

    // ---- RR manifold at the base point, reduced chart (pt_sensitivity_detail.hpp:53-100)
    double xt[NC], yt[NC], zt[NC];
    {
        double xs = 0, ys = 0, zs = 0;
        for (int i = 0; i + 1 < NC; ++i) {
            xt[i] = x[i]; yt[i] = y[i]; zt[i] = feed[i];
            xs += x[i]; ys += y[i]; zs += feed[i];
        }
        xt[NC - 1] = 1.0 - xs;
        yt[NC - 1] = 1.0 - ys;
        zt[NC - 1] = 1.0 - zs;
    }
    for (int i = 0; i < NC; ++i) {
        if (!(zt[i] > 0) || !(xt[i] > 0) || !(yt[i] > 0) || !finite(xt[i]) || !finite(yt[i])) {
            return sens_fail(CR_PHASE_BOUNDARY);
        }
    }
    double h = 0;
    for (int i = 0; i < NC; ++i) { const double d = yt[i] - xt[i]; h += d * d / zt[i]; }
    if (!finite(h) || !(h > 0)) { return sens_fail(CR_ILL_CONDITIONED); }

    // ---- one dual pass per phase over its chart
    double JL[NC * Q], JV[NC * Q], gl[Q], gv[Q], lnl[NC], lnv[NC], zlr = nan, zvr = nan;
    {
        int f = phase_dual<NC>(m, P, T, xt, bl, zlr, gl, lnl, JL);
        if (f == PF_OK) { f = phase_dual<NC>(m, P, T, yt, bv, zvr, gv, lnv, JV); }
        if (f == PF_INDEX) { return word(CP_INDETERMINATE, 1, CR_NONE) | kClosureThrown; }
        if (f == PF_PHASE || f == PF_RANGE) { return sens_fail(CR_PROPERTY_FAILURE); }
        if (f == PF_DOMAIN) { return sens_fail(CR_ARITHMETIC_FAILURE); }
    }
    double rn = 0;
    for (int i = 0; i < NC; ++i) {
        const double F = ((log(xt[i]) - log(yt[i])) + lnl[i]) - lnv[i];
        rn = smax(rn, __builtin_fabs(F));
    }
    if (out.diag) { out.diag[2 * n + c] = rn; }
    const double ascale = (8192.0 * kEps) * static_cast<double>(NC);
    if (rn > opt.fug_tol + ascale) { return sens_fail(CR_NOT_ACCEPTED); }

    // RR linearization map: seed (dlogK, dz with dz_last = -sum) -> d beta, dx, dy
    auto rr_map = [&](const double* dlk, const double* dz, double& db, double* dx, double* dy) {
        double num = 0;
        for (int i = 0; i < NC; ++i) {
            num += ((xt[i] * yt[i]) / zt[i]) * dlk[i] + ((yt[i] - xt[i]) / zt[i]) * dz[i];
        }
        db = num / h;
        double sx = 0, sy = 0;
        for (int i = 0; i + 1 < NC; ++i) {
            const double dlx = (dz[i] / zt[i] - ((yt[i] - xt[i]) / zt[i]) * db) - (beta * (yt[i] / zt[i])) * dlk[i];
            dx[i] = xt[i] * dlx;
            dy[i] = yt[i] * (dlx + dlk[i]);
            sx += dx[i];
            sy += dy[i];
        }
        dx[NC - 1] = 0.0 - sx;
        dy[NC - 1] = 0.0 - sy;
    };
    // dF for a seed, given its dx, dy and direct (dp, dT)
    auto dF = [&](int i, const double* dx, const double* dy, double dp, double dt) {
        double lp = JL[i * Q + 0] * dp + JL[i * Q + 1] * dt;
        double vp = JV[i * Q + 0] * dp + JV[i * Q + 1] * dt;
        for (int k = 0; k + 1 < NC; ++k) {
            lp += JL[i * Q + 2 + k] * dx[k];
            vp += JV[i * Q + 2 + k] * dy[k];
        }
        return ((dx[i] / xt[i] - dy[i] / yt[i]) + lp) - vp;
    };

    // ---- partial derivatives of the RR-manifold outputs (beta, x, y) for each seed
    // s = logK_0..logK_{nc-1}, z_0..z_{nc-2}: one unit seed at a time, so every partial
    // is the same number the reference's Dual lane carries. From them the equilibrium
    // Jacobian A = dF/dlogK and B = dF/dq (pr76_sensitivity.hpp:228-247).
    constexpr int S = 2 * NC - 1;
    double Pb[S], Px[NC * S], Py[NC * S];
    double Am[NC * NC], Bm[NC * Q];
    {
        double dlk[NC], dz[NC], dx[NC], dy[NC], db;
        #pragma unroll 1
        for (int sd = 0; sd < S; ++sd) {
            for (int i = 0; i < NC; ++i) {
                dlk[i] = i == sd ? 1.0 : 0.0;
                dz[i] = sd < NC ? 0.0 : (i == sd - NC ? 1.0 : (i == NC - 1 ? -1.0 : 0.0));
            }
            rr_map(dlk, dz, db, dx, dy);
            Pb[sd] = db;
            for (int i = 0; i < NC; ++i) {
                Px[i * S + sd] = dx[i];
                Py[i * S + sd] = dy[i];
                const double f = dF(i, dx, dy, 0.0, 0.0);
                if (sd < NC) { Am[i * NC + sd] = f; } else { Bm[i * Q + 2 + (sd - NC)] = f; }
            }
        }
        for (int i = 0; i < NC; ++i) {
            Bm[i * Q + 0] = JL[i * Q + 0] - JV[i * Q + 0];
            Bm[i * Q + 1] = JL[i * Q + 1] - JV[i * Q + 1];
        }
    }
    bool fin_all = true;
    for (int k = 0; k < NC * NC; ++k) { fin_all = fin_all && finite(Am[k]); }
    for (int k = 0; k < NC * Q; ++k) { fin_all = fin_all && finite(Bm[k]); }
    if (!fin_all) { return sens_fail(CR_ARITHMETIC_FAILURE); }

    // ---- factor, condition, implicit solve dlogK/dq = -A^-1 B (:249-310)
    double LU[NC * NC];
    int piv[NC];
    double norm;
    for (int k = 0; k < NC * NC; ++k) { LU[k] = Am[k]; }
    if (!lu_factor<NC>(LU, piv, norm)) { return sens_fail(CR_ILL_CONDITIONED); }
    double rcond = 0.0;
    {
        double rows[NC];
        for (int r = 0; r < NC; ++r) { rows[r] = 0.0; }
        bool ok = true;
        #pragma unroll 1
        for (int col = 0; col < NC; ++col) {
            double rhs[NC];
            for (int r = 0; r < NC; ++r) { rhs[r] = r == col ? 1.0 : 0.0; }
            ok = ok && lu_solve<NC>(LU, piv, rhs);
            for (int r = 0; r < NC; ++r) { rows[r] += __builtin_fabs(rhs[r]); }
        }
        if (ok) {
            double inv = rows[0];
            for (int r = 1; r < NC; ++r) { inv = smax(inv, rows[r]); }
            const double prod = norm * inv;
            rcond = (prod > 0 && finite(prod)) ? 1.0 / prod : 0.0;
        }
    }
    if (out.diag) { out.diag[0 * n + c] = rcond; }
    const double need = smax((256.0 * kEps) * static_cast<double>(NC), opt.min_rcond);
    if (!finite(rcond) || !(rcond > need)) { return sens_fail(CR_ILL_CONDITIONED); }
    double Dm[NC * Q];
    #pragma unroll 1
    for (int col = 0; col < Q; ++col) {
        double rhs[NC];
        for (int r = 0; r < NC; ++r) { rhs[r] = -Bm[r * Q + col]; }
        if (!lu_solve<NC>(LU, piv, rhs)) { return sens_fail(CR_ILL_CONDITIONED); }
        for (int r = 0; r < NC; ++r) { Dm[r * Q + col] = rhs[r]; }
    }
    double be = 0;
    #pragma unroll 1
    for (int col = 0; col < Q; ++col) {
        for (int r = 0; r < NC; ++r) {
            double res = Bm[r * Q + col];
            double sc = __builtin_fabs(res);
            #pragma unroll 1
            for (int j = 0; j < NC; ++j) {
                const double term = Am[r * NC + j] * Dm[j * Q + col];
                res += term;
                sc += __builtin_fabs(term);
            }
            be = smax(be, __builtin_fabs(res) / smax(1.0, sc));
        }
    }
    if (out.diag) { out.diag[1 * n + c] = be; }
    if (!finite(be) || be > ascale) { return sens_fail(CR_ARITHMETIC_FAILURE); }

    // ---- totals per column, checks, closure rows (:312-389; closure :287-357)
    bool ok = true;
    for (int k = 0; k < NC * Q; ++k) { ok = ok && finite(Dm[k]); }
    const int64_t rows = 2 * NC + 5;
    auto put = [&](int row, int col, double v) { out.jac[(static_cast<int64_t>(row) * Q + col) * n + c] = v; };
    #pragma unroll 1
    for (int col = 0; col < Q && ok; ++col) {
        // total = partial wrt q + sum_j partial wrt logK_j * dlogK_j/dq, in the reference's
        // order (pr76_sensitivity.hpp:312-323); RR outputs have no direct p, T partial
        double dx[NC], dy[NC], dz[NC];
        const int qs = col < 2 ? -1 : NC + (col - 2);
        double db = qs < 0 ? 0.0 : Pb[qs];
        for (int j = 0; j < NC; ++j) { db += Pb[j] * Dm[j * Q + col]; }
        for (int i = 0; i < NC; ++i) {
            double vx = qs < 0 ? 0.0 : Px[i * S + qs];
            double vy = qs < 0 ? 0.0 : Py[i * S + qs];
            for (int j = 0; j < NC; ++j) {
                vx += Px[i * S + j] * Dm[j * Q + col];
                vy += Py[i * S + j] * Dm[j * Q + col];
            }
            dx[i] = vx;
            dy[i] = vy;
            dz[i] = col < 2 ? 0.0 : (i == col - 2 ? 1.0 : (i == NC - 1 ? -1.0 : 0.0));
        }
        double ls = 0, vs = 0;
        ok = ok && finite(db);
        for (int i = 0; i < NC; ++i) {
            ok = ok && finite(dx[i]) && finite(dy[i]);
            ls += dx[i];
            vs += dy[i];
        }
        ok = ok && !(__builtin_fabs(ls) > ascale) && !(__builtin_fabs(vs) > ascale);
        for (int i = 0; i < NC; ++i) {
            const double der = (((1.0 - beta) * dx[i] + beta * dy[i]) + (yt[i] - xt[i]) * db) - dz[i];
            const double sc = smax(1.0, ((__builtin_fabs(dx[i]) + __builtin_fabs(dy[i])) + __builtin_fabs(db)) +
                                            __builtin_fabs(dz[i]));
            ok = ok && !(__builtin_fabs(der) > ascale * sc);
        }
        // pr76_phase_z_partials chained through the composition totals
        double dzl = col == 0 ? gl[0] : (col == 1 ? gl[1] : 0.0);
        double dzv = col == 0 ? gv[0] : (col == 1 ? gv[1] : 0.0);
        for (int k = 0; k + 1 < NC; ++k) {
            dzl += gl[2 + k] * dx[k];
            dzv += gv[2 + k] * dy[k];
        }
        const double dp = col == 0 ? 1.0 / P : 0.0;
        const double dt = col == 1 ? 1.0 / T : 0.0;
        const double drl = rhol * ((dp - dzl / zl) - dt);
        const double drv = rhov * ((dp - dzv / zv) - dt);
        put(0, col, db);
        for (int i = 0; i < NC; ++i) {
            put(1 + i, col, dx[i]);
            put(1 + NC + i, col, dy[i]);
        }
        put(2 * NC + 1, col, dzl);
        put(2 * NC + 2, col, dzv);
        put(2 * NC + 3, col, drl);
        put(2 * NC + 4, col, drv);
        ok = ok && finite(dzl) && finite(dzv) && finite(drl) && finite(drv);
    }
    if (!ok) {
        for (int64_t k = 0; k < rows * Q; ++k) { out.jac[k * n + c] = nan; }
        return sens_fail(CR_ARITHMETIC_FAILURE);
    }
    if (out.log_k_jac) {
        for (int k = 0; k < NC * Q; ++k) { out.log_k_jac[static_cast<int64_t>(k) * n + c] = Dm[k]; }
    }
    if (!closure_same_z(zlr, zl) || !closure_same_z(zvr, zv)) {
        for (int64_t k = 0; k < rows * Q; ++k) { out.jac[k * n + c] = nan; }
        return word(CP_VALID, 1, CR_ARITHMETIC_FAILURE);
    }
    return word(CP_VALID, 0, CR_NONE);
}

// Status of every cell; the accepted two-phase cells are appended to `list` (one
// atomic per wave) for k_closure_eval. Outputs of every cell start as NaN.
template <int NC>
__global__ __launch_bounds__(64) void k_closure_classify(int64_t n, ClosureIn in, ClosureOut out, int32_t* list,
                                                         unsigned long long* count) {
    constexpr int Q = NC + 1;
    const double nan = __builtin_nan("");
    for (int64_t base = static_cast<int64_t>(blockIdx.x) * blockDim.x; base < n;
         base += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t c = base + threadIdx.x;
        bool eligible = false;
        if (c < n) {
            const int32_t s = in.status[c];
            const int set = s & 0xF, max3 = (s >> 4) & 0xF, split = (s >> 8) & 0xF;
            int32_t w;
            if (s & kClosureThrown) {
                w = CP_INDETERMINATE | (1 << 4) | kClosureThrown;
            } else {
                const bool vle = max3 == 0xF;
                const bool two = vle ? split == V_TWO : max3 == M_TWO;
                const bool single = vle ? split == V_SINGLE : (max3 == M_SINGLE || max3 == M_THREE);
                eligible = two && set == 0 && in.phase_count[c] == 2;
                w = (single && set == 0) ? (CP_NOT_IMPLEMENTED | (1 << 4)) : (CP_INDETERMINATE | (1 << 4));
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) { out.primal[k * n + c] = nan; }
            for (int k = 0; k < (2 * NC + 5) * Q; ++k) { out.jac[static_cast<int64_t>(k) * n + c] = nan; }
            if (out.log_k_jac) {
                for (int k = 0; k < NC * Q; ++k) { out.log_k_jac[static_cast<int64_t>(k) * n + c] = nan; }
            }
            if (out.diag) {
#pragma unroll
                for (int k = 0; k < 3; ++k) { out.diag[k * n + c] = nan; }
            }
            if (!eligible) { out.status[c] = w; }
        }
        const unsigned long long mask = __ballot(eligible);
        if (mask) {
            const int leader = __ffsll(mask) - 1;
            unsigned long long slot = 0;
            if (static_cast<int>(__lane_id()) == leader) { slot = atomicAdd(count, static_cast<unsigned long long>(__popcll(mask))); }
            slot = __shfl(slot, leader);
            if (eligible) {
                const int rank = __popcll(mask & ((1ull << __lane_id()) - 1ull));
                list[slot + rank] = static_cast<int32_t>(c);
            }
        }
    }
}

template <int NC>
__global__ __launch_bounds__(64) void k_closure_eval(const double* __restrict__ mixc, int64_t n,
                                                     const double* __restrict__ p, const double* __restrict__ t,
                                                     const double* __restrict__ z, ClosureIn in, ClosureOut out,
                                                     ClosureOpt opt, const int32_t* __restrict__ list,
                                                     const unsigned long long* __restrict__ count) {
    __shared__ Mix<NC> smix;
    load_mix<NC>(mixc, smix);
    __syncthreads();
    const int64_t total = static_cast<int64_t>(*count);
    for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < total;
         k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t c = list[k];
        out.status[c] = closure_cell<NC>(smix, n, c, p[c], t[c], z, in, out, opt);
    }
}

static __global__ void k_closure_zero(unsigned long long* count) {
    if (threadIdx.x == 0) { *count = 0ull; }
}

// Workspace: the eligible-cell list (int32 per cell) and one counter.
inline size_t closure_workspace_bytes(int64_t n) {
    return 256 + ((static_cast<size_t>(n) * sizeof(int32_t) + 255) / 256) * 256;
}

template <int NC>
hipError_t launch_closure(const double* mixc, void* ws, size_t ws_bytes, int64_t n, const double* p, const double* t,
                          const double* z, const ClosureIn& in, const ClosureOut& out, const ClosureOpt& opt,
                          hipStream_t s) {
    if (!ws || ws_bytes < closure_workspace_bytes(n)) { return hipErrorInvalidValue; }
    auto* count = static_cast<unsigned long long*>(ws);
    auto* list = reinterpret_cast<int32_t*>(static_cast<char*>(ws) + 256);
    const unsigned blocks = static_cast<unsigned>(std::min<int64_t>((n + 63) / 64, 16384));
    static const unsigned eval_blocks = [] {
        int dev = 0, cus = 0, per_cu = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_closure_eval<NC>, 64, 0) != hipSuccess ||
            per_cu < 1) {
            per_cu = 1;
        }
        // 13 KB of private memory per lane at nc = 12: cap the grid (256 waves) so one
        // dispatch reserves little scratch; the loop strides over the eligible cells
        return std::min(static_cast<unsigned>((cus < 1 ? 1 : cus) * per_cu), 256u);
    }();
    hipLaunchKernelGGL(k_closure_zero, dim3(1), dim3(64), 0, s, count);
    hipLaunchKernelGGL(k_closure_classify<NC>, dim3(blocks), dim3(64), 0, s, n, in, out, list, count);
    hipLaunchKernelGGL(k_closure_eval<NC>, dim3(eval_blocks), dim3(64), 0, s, mixc, n, p, t, z, in, out, opt, list,
                       count);
    return hipGetLastError();
}

#define MPMC_FOR_EACH_NC2(X) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12)
#define MPMC_CLOSURE_SIG(N)                                                                                       \
    hipError_t launch_closure<N>(const double*, void*, size_t, int64_t, const double*, const double*,           \
                                 const double*, const ClosureIn&, const ClosureOut&, const ClosureOpt&, hipStream_t)
#define MPMC_DECLARE_CLOSURE(N) extern template MPMC_CLOSURE_SIG(N);
#define MPMC_INSTANTIATE_CLOSURE(N) template MPMC_CLOSURE_SIG(N);

} // namespace mpmc::dev
