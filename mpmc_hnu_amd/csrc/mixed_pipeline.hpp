// Mixed-precision stages of the staged batch flash (precision = 1).
//
// The FP64 schedule (flash_pipeline.hpp) runs every TPD descent and every split attempt
// of the reference in FP64 to the reference's tolerances. Here the searches run in the
// reduced precision of mixed_device.hpp, and FP64 is spent only where a result is
// decided or published:
//
//   k_prep            FP64, as in the FP64 schedule (reference phase of the feed)
//   k_tpd32(initial)  per (trial, cell): FP32 descent, classified with margins
//   k_tpd(resume)     FP64 continuation of every trial the FP32 descent left ambiguous,
//                     from its end point and iteration count, under the reference's rules
//   k_init_reduce     shared: witnesses in trial order
//   k_split32         per (attempt, unstable cell): FP32 successive substitution
//   k_split_gate      per unstable cell: when no attempt converged in FP32, its ambiguous
//                     attempts are handed to the FP64 continuation
//   k_split(resume)   FP64 continuation of those attempts from their logK
//   k_polish          per unstable cell: converged attempts in Gibbs order, each polished
//                     by FP64 Newton on logK with the Dual Jacobian until one passes the
//                     reference's acceptance tests in FP64; then the feed-Gibbs comparison
//                     and the review tolerance
//   k_tpd32(review)   per (trial, unstable cell): FP32 descent against the common tangent
//   k_tpd(resume)     FP64 continuation of the ambiguous review trials
//   k_review_reduce, k_max3_tail (FP64), k_publish, k_fallback: shared with the FP64 schedule
//
// A trial or attempt is "ambiguous" whenever its outcome is within reduced-precision
// reach of one of the reference's thresholds, or its search is slow enough to meet the
// reference's iteration limit; its decision is then taken by the FP64 continuation.
// The published phase set always comes from FP64 values.
#pragma once

namespace mpmc::dev {

// Margins of the reduced-precision classification.
// FP32 rounding of ln(phi) (terms of order 10) leaves a noise floor of about 1e-5 on the
// TPD residual and the split's fugacity residual: a search is converged in FP32 at the
// target below, or at the first rejected step once within the floor bound (its line
// search then only sees rounding).
constexpr float kStat32 = 2e-5f;        // TPD stationarity target
constexpr float kStatFloor32 = 1e-4f;   // TPD stationarity at which a rejected step ends the descent
constexpr float kTrivial32 = 1e-2f;     // max |ln w - ln w_trivial| of a trivial end point
constexpr float kNorm32 = 5e-5f;        // split fugacity residual target
constexpr float kNormFloor32 = 1e-3f;   // split residual at which a rejected step ends the attempt
// Evaluation caps of one FP32 trial / attempt: a longer search is left to FP64 (a TPD
// trial) or dropped from the candidate set (a split attempt), which bounds the tail of
// the persistent kernels.
constexpr uint32_t kTrialEvals32 = 300;
// A descent slower than this many steps may end at the reference's 512-step limit in FP64
// (an indeterminate trial); its outcome is left to FP64.
constexpr int kSlowIt32 = 128;
constexpr float kSlowTotal = 384.0f;
constexpr uint32_t kAttemptEvals32 = 400;

// ------------------------------------------------------------------ TPD (FP32)
template <int NC>
struct Tpd32Lane {
    int64_t slot, cell;
    int mid, k, mode, bt, it;
    uint32_t evals;
    double tol;
    Cell32<NC> c;
    double yprev[2];
    float d[NC];        // ln z_i + ln phi_i(reference) (initial) or the common log-activity (review)
    float w[NC], lw[NC], res[NC];
    float value, guard, stat;
    int branch;
    float dir[NC], cand[NC], candlw[NC];  // candidate w and ln w (w may underflow, ln w is kept)
    float slope, scale, alpha;
    float lrate;        // running mean of ln(stat_k / stat_k-1) over accepted steps
    uint32_t zmask;
    bool has_pt;        // an accepted point exists (else the candidate is the start point)
};

// Tangent-plane distance of the candidate in FP32; false when it cannot be formed.
// The composition is carried in log space: a trace component far below FP32's range keeps
// its ln w, and its w (underflowed to 0) drops out of the sums as it does to rounding in FP64.
template <int NC>
__device__ __forceinline__ bool tpd32_point(const Tpd32Lane<NC>& L, const float* w, const float* lwc, const float* lp,
                                            float* res, float& value, float& guard, float& stat) {
    float q[NC];
    float sum = 0.0f, val = 0.0f, mag = 1.0f;
    bool ok = true;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        q[i] = 0.0f;
        if (!((L.zmask >> i) & 1u)) { continue; }
        const float lw = lwc[i];
        ok = ok && __builtin_isfinite(lw);
        q[i] = (lw + lp[i]) - L.d[i];
        sum += w[i];
        val = __builtin_fmaf(w[i], q[i], val);
        mag += w[i] * (__builtin_fabsf(lw) + __builtin_fabsf(L.d[i]) + __builtin_fabsf(lp[i]));
    }
    const float mean = val / sum;
    float st = 0.0f;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        res[i] = 0.0f;
        if (!((L.zmask >> i) & 1u)) { continue; }
        res[i] = q[i] - mean;
        st = fmaxf(st, __builtin_fabsf(res[i]));
    }
    value = val;
    // each q_i carries a few FP32 roundings of terms bounded by the magnitude sum
    guard = 8.0f * kEps32 * mag;
    stat = st;
    return ok && __builtin_isfinite(val) && __builtin_isfinite(st) && __builtin_isfinite(mag);
}

// max_i |ln w_i - ln v_i| over the support; v from a [NC][n] double array (take_log) or
// its logarithms from a [NC][n] float array.
template <int NC>
__device__ __forceinline__ float log_gap32(const Tpd32Lane<NC>& L, const double* v, int64_t n, int64_t slot) {
    float g = 0.0f;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        if (!((L.zmask >> i) & 1u)) { continue; }
        g = fmaxf(g, __builtin_fabsf(L.lw[i] - __logf(static_cast<float>(v[i * n + slot]))));
    }
    return g;
}
template <int NC>
__device__ __forceinline__ float log_gap32(const Tpd32Lane<NC>& L, const float* lv, int64_t n, int64_t slot) {
    float g = 0.0f;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        if (!((L.zmask >> i) & 1u)) { continue; }
        g = fmaxf(g, __builtin_fabsf(L.lw[i] - lv[i * n + slot]));
    }
    return g;
}

// Classification of an end point of the FP32 descent: negative beyond rounding reach,
// stationary beyond rounding reach (positive, or at a trivial point), else ambiguous.
template <int NC>
__device__ __forceinline__ int tpd32_classify(const Pipe<NC>& P, int stage, const Tpd32Lane<NC>& L) {
    if (L.it >= kSlowIt32) { return T_AMBIG; }
    // Decision-exact schedule: the reference stops a descent only at stationarity 1e-8 or
    // after 512 steps. A descent whose recent contraction rate, extrapolated from the FP32
    // end point down to 1e-8, would need more than kSlowTotal steps in all may end at the
    // iteration limit in FP64: FP64 decides it.
    if (P.exact && L.it > 0) {
        const float need = __logf(fmaxf(L.stat, 1e-8f) * 1e8f);
        if (!(L.lrate < 0.0f) || static_cast<float>(L.it) - need / L.lrate > kSlowTotal) { return T_AMBIG; }
    }
    const float margin = 4.0f * L.guard + 1e-6f;
    if (L.value < -(static_cast<float>(L.tol) + margin)) { return T_NEGATIVE; }
    if (L.value > margin) { return T_STATIONARY; }
    const int64_t n = P.n;
    float gap;
    if (stage == 0) {
        gap = log_gap32<NC>(L, P.lz32, n, L.cell);
    } else {
        const int sel = P.u_sel[L.slot];
        gap = fminf(log_gap32<NC>(L, P.a_v + (static_cast<int64_t>(sel) * 5 + 0) * NC * n, n, L.slot),
                    log_gap32<NC>(L, P.a_v + (static_cast<int64_t>(sel) * 5 + 1) * NC * n, n, L.slot));
    }
    return gap <= kTrivial32 ? T_STATIONARY : T_AMBIG;
}

template <int NC>
__device__ bool tpd32_begin(const Pipe<NC>& P, const Mix<NC>* smix, int stage, int k, int64_t slot,
                            Tpd32Lane<NC>& L) {
    const int64_t n = P.n;
    int32_t* st_arr = stage == 0 ? P.tst : P.fst;
    int64_t c;
    int ntr;
    const int32_t info0 = stage == 0 ? P.refinfo[slot] : 0;
    if (stage == 0) {
        c = slot;
        if (P.flags[c] & CF_REJECTED) { return false; }
        const int active = active_count<NC>(info0);
        ntr = (active > 1 ? active + 2 : 1) + hint_count<NC>(info0);
        if (!(info0 & 0x200) || k >= ntr) { st_arr[k * n + slot] = -1; return false; }
    } else {
        const int sel = P.u_sel[slot];
        if (sel < 0 || !finite(P.u_ftol[slot])) { st_arr[k * n + slot] = -1; return false; }
        c = P.u_cell[slot];
        if (P.flags[c] & (CF_REJECTED | CF_FALLBACK)) { st_arr[k * n + slot] = -1; return false; }
        ntr = review_trials<NC>(P.refinfo[c]);
        if (k >= ntr) { st_arr[k * n + slot] = -1; return false; }
    }
    L.slot = slot;
    L.cell = c;
    L.mid = mix_at(P, c);
    L.k = k;
    double z[NC];
    L.zmask = 0u;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        z[i] = P.feed[i * n + c];
        if (z[i] > 0) { L.zmask |= 1u << i; }
    }
    const int active = active_count<NC>(P.refinfo[c]);
    const int gen = active > 1 ? active + 2 : 1;
    bool fin = true;
    if (stage == 0) {
        L.tol = 1e-10;
#pragma unroll
        for (int i = 0; i < NC; ++i) { L.d[i] = z[i] > 0 ? P.d32[i * n + c] : 0.0f; }
    } else {
        const int sel = P.u_sel[slot];
        L.tol = P.u_ftol[slot];
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            L.d[i] = z[i] > 0 ? static_cast<float>(P.a_v[((sel * 5 + 4) * NC + i) * n + slot]) : 0.0f;
            fin = fin && __builtin_isfinite(L.d[i]);
        }
    }
    double w0[NC];
    if (k < gen) {
        trial_start<NC>(z, active, k, w0);
    } else if (stage == 0) {
        hint_start<NC>(P, c, k - gen, w0);
    } else if (k - gen < final_hint_count<NC>(P.refinfo[c])) {
        nth_hint<NC>(P.fhints, P.nfhint, n, c, k - gen, w0);
    } else {
        const int sel = P.u_sel[slot];
        const int q = k - gen - final_hint_count<NC>(P.refinfo[c]);
#pragma unroll
        for (int i = 0; i < NC; ++i) { w0[i] = P.a_v[((sel * 5 + q) * NC + i) * n + slot]; }
    }
    const double s = ksum<NC>(w0);
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        L.cand[i] = (L.zmask >> i) & 1u ? static_cast<float>(w0[i] / s) : 0.0f;
        L.candlw[i] = (L.zmask >> i) & 1u ? __logf(L.cand[i]) : 0.0f;
        fin = fin && ((L.zmask >> i) & 1u ? L.cand[i] > 0.0f : true);
    }
    fin = fin && cell32_load<NC>(smix[L.mid], P.ra32, n, c, P.p[c], P.t[c], L.c);
    L.yprev[0] = -1.0;
    L.yprev[1] = -1.0;
    L.branch = 0;
    L.has_pt = false;
    L.evals = 0;
    L.it = 0;
    L.lrate = 0.0f;
    L.mode = TM_START;
    if (!fin) {
        st_arr[k * n + slot] = T_AMBIG;
        (stage == 0 ? P.tev : P.fev)[k * n + slot] = 0;
        return false;
    }
    return true;
}

// An ambiguous trial is handed to the FP64 continuation (T_RESUME) from its current point,
// or, in the decision-exact schedule, to an FP64 search from its start (T_REDO).
template <int NC>
__device__ void tpd32_finish(const Pipe<NC>& P, int stage, const Tpd32Lane<NC>& L, int status) {
    const int64_t n = P.n;
    const int64_t idx = L.k * n + L.slot;
    if (P.exact && status == T_AMBIG) {
        (stage == 0 ? P.tst : P.fst)[idx] = T_REDO;
        (stage == 0 ? P.tev : P.fev)[idx] = L.evals;
        atomicAdd(&P.cnt[0 * n + L.cell], L.evals);
        atomicAdd(&P.cnt[1 * n + L.cell], L.evals);
        return;
    }
    const bool resume = status == T_AMBIG;
    (stage == 0 ? P.tst : P.fst)[idx] = resume ? (T_RESUME | (L.it << 16)) : (status | (L.branch << 8));
    (stage == 0 ? P.tev : P.fev)[idx] = L.evals;
    if (status == T_NEGATIVE || resume) {
        const float* lw = L.has_pt ? L.lw : L.candlw;
        double w[NC];
#pragma unroll
        for (int i = 0; i < NC; ++i) { w[i] = (L.zmask >> i) & 1u ? exp(static_cast<double>(lw[i])) : 0.0; }
        const double s = ksum<NC>(w);
        double* dst = stage == 0 ? P.tw : P.fw;
#pragma unroll
        for (int i = 0; i < NC; ++i) { dst[(L.k * NC + i) * n + L.slot] = w[i] / s; }
    }
    atomicAdd(&P.cnt[0 * n + L.cell], L.evals);
    atomicAdd(&P.cnt[1 * n + L.cell], L.evals);
}

// Termination tests at an accepted point and the next descent direction (the reference's
// rules, pt_stability.hpp:528-561, with reduced-precision classification). Returns a
// trial status or -1 to start backtracking.
template <int NC>
__device__ __forceinline__ int tpd32_check(const Pipe<NC>& P, int stage, Tpd32Lane<NC>& L) {
    if (L.value < -(static_cast<float>(L.tol) + 4.0f * L.guard + 1e-6f)) { return T_NEGATIVE; }
    if (L.stat <= kStat32) { return tpd32_classify<NC>(P, stage, L); }
    if (L.it >= kSlowIt32) { return T_AMBIG; }
    if (L.it >= 512 || L.evals >= kTrialEvals32) { return T_AMBIG; }
    L.scale = fmaxf(1.0f, L.stat / 4.0f);
    const float inv = 1.0f / L.scale;
    float slope = 0.0f;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        L.dir[i] = -L.res[i] * inv;
        slope = __builtin_fmaf(L.w[i] * L.res[i], L.dir[i], slope);
    }
    // a zero slope (the residual left only on components whose w underflowed) still descends
    if (!(slope <= 0.0f)) { return L.stat <= kStatFloor32 ? tpd32_classify<NC>(P, stage, L) : T_AMBIG; }
    L.slope = slope;
    L.mode = TM_BT;
    L.bt = 0;
    L.alpha = 1.0f;
    return -1;
}

template <int NC>
__global__ __launch_bounds__(64) MPMC_SEARCH_ATTR void k_tpd32(Pipe<NC> P, const double* __restrict__ mixc, int stage) {
    Mix<NC>* const smix = stage_mixes<NC>(mixc, P.nmix);
    const int tk = stage == 0 ? NC + 2 + P.nhint : NC + 4 + P.nfhint;
    const int64_t nslot = stage == 0 ? P.n : static_cast<int64_t>(P.g[G_NU]);
    const int64_t total = tk * nslot;
    unsigned long long* counter = &P.g[stage == 0 ? G_TASK_INIT32 : G_TASK_REVIEW32];
    Tpd32Lane<NC> L;
    bool active = false, exhausted = false;
    for (;;) {
        const bool want = !active && !exhausted;
        const int64_t task = fetch_task(counter, want);
        if (want) {
            if (task >= total) { exhausted = true; }
            else {
                int k;
                int64_t slot;
                task_index<MPMC_TPD_ORDER>(task, nslot, tk, k, slot);
                active = tpd32_begin<NC>(P, smix, stage, k, slot, L);
            }
        }
        if (!__any(active)) {
            if (!__any(!exhausted)) { break; }
            continue;
        }
        if (!active) { continue; }
        const bool do_eval = true;
        if (L.mode == TM_BT) {
            float largest = -__builtin_huge_valf();
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                if ((L.zmask >> i) & 1u) {
                    L.cand[i] = __builtin_fmaf(L.alpha, L.dir[i], L.lw[i]);
                    largest = fmaxf(largest, L.cand[i]);
                }
            }
            float s = 0.0f;
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                L.candlw[i] = L.cand[i] - largest;
                L.cand[i] = (L.zmask >> i) & 1u ? __expf(L.candlw[i]) : 0.0f;
                s += L.cand[i];
            }
            const float inv = 1.0f / s, ls = __logf(s);
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                L.cand[i] *= inv;
                L.candlw[i] = (L.zmask >> i) & 1u ? L.candlw[i] - ls : 0.0f;
            }
        }
        float lp[NC], nres[NC], nval = 0.0f, ngd = 0.0f, nst = 0.0f;
        int br = 0;
        bool ok_eval = false;
        if (do_eval) {
            ++L.evals;
            double zz;
            ok_eval = provider32<NC>(smix[L.mid], L.c, L.cand, L.yprev, lp, br, zz) == ISSUE_NONE &&
                      tpd32_point<NC>(L, L.cand, L.candlw, lp, nres, nval, ngd, nst);
        }
        int done = -1;
        if (L.mode == TM_START) {
            if (!ok_eval) { done = T_AMBIG; }
            else {
#pragma unroll
                for (int i = 0; i < NC; ++i) { L.w[i] = L.cand[i]; L.lw[i] = L.candlw[i]; L.res[i] = nres[i]; }
                L.value = nval; L.guard = ngd; L.stat = nst; L.branch = br;
                L.has_pt = true;
                done = tpd32_check<NC>(P, stage, L);
            }
        } else {
            bool ok = false;
            if (ok_eval) {
                ok = nval < -(static_cast<float>(L.tol) + 4.0f * ngd + 1e-6f);
                if (!ok) {
                    // stability_accept_step (pt_stability.hpp:312-326) at FP32 resolution: the
                    // Armijo decrease is accepted at any step size (the FP64 rule's guard is
                    // below every step the search takes), residual progress within the guard
                    const float pred = -L.alpha * L.slope;
                    const float gd = L.guard + ngd;
                    ok = nval <= L.value - 1e-4f * pred || (pred <= gd && nval <= L.value + gd && nst < L.stat);
                }
            }
            if (ok) {
                const float lr = nst > 0.0f && L.stat > 0.0f ? __logf(nst / L.stat) : 0.0f;
                L.lrate = L.it == 0 ? lr : 0.7f * L.lrate + 0.3f * lr;
#pragma unroll
                for (int i = 0; i < NC; ++i) { L.w[i] = L.cand[i]; L.lw[i] = L.candlw[i]; L.res[i] = nres[i]; }
                L.value = nval; L.guard = ngd; L.stat = nst; L.branch = br;
                ++L.it;
                done = tpd32_check<NC>(P, stage, L);
            } else if (L.stat <= kStatFloor32) {
                done = tpd32_classify<NC>(P, stage, L);
            } else {
                ++L.bt;
                L.alpha *= 0.5f;
                if (L.bt >= 32 || L.evals >= kTrialEvals32) { done = T_AMBIG; }
            }
        }
        if (done >= 0) {
            tpd32_finish<NC>(P, stage, L, done);
            active = false;
        }
    }
}

// ------------------------------------------------------------------ split (FP32)
template <int NC>
struct Split32Lane {
    int64_t u, cell;
    int mid, j, mode, bt, it;
    uint32_t evals, rr;
    Cell32<NC> c;
    double yl, yv;
    float z[NC];
    float lk[NC], r[NC], x[NC], y[NC];
    float beta, norm, gibbs, guard, zl, zv;
    float scale, slope, alpha;
    float cand[NC];
};

// One evaluation of iterate_pt_split's evaluate() (pt_split.hpp:300-341) in reduced
// precision. Returns -1 (point formed), A_NO_INTERIOR, or A_AMBIG.
template <int NC>
__device__ __forceinline__ int split32_eval(const Mix<NC>& m, Split32Lane<NC>& L, const float* lk, float guess,
                                            float* nlk, float* nr, float* nx, float* ny, float& nbeta, float& nnorm,
                                            float& ngibbs, float& nguard, float& nzl, float& nzv) {
    ++L.rr;
    float lx[NC], ly[NC];
    const int e = rr32<NC>(L.z, lk, guess, nbeta, nx, ny, lx, ly);
    if (e == 1) { return A_NO_INTERIOR; }
    if (e != 0) { return A_AMBIG; }
    float ll[NC], lv[NC];
    double zl, zv;
    L.evals += 2;
    if (side32<NC>(m, L.c, nx, 0, L.yl, ll, zl) != ISSUE_NONE) { return A_AMBIG; }
    if (side32<NC>(m, L.c, ny, 1, L.yv, lv, zv) != ISSUE_NONE) { return A_AMBIG; }
    float g = 0.0f, mag = 1.0f, nm = 0.0f;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        nlk[i] = 0.0f;
        nr[i] = 0.0f;
        if (!(L.z[i] > 0.0f)) { continue; }
        const float mul = lx[i] + ll[i], muv = ly[i] + lv[i];
        nlk[i] = ly[i] - lx[i];
        nr[i] = (lx[i] - ly[i]) + (ll[i] - lv[i]);
        nm = fmaxf(nm, __builtin_fabsf(nr[i]));
        g += (1.0f - nbeta) * nx[i] * mul + nbeta * ny[i] * muv;
        mag += (1.0f - nbeta) * nx[i] * (__builtin_fabsf(lx[i]) + __builtin_fabsf(ll[i])) +
               nbeta * ny[i] * (__builtin_fabsf(ly[i]) + __builtin_fabsf(lv[i]));
    }
    nnorm = nm;
    ngibbs = g;
    // the split's Gibbs sum carries a few roundings per term: a tight guard keeps the FP32
    // line search rejecting the Gibbs increases the reference's FP64 Armijo test rejects
    nguard = 8.0f * kEps32 * mag;
    nzl = static_cast<float>(zl);
    nzv = static_cast<float>(zv);
    if (!__builtin_isfinite(g) || !__builtin_isfinite(nm) || !(nzl > 0.0f) || !(nzv > 0.0f)) { return A_AMBIG; }
    return -1;
}

// Converged in reduced precision: the reference's acceptance tests (pt_split.hpp:349-361),
// each either decided beyond rounding reach or left to FP64.
template <int NC>
__device__ __forceinline__ int split32_accept(const Split32Lane<NC>& L) {
    float contrast = 0.0f;
#pragma unroll
    for (int i = 0; i < NC; ++i) { if (L.z[i] > 0.0f) { contrast = fmaxf(contrast, __builtin_fabsf(L.lk[i])); } }
    // An attempt closing in on the trivial solution (K -> 1) reaches the FP32 floor with a
    // small contrast: not a two-phase candidate. A near-critical split that FP64 would
    // accept is lost only if no other attempt converges, and then the cell goes to FP64.
    if (contrast < 1e-2f) { return A_INDISTINGUISHABLE; }
    if (L.beta <= 1e-5f || L.beta >= 1.0f - 1e-5f) { return A_AMBIG; }
    // both candidates on one root at FP32 resolution: the reference's indistinguishable pair
    if (L.zv - L.zl <= 1e-4f * fmaxf(L.zv, L.zl)) { return A_INDISTINGUISHABLE; }
    return A_CONVERGED;
}

template <int NC>
__device__ __forceinline__ int split32_check(Split32Lane<NC>& L) {
    if (L.norm <= kNorm32) { return split32_accept<NC>(L); }
    if (L.it >= 512 || L.evals >= kAttemptEvals32) { return A_AMBIG; }
    L.scale = fmaxf(1.0f, L.norm / 2.0f);
    float h = 0.0f, f = 0.0f, s2 = 0.0f;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        if (!(L.z[i] > 0.0f)) { continue; }
        const float c = (L.x[i] / L.z[i]) * L.y[i];
        const float d = L.y[i] - L.x[i];
        h += d * (d / L.z[i]);
        f += c * L.r[i];
        s2 += (c * L.r[i]) * L.r[i];
    }
    const float slope = -(L.beta * (1.0f - L.beta) * s2 + (f / h) * f) / L.scale;
    if (!(h > 0.0f) || !__builtin_isfinite(slope) || !(slope < 0.0f)) {
        return L.norm <= kNormFloor32 ? split32_accept<NC>(L) : A_AMBIG;
    }
    L.slope = slope;
    L.mode = TM_BT;
    L.bt = 0;
    L.alpha = 1.0f;
    return -1;
}

template <int NC>
__device__ bool split32_begin(const Pipe<NC>& P, const Mix<NC>* smix, int j, int64_t u, Split32Lane<NC>& L) {
    const int64_t n = P.n;
    if (j >= 2 * P.u_nwit[u]) { P.a_st[j * n + u] = -1; P.a_ev[j * n + u] = 0; return false; }
    const int64_t c = P.u_cell[u];
    L.u = u;
    L.cell = c;
    L.j = j;
    L.mid = mix_at(P, c);
    double z[NC], w[NC], lk0[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        z[i] = P.feed[i * n + c];
        L.z[i] = static_cast<float>(z[i]);
    }
    const int k = P.u_wit[(j / 2) * n + u];
#pragma unroll
    for (int i = 0; i < NC; ++i) { w[i] = P.tw[(k * NC + i) * n + c]; }
    L.evals = 0;
    L.rr = 0;
    L.it = 0;
    L.mode = TM_START;
    L.yl = -1.0;
    L.yv = -1.0;
    L.beta = 0.5f;
    const bool ok = cell32_load<NC>(smix[L.mid], P.ra32, n, c, P.p[c], P.t[c], L.c);
    if (!split_seed<NC>(z, w, (j % 2) == 0, lk0)) {
        P.a_st[j * n + u] = A_UNREPRESENTABLE_SEED;
        P.a_ev[j * n + u] = 0;
        return false;
    }
    if (!ok) { P.a_st[j * n + u] = A_AMBIG; P.a_ev[j * n + u] = 0; return false; }
#pragma unroll
    for (int i = 0; i < NC; ++i) { L.cand[i] = static_cast<float>(lk0[i]); }
    return true;
}

// Converged attempts are marked kReduced (their logK is in the fifth vector slot);
// ambiguous ones keep their current logK and iteration count for a possible FP64
// continuation.
constexpr int32_t kReduced = 1 << 24;
template <int NC>
__device__ void split32_finish(const Pipe<NC>& P, const Split32Lane<NC>& L, int status) {
    const int64_t n = P.n;
    const int64_t idx = L.j * n + L.u;
    P.a_ev[idx] = L.evals;
    const bool keep = status == A_AMBIG || status == A_INDISTINGUISHABLE;
    P.a_st[idx] = status == A_CONVERGED ? (status | kReduced) : (keep ? (status | (L.it << 16)) : status);
    if (keep) {
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            P.a_v[((L.j * 5 + 4) * NC + i) * n + L.u] = L.mode == TM_START ? L.cand[i] : L.lk[i];
        }
    }
    if (status == A_CONVERGED) {
        const double s[6] = {L.gibbs, L.beta, L.zl, L.zv, L.norm, L.guard};
#pragma unroll
        for (int q = 0; q < 6; ++q) { P.a_s[(L.j * 6 + q) * n + L.u] = s[q]; }
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            P.a_v[((L.j * 5 + 0) * NC + i) * n + L.u] = L.x[i];
            P.a_v[((L.j * 5 + 1) * NC + i) * n + L.u] = L.y[i];
            P.a_v[((L.j * 5 + 4) * NC + i) * n + L.u] = L.lk[i];
        }
    }
    atomicAdd(&P.cnt[0 * n + L.cell], L.evals);
    atomicAdd(&P.cnt[2 * n + L.cell], L.rr);
}

template <int NC>
__global__ __launch_bounds__(64) MPMC_SEARCH_ATTR void k_split32(Pipe<NC> P, const double* __restrict__ mixc) {
    Mix<NC>* const smix = stage_mixes<NC>(mixc, P.nmix);
    const int64_t nu = static_cast<int64_t>(P.g[G_NU]);
    const int64_t total = kAttempts * nu;
    Split32Lane<NC> L;
    bool active = false, exhausted = false;
    for (;;) {
        const bool want = !active && !exhausted;
        const int64_t task = fetch_task(&P.g[G_TASK_SPLIT32], want);
        if (want) {
            if (task >= total) { exhausted = true; }
            else {
                int k;
                int64_t u;
                task_index<MPMC_SPLIT_ORDER>(task, nu, kAttempts, k, u);
                if (MPMC_SPLIT_LPT) { u = P.u_order[u]; }
                active = split32_begin<NC>(P, smix, k, u, L);
            }
        }
        if (!__any(active)) {
            if (!__any(!exhausted)) { break; }
            continue;
        }
        if (!active) { continue; }
        if (L.mode == TM_BT) {
            const float step = L.alpha / L.scale;
#pragma unroll
            for (int i = 0; i < NC; ++i) { L.cand[i] = __builtin_fmaf(step, L.r[i], L.lk[i]); }
        }
        float nlk[NC], nr[NC], nx[NC], ny[NC], nb = 0.0f, nn = 0.0f, ng = 0.0f, ngd = 0.0f, nzl = 0.0f, nzv = 0.0f;
        const int e = split32_eval<NC>(smix[L.mid], L, L.cand, L.beta, nlk, nr, nx, ny, nb, nn, ng, ngd, nzl, nzv);
        int done = -1;
        bool accept = false;
        if (L.mode == TM_START) {
            if (e != -1) { done = e; }
            else { accept = true; }
        } else if (e == -1) {
            const float pred = -L.alpha * L.slope;
            const float gd = L.guard + ngd;
            // split_accept_step (pt_split.hpp:232-244) at FP32 resolution, the Armijo decrease
            // accepted at any step size
            accept = ng <= L.gibbs - 1e-4f * pred ||
                     (pred <= gd && ng <= L.gibbs + gd &&
                      (nn <= kNorm32 || L.norm - nn > 1e-4f * (L.alpha / L.scale) * L.norm));
        }
        if (accept) {
#pragma unroll
            for (int i = 0; i < NC; ++i) { L.lk[i] = nlk[i]; L.r[i] = nr[i]; L.x[i] = nx[i]; L.y[i] = ny[i]; }
            L.beta = nb; L.norm = nn; L.gibbs = ng; L.guard = ngd; L.zl = nzl; L.zv = nzv;
            if (L.mode != TM_START) { ++L.it; }
            done = split32_check<NC>(L);
        } else if (done < 0 && L.norm <= kNormFloor32) {
            done = split32_accept<NC>(L);
        } else if (done < 0) {
            ++L.bt;
            L.alpha *= 0.5f;
            if (L.bt >= 24 || L.evals >= kAttemptEvals32) { done = A_AMBIG; }
        }
        if (done >= 0) {
            split32_finish<NC>(P, L, done);
            active = false;
        }
    }
}

// ------------------------------------------------------------------ FP64 polish
// dF/dlogK of the equilibrium residual F = ln x - ln y + ln phi_L(x) - ln phi_V(y) on the
// Rachford-Rice manifold (the system of pt_sensitivity_detail.hpp:45-143 with dz = 0), from
// one Dual<nc+1> pass per phase.
template <int NC>
__device__ __noinline__ bool split_jacobian(const Mix<NC>& m, double p, double t, const double* z,
                                            const SplitState<NC>& s, double* A) {
    constexpr int Q = NC + 1;
    double JL[NC * Q], JV[NC * Q], gl[Q], gv[Q], lnl[NC], lnv[NC], zlr, zvr;
    if (phase_dual<NC>(m, p, t, s.fr.x, s.liq.branch, zlr, gl, lnl, JL) != PF_OK) { return false; }
    if (phase_dual<NC>(m, p, t, s.fr.y, s.vap.branch, zvr, gv, lnv, JV) != PF_OK) { return false; }
    const double beta = s.fr.beta;
    const double* x = s.fr.x;
    const double* y = s.fr.y;
    double h = 0;
#pragma unroll
    for (int i = 0; i < NC; ++i) { const double d = y[i] - x[i]; h += d * d / z[i]; }
    if (!finite(h) || !(h > 0)) { return false; }
#pragma unroll 1
    for (int j = 0; j < NC; ++j) {
        const double db = ((x[j] * y[j]) / z[j]) / h;
        double dx[NC], dy[NC], sx = 0, sy = 0;
        for (int i = 0; i + 1 < NC; ++i) {
            const double dlx = -((y[i] - x[i]) / z[i]) * db - (i == j ? beta * (y[i] / z[i]) : 0.0);
            dx[i] = x[i] * dlx;
            dy[i] = y[i] * (dlx + (i == j ? 1.0 : 0.0));
            sx += dx[i];
            sy += dy[i];
        }
        dx[NC - 1] = -sx;
        dy[NC - 1] = -sy;
        for (int i = 0; i < NC; ++i) {
            double a = dx[i] / x[i] - dy[i] / y[i];
            for (int k = 0; k + 1 < NC; ++k) { a += JL[i * Q + 2 + k] * dx[k] - JV[i * Q + 2 + k] * dy[k]; }
            A[i * NC + j] = a;
        }
    }
    return true;
}

// Per unstable cell: when no attempt converged in reduced precision, its ambiguous and
// indistinguishable attempts are continued in FP64 from their current logK (an FP32
// "indistinguishable" end can be a near-critical split the reference accepts).
template <int NC>
__global__ __launch_bounds__(64) void k_split_gate(Pipe<NC> P) {
    const int64_t n = P.n;
    const int64_t nu = static_cast<int64_t>(P.g[G_NU]);
    for (int64_t u = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; u < nu;
         u += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int na = 2 * P.u_nwit[u];
        bool conv = false;
        for (int j = 0; j < na; ++j) { conv = conv || (P.a_st[j * n + u] & 0xFF) == A_CONVERGED; }
        if (conv) { continue; }
        for (int j = 0; j < na; ++j) {
            const int32_t st = P.a_st[j * n + u];
            if ((st & 0xFF) == A_AMBIG || (st & 0xFF) == A_INDISTINGUISHABLE) { P.a_st[j * n + u] = A_RESUME | (st & ~0xFF); }
        }
    }
}

// logK of a converged attempt: its reduced-precision iterate, or ln y - ln x of an FP64 one.
template <int NC>
__device__ __forceinline__ void attempt_logk(const Pipe<NC>& P, int j, int64_t u, double* lk) {
    const int64_t n = P.n;
    const bool reduced = (P.a_st[j * n + u] & kReduced) != 0;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        lk[i] = reduced ? P.a_v[((j * 5 + 4) * NC + i) * n + u]
                        : log(P.a_v[((j * 5 + 1) * NC + i) * n + u]) - log(P.a_v[((j * 5 + 0) * NC + i) * n + u]);
    }
}

// Newton on logK with step halving, from lk: converged at 1e-2 fug_tol, or within fug_tol
// once a step no longer reduces the residual by 4x (the FP64 rounding floor).
template <int NC>
__device__ __noinline__ bool polish_split(const Mix<NC>& m, double p, double t, const double* z, double* lk,
                                          double beta0, double fug_tol, SplitState<NC>& s, Counters& cnt,
                                          bool& rejected) {
    uint32_t evals = 0;
    SplitState<NC> trial;
    if (split_evaluate<NC, false>(m, p, t, z, lk, evals, 0xFFFFFFFFu, s, cnt, rejected, beta0) != -1 || rejected) {
        return false;
    }
    for (int it = 0; it < 12; ++it) {
        if (s.norm <= 1e-2 * fug_tol) { return true; }
        double Am[NC * NC], dl[NC];
        int piv[NC];
        double nrm;
        if (!split_jacobian<NC>(m, p, t, z, s, Am) || !lu_factor<NC>(Am, piv, nrm)) { return false; }
#pragma unroll
        for (int i = 0; i < NC; ++i) { dl[i] = -s.r[i]; }
        if (!lu_solve<NC>(Am, piv, dl)) { return false; }
        double step = 1.0;
        bool took = false;
        for (int bt = 0; bt < 5 && !took; ++bt, step *= 0.5) {
#pragma unroll
            for (int i = 0; i < NC; ++i) { lk[i] = s.lk[i] + step * dl[i]; }
            const int e = split_evaluate<NC, false>(m, p, t, z, lk, evals, 0xFFFFFFFFu, trial, cnt, rejected, s.fr.beta);
            if (rejected) { return false; }
            took = e == -1 && trial.norm < s.norm;
        }
        if (!took) { return s.norm <= fug_tol; }
        const bool stalled = trial.norm > 0.25 * s.norm;
        s = trial;
        if (stalled && s.norm <= fug_tol) { return true; }
    }
    return false;
}

// Per unstable cell: the converged attempts in Gibbs order (at most three tried), each
// polished by FP64 Newton on logK until one passes the reference's acceptance tests
// (pt_split.hpp:349-361) in FP64; then the selection epilogue of k_split_select
// (pt_split.hpp:459-507). No converged attempt: the reference's indeterminate outcome.
template <int NC>
__global__ __launch_bounds__(64) void k_polish(Pipe<NC> P, const double* __restrict__ mixc) {
    Mix<NC>* const smix = stage_mixes<NC>(mixc, P.nmix);
    const int64_t n = P.n;
    const int64_t nu = static_cast<int64_t>(P.g[G_NU]);
    for (int64_t u = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; u < nu;
         u += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t c = P.u_cell[u];
        P.u_sel[u] = -1;
        P.u_v[u] = V_INDETERMINATE;
        P.u_ftol[u] = __builtin_nan("");
        P.u_dgbad[u] = 0;
        if (P.flags[c] & CF_REJECTED) { continue; }
        const int na = 2 * P.u_nwit[u];
        bool disappearance = false;
        for (int j = 0; j < na; ++j) { disappearance = disappearance || (P.a_st[j * n + u] & 0xFF) == A_DISAPPEARANCE; }
        P.u_dgbad[u] = disappearance ? 2 : 0;
        double z[NC];
        bool support = true;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            z[i] = P.feed[i * n + c];
            support = support && z[i] > 0;
        }
        bool any = false;
        for (int j = 0; j < na; ++j) { any = any || (P.a_st[j * n + u] & 0xFF) == A_CONVERGED; }
        if (!any) { continue; }
        if (!support) { atomicOr(&P.flags[c], CF_FALLBACK); continue; }
        const Mix<NC>& m = smix[mix_at(P, c)];
        const double pc = P.p[c], tc = P.t[c];
        Counters cnt{0, 0, 0, 0};
        bool rejected = false;
        SplitState<NC> s;
        int sel = -1;
        uint32_t tried = 0;
        for (int attempt = 0; attempt < 3 && sel < 0; ++attempt) {
            int best = -1;
            double best_g = 0;
            for (int j = 0; j < na; ++j) {
                if ((P.a_st[j * n + u] & 0xFF) != A_CONVERGED || ((tried >> j) & 1u)) { continue; }
                const double g = P.a_s[(j * 6 + 0) * n + u];
                if (best < 0 || g < best_g) { best = j; best_g = g; }
            }
            if (best < 0) { break; }
            tried |= 1u << best;
            double lk[NC];
            attempt_logk<NC>(P, best, u, lk);
            if (!polish_split<NC>(m, pc, tc, z, lk, P.a_s[(best * 6 + 1) * n + u], P.fug_tol, s, cnt, rejected)) {
                if (rejected) { break; }
                continue;
            }
            double contrast = 0;
#pragma unroll
            for (int i = 0; i < NC; ++i) { contrast = smax(contrast, __builtin_fabs(s.lk[i])); }
            if (!(s.norm <= P.fug_tol) || !(s.fr.mass_abs <= 1e-12 && s.fr.mass_rel <= 1e-10) || s.fr.beta <= 1e-10 ||
                1 - s.fr.beta <= 1e-10 || contrast <= 1e-7 || s.zv - s.zl <= 1e-8 * smax(s.zv, s.zl)) {
                continue;
            }
            sel = best;
        }
        atomicAdd(&P.cnt[0 * n + c], cnt.prop);
        atomicAdd(&P.cnt[2 * n + c], cnt.rr);
        if (rejected) { atomicOr(&P.flags[c], CF_REJECTED); continue; }
        // no candidate holds in FP64: the reference's outcome without a converged attempt
        if (sel < 0) { continue; }
        P.a_st[sel * n + u] = A_CONVERGED | (s.liq.branch << 8) | (s.vap.branch << 16);
        const double sv[6] = {s.gibbs, s.fr.beta, s.zl, s.zv, s.norm, s.guard};
#pragma unroll
        for (int q = 0; q < 6; ++q) { P.a_s[(sel * 6 + q) * n + u] = sv[q]; }
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            P.a_v[((sel * 5 + 0) * NC + i) * n + u] = s.fr.x[i];
            P.a_v[((sel * 5 + 1) * NC + i) * n + u] = s.fr.y[i];
            P.a_v[((sel * 5 + 2) * NC + i) * n + u] = s.liq.lnphi[i];
            P.a_v[((sel * 5 + 3) * NC + i) * n + u] = s.vap.lnphi[i];
            P.a_v[((sel * 5 + 4) * NC + i) * n + u] = s.common[i];
        }
        P.u_sel[u] = sel;
        double feed_g = 0, corr = 0, mag = 1;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            const double lz = log(z[i]);
            const double ph = P.refl[i * n + c];
            kadd(z[i] * (lz + ph), feed_g, corr);
            mag += z[i] * (__builtin_fabs(lz) + __builtin_fabs(ph));
        }
        const double dg = s.gibbs - feed_g;
        const double guard = s.guard + 256 * kEps * mag;
        const double ftol = 1e-10 + 0.5 * s.norm;
        if (!finite(dg) || !finite(guard) || !finite(ftol)) { continue; }
        P.u_v[u] = -1;
        P.u_ftol[u] = ftol;
        P.u_dgbad[u] |= dg > guard ? 1 : 0;
    }
}

} // namespace mpmc::dev
