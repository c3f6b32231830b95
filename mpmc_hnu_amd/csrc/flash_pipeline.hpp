// Staged batched flash for gfx950.
//
// The reference's per-cell search (solve_pr76_pt_max3, pr76_three_phase.hpp:328) is a
// nest of independent local searches: nc+2 TPD trials from fixed starts
// (pt_stability.hpp:442-466), then up to 16 split attempts seeded by the negative
// trials (pt_split.hpp:459-486), then nc+4 review trials against the common tangent
// (pt_split.hpp:508-512). Each of these only depends on the stage before it, so the
// batch runs stage by stage, one lane per (cell, trial) or (cell, attempt), instead of
// one lane per cell:
//
//   k_prep          per cell: normalised feed, reference phase (1 evaluation)
//   k_tpd(initial)  per (trial, cell): one TPD descent
//   k_init_reduce   per cell: stability status; unstable cells get a compact index u
//   k_split         per (attempt, u): one logK-SSI split attempt
//   k_split_select  per u: lowest-Gibbs converged attempt, review tolerance
//   k_tpd(review)   per (trial, u): one TPD descent against the imposed tangent
//   k_review_reduce per u: two-phase / unstable / indeterminate
//   k_max3_tail     per u still unstable: three-phase attempts (serial, rare)
//   k_publish       per cell: phase-set publication
//   k_fallback      per cell the stages cannot reproduce exactly: the serial path
//
// The TPD and split kernels are persistent: each lane runs a small state machine whose
// loop body performs exactly one EOS evaluation (or one backtrack with none), and a
// lane that finishes its search takes the next task from a device work counter (one
// atomic per wave), so lanes of a wave stay busy on the same code path. Stage sizes live
// in device memory; no host synchronisation is needed between stages.
//
// Evaluation budgets of the reference are shared across searches run in order
// (stability: 100000 per test, pt_stability.hpp:206; split: 20000 per flash,
// pt_split.hpp:41 and :472). Each search here runs with the full budget; the reduce
// kernels check the cumulative count in the reference's order and send every cell
// where a budget would have cut a search to k_fallback, which runs the serial path.
#pragma once

#include "closure_device.hpp"
#include "flash_kernels.hpp"
#include "mixed_device.hpp"

#include <algorithm>
#include <type_traits>

namespace mpmc::dev {

// Occupancy hint for the persistent search kernels (waves per SIMD the register
// allocation must allow); 0 leaves the choice to the compiler.
#ifndef MPMC_SEARCH_WAVES_PER_EU
#define MPMC_SEARCH_WAVES_PER_EU 0
#endif
#if MPMC_SEARCH_WAVES_PER_EU > 0
#define MPMC_SEARCH_ATTR __attribute__((amdgpu_waves_per_eu(MPMC_SEARCH_WAVES_PER_EU)))
#else
#define MPMC_SEARCH_ATTR
#endif
// The same hint for the FP64 stability (k_tpd) and split (k_split) kernels alone.
#ifndef MPMC_TPD_WAVES_PER_EU
#define MPMC_TPD_WAVES_PER_EU MPMC_SEARCH_WAVES_PER_EU
#endif
#ifndef MPMC_SPLIT_WAVES_PER_EU
#define MPMC_SPLIT_WAVES_PER_EU MPMC_SEARCH_WAVES_PER_EU
#endif
#if MPMC_TPD_WAVES_PER_EU > 0
#define MPMC_TPD_ATTR __attribute__((amdgpu_waves_per_eu(MPMC_TPD_WAVES_PER_EU)))
#else
#define MPMC_TPD_ATTR
#endif
#if MPMC_SPLIT_WAVES_PER_EU > 0
#define MPMC_SPLIT_ATTR __attribute__((amdgpu_waves_per_eu(MPMC_SPLIT_WAVES_PER_EU)))
#else
#define MPMC_SPLIT_ATTR
#endif

// Order in which a persistent queue hands out its (k, slot) tasks: 0 slot-fastest (the k-th
// trial or attempt of every cell, then the next k), 1 k-fastest (all trials or attempts of
// one cell together), 2 slot-fastest from the last k down. Results do not depend on it:
// every task writes its own record and budgets are replayed in the reference's order
// afterwards. The split queues take all attempts of a cell together: in attempt-major
// order the high-index attempts, which only cells with many negative trials have, all
// start at the end of the queue and lengthen its tail (measured on the 1M-cell nc = 6
// grid: FP64 flash 164 -> 155 ms, decision-exact mixed 133 -> 123 ms, identical records).
#ifndef MPMC_SPLIT_ORDER
#define MPMC_SPLIT_ORDER 1
#endif
#ifndef MPMC_TPD_ORDER
#define MPMC_TPD_ORDER 0
#endif
// 1: the split queues take the unstable cells in decreasing order of their witness count
// (u_order), the cells with the most attempts first, so that the long work starts early
// and short attempts fill the end (measured on the 1M-cell nc = 6 grid: FP64 flash
// 154 -> 148 ms, decision-exact mixed 122 -> 117 ms, identical records).
#ifndef MPMC_SPLIT_LPT
#define MPMC_SPLIT_LPT 1
#endif
template <int kOrder>
__device__ __forceinline__ void task_index(int64_t task, int64_t nslot, int per_slot, int& k, int64_t& slot) {
    if (kOrder == 1) {
        k = static_cast<int>(task % per_slot);
        slot = task / per_slot;
    } else if (kOrder == 2) {
        k = per_slot - 1 - static_cast<int>(task / nslot);
        slot = task % nslot;
    } else {
        k = static_cast<int>(task / nslot);
        slot = task % nslot;
    }
}

constexpr int kAttempts = 16;   // PtSplitOptions::max_split_attempts (pt_split.hpp:49)
constexpr int kMaxHints = 6;    // warm-start compositions per cell (extra initial stability starts)
constexpr int kWitnesses = 8;   // attempts per witness = 2 (vapor, liquid)

// Outcomes of a reduced-precision trial or split attempt that FP64 must decide: an
// ambiguous end, or a search handed to the FP64 kernels to be continued from its current
// point (iteration count in bits 16..).
constexpr int32_t T_AMBIG = 0x7D;
constexpr int32_t T_RESUME = 0x7C;
constexpr int32_t T_REDO = 0x7B;      // review trial to be searched again in FP64 from its start
constexpr int A_AMBIG = 0x7D;
constexpr int A_RESUME = 0x7C;

// CF_FALLBACK: the cell is finished by the serial FP64 path.
enum CellFlag : int32_t { CF_REJECTED = 1, CF_FALLBACK = 2 };
enum GlobalSlot : int { G_NU = 0, G_NFALLBACK = 1, G_TASK_INIT = 2, G_TASK_SPLIT = 3, G_TASK_REVIEW = 4,
                        G_TASK_MAX3 = 5, G_TASK_FALLBACK = 6, G_TASK_INIT32 = 7, G_TASK_REVIEW32 = 8,
                        G_TASK_SPLIT32 = 9, G_TASK_INIT_RESUME = 10, G_TASK_REVIEW_RESUME = 11,
                        G_TASK_SPLIT_RESUME = 12, G_TASK_REVIEW_REDO = 13, G_TASK_REVIEW_REDO2 = 14,
                        G_BUCKET = 16, G_CURSOR = 25, G_SLOTS = 34 };

// Device workspace carved from one allocation (pipe_carve). n = cells; arrays
// indexed by the compact unstable index u use capacity n.
template <int NC>
struct Pipe {
    int64_t n;
    int max_phases;
    double fug_tol;
    const double* p;
    const double* t;
    const double* zin;
    FlashOut out;
    double* feed;       // [NC][n]
    double* refl;       // [NC][n]
    int32_t* refinfo;   // [n] branch | smooth << 8 | ok << 9 | active << 16 | hints << 24 | final hints << 27
    int32_t* flags;     // [n]
    int32_t* cstat;     // [n] initial stability status
    int32_t* cell_u;    // [n]
    uint32_t* cnt;      // [4][n]
    int32_t* tst;       // [NC+2+kMaxHints][n] status | branch << 8 (-1: no such trial)
    uint32_t* tev;      // [NC+2+kMaxHints][n]
    double* tw;         // [NC+2+kMaxHints][NC][n] final point of negative trials
    int32_t* u_cell;    // [n]
    int32_t* u_order;   // [n] split queue order of the unstable cells (MPMC_SPLIT_LPT)
    int32_t* u_nwit;    // [n]
    int32_t* u_wit;     // [8][n]
    int32_t* a_st;      // [16][n] status | liquid branch << 8 | vapor branch << 16
    uint32_t* a_ev;     // [16][n]
    double* a_s;        // [16][6][n] gibbs, beta, zl, zv, norm, guard
    double* a_v;        // [16][5][NC][n] x, y, lnphi_l, lnphi_v, common
    int32_t* u_sel;     // [n]
    int32_t* u_v;       // [n] VLE status (-1 pending)
    double* u_ftol;     // [n]
    int32_t* u_dgbad;   // [n]
    int32_t* fst;       // [NC+4+kMaxHints][n]
    uint32_t* fev;      // [NC+4+kMaxHints][n]
    double* fw;         // [NC+4+kMaxHints][NC][n]
    int32_t* m_st;      // [n] max3 status | neighbour << 8
    double* m_s;        // [6][n] beta[3], z[3]
    double* m_v;        // [6][NC][n] x[3], lnphi[3]
    int32_t* m_br;      // [3][n]
    unsigned long long* g; // [16]
    int32_t* fb_list;   // [n]
    // per-cell constants of the reduced-precision searches (mixed precision only)
    float* ra32;        // [NC][n] sqrt(a_i(T))
    float* lz32;        // [NC][n] ln z_i
    float* d32;         // [NC][n] ln z_i + ln phi_i(reference phase)
    // Ragged multi-mixture batches: cell c belongs to mixture mix_of[c] of the launch's
    // table of nmix mixtures (all with NC components); mix_of == nullptr: one mixture.
    const int32_t* mix_of;
    int nmix;
    int mixed;          // mixed-precision schedules: 1, then 2 after the review redo
    int exact;          // precision 1: every decision-bearing search is (re)run in FP64 from its start
    // warm-start hints: nhint compositions per cell, hints[(h*NC + i)*n + c], appended to the
    // initial stability test's starts in order (NaN in component 0: no hint h for cell c)
    const double* hints;
    int nhint;
    // final-review hints, the same layout: extra starts of the common-tangent review, placed
    // before the selected pair's two compositions (pt_split.hpp:508-510)
    const double* fhints;
    int nfhint;
    // two caller-supplied three-phase start slots per cell (slot 0: a continuation's carried
    // state, slot 1: a fixed start; tried in that order): compositions
    // tpz[((s*3 + ph)*NC + i)*n + c] and phase-fraction seed tpf[(s*2 + k)*n + c] of phases
    // 1, 2 (NaN in tpf[(s*2)*n + c]: slot s empty for cell c); nullptr: no slots
    const double* tpz;
    const double* tpf;
};

// A cell's supplied three-phase start slots, read from Pipe::tpz / tpf into registers.
template <int NC>
struct SuppliedStarts {
    bool has0 = false, has1 = false;
    double st0[3 * NC], st1[3 * NC];
    double sf0[2], sf1[2];
    __device__ void load(const Pipe<NC>& P, int64_t c) {
        if (!P.tpz) { return; }
        const int64_t n = P.n;
        has0 = !__builtin_isnan(P.tpf[c]);
        has1 = !__builtin_isnan(P.tpf[2 * n + c]);
        sf0[0] = P.tpf[c];
        sf0[1] = P.tpf[n + c];
        sf1[0] = P.tpf[2 * n + c];
        sf1[1] = P.tpf[3 * n + c];
#pragma unroll
        for (int k = 0; k < 3 * NC; ++k) {
            st0[k] = P.tpz[k * n + c];
            st1[k] = P.tpz[(3 * NC + k) * n + c];
        }
    }
    __device__ const double* start(int slot) const {
        return slot == 0 ? (has0 ? st0 : nullptr) : (has1 ? st1 : nullptr);
    }
};

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ int64_t fetch_task(unsigned long long* counter, bool want) {
    const unsigned long long mask = __ballot(want);
    if (mask == 0ull) { return -1; }
    const int leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
    long long base = 0;
    if (static_cast<int>(__lane_id()) == leader) {
        base = static_cast<long long>(atomicAdd(counter, static_cast<unsigned long long>(__popcll(mask))));
    }
    base = __shfl(base, leader);
    const int rank = __popcll(mask & ((1ull << __lane_id()) - 1ull));
    return want ? static_cast<int64_t>(base) + rank : -1;
}

// The launch's mixture table (nmix consecutive Mix<NC> records) staged in LDS; the
// launch passes nmix * sizeof(Mix<NC>) bytes of dynamic LDS.
template <int NC>
__device__ __forceinline__ Mix<NC>* stage_mixes(const double* __restrict__ mixc, int nmix) {
    extern __shared__ double lds_mix[];
    constexpr int kd = static_cast<int>(sizeof(Mix<NC>) / sizeof(double));
    const int total = (nmix > 1 ? nmix : 1) * kd;
    for (int k = threadIdx.x; k < total; k += blockDim.x) { lds_mix[k] = mixc[k]; }
    __syncthreads();
    return reinterpret_cast<Mix<NC>*>(lds_mix);
}

template <int NC>
__device__ __forceinline__ int mix_at(const Pipe<NC>& P, int64_t c) {
    return P.mix_of ? P.mix_of[c] : 0;
}

template <int NC>
__device__ __forceinline__ int active_count(int32_t info) { return (info >> 16) & 0xFF; }

template <int NC>
__device__ __forceinline__ int hint_count(int32_t info) { return (info >> 24) & 0x7; }

template <int NC>
__device__ __forceinline__ int final_hint_count(int32_t info) { return (info >> 27) & 0x7; }

// The m-th hint present for cell c in an [nh][NC][n] hint array (hints absent for a cell,
// NaN in component 0, are skipped).
template <int NC>
__device__ __forceinline__ void nth_hint(const double* H, int nh, int64_t n, int64_t c, int m, double* w) {
    int seen = -1;
    for (int h = 0; h < nh; ++h) {
        if (__builtin_isnan(H[(h * NC + 0) * n + c])) { continue; }
        if (++seen == m) {
#pragma unroll
            for (int i = 0; i < NC; ++i) { w[i] = H[(h * NC + i) * n + c]; }
            return;
        }
    }
}

template <int NC>
__device__ __forceinline__ void hint_start(const Pipe<NC>& P, int64_t c, int m, double* w) {
    nth_hint<NC>(P.hints, P.nhint, P.n, c, m, w);
}

// Review trial count of the unstable cell with info word `info`: the generated starts, the
// final-review hints, the selected pair's liquid and vapor compositions.
template <int NC>
__device__ __forceinline__ int review_trials(int32_t info) {
    const int active = active_count<NC>(info);
    return (active > 1 ? active + 2 : 1) + final_hint_count<NC>(info) + 2;
}

// Start composition of trial k (pt_stability.hpp:442-466): feed, uniform over the active
// support, 0.1 feed + 0.9 e_i per active i; then the extra starts.
template <int NC>
__device__ __forceinline__ void trial_start(const double* z, int active, int k, double* w) {
    if (k == 0) {
#pragma unroll
        for (int i = 0; i < NC; ++i) { w[i] = z[i]; }
    } else if (k == 1) {
#pragma unroll
        for (int i = 0; i < NC; ++i) { w[i] = z[i] > 0 ? 1.0 / static_cast<double>(active) : 0.0; }
    } else {
        int seen = 0, vert = -1;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            if (z[i] != 0) { if (seen == k - 2) { vert = i; } ++seen; }
        }
#pragma unroll
        for (int i = 0; i < NC; ++i) { w[i] = z[i] * 0.1 + (i == vert ? 0.9 : 0.0); }
    }
}

// ------------------------------------------------------------------ k_reset / k_prep
static __global__ void k_reset(unsigned long long* g) {
    if (threadIdx.x < G_SLOTS) { g[threadIdx.x] = 0; }
}

template <int NC, bool kN>
__global__ __launch_bounds__(64) void k_prep(Pipe<NC> P, const double* __restrict__ mixc) {
    Mix<NC>* const smix = stage_mixes<NC>(mixc, P.nmix);
    const int64_t n = P.n;
    for (int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < n;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        double z[NC];
#pragma unroll
        for (int i = 0; i < NC; ++i) { z[i] = P.zin[i * n + c]; }
        const double p = P.p[c], t = P.t[c];
        bool rejected = false;
        // solve_pr76_pt_max3 / solve_pt_vle / test_pt_stability_impl input contracts
        const double isum = check_composition<NC>(z, rejected);
        if (!finite(p) || p <= 0 || !finite(t) || t <= 0) { rejected = true; }
        int active = 0;
        bool lost = false;
        double feed[NC];
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            if (z[i] > 0) { ++active; }
            feed[i] = z[i] / isum;
            lost = lost || (z[i] > 0 && feed[i] == 0);
        }
        uint32_t prop = 0;
        int32_t info = active << 16;
        // warm-start hints: validated before any provider call, as the reference validates
        // every supplied start (pt_stability.hpp:264-271; final starts pt_split.hpp:440-446)
        int nh = 0, nfh = 0;
        for (int h = 0; h < P.nhint + P.nfhint; ++h) {
            const bool fin = h >= P.nhint;
            const double* H = fin ? P.fhints : P.hints;
            const int hh = fin ? h - P.nhint : h;
            if (__builtin_isnan(H[(hh * NC + 0) * n + c])) { continue; }
            double w[NC];
#pragma unroll
            for (int i = 0; i < NC; ++i) { w[i] = H[(hh * NC + i) * n + c]; }
            (void)check_composition<NC>(w, rejected);
#pragma unroll
            for (int i = 0; i < NC; ++i) { if ((z[i] > 0) != (w[i] > 0)) { rejected = true; } }
            if (fin) { ++nfh; } else { ++nh; }
        }
        info |= (nh << 24) | (nfh << 27);
        if (P.tpz) {
#pragma unroll
            for (int sl = 0; sl < 2; ++sl) {
                if (__builtin_isnan(P.tpf[(sl * 2) * n + c])) { continue; }
                double st3[3 * NC];
                const double sf[2] = {P.tpf[(sl * 2) * n + c], P.tpf[(sl * 2 + 1) * n + c]};
#pragma unroll
                for (int k = 0; k < 3 * NC; ++k) { st3[k] = P.tpz[(sl * 3 * NC + k) * n + c]; }
                check_three_phase_start<NC>(z, st3, sf, rejected);
            }
        }
        Activity<NC> ref;
        if (!rejected && !lost) {
            ++prop;
            const int e = stability_provider<NC, kN>(smix[mix_at(P, c)], p, t, feed, ref);
            if (e == REJECTED) { rejected = true; }
            else if (e == ISSUE_NONE) { info |= ref.branch | (ref.smooth ? 0x100 : 0) | 0x200; }
        }
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            P.feed[i * n + c] = feed[i];
            P.refl[i * n + c] = (info & 0x200) ? ref.lnphi[i] : 0.0;
        }
        P.refinfo[c] = info;
        P.flags[c] = rejected ? CF_REJECTED : 0;
        if (P.mixed) {
            const Mix<NC>& m = smix[mix_at(P, c)];
            const double st = __builtin_sqrt(t);
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                const double br = 1.0 + m.kappa[i] * (1.0 - st / m.sqrt_tc[i]);
                const double a = (br * m.ac[i]) * br;
                const double lz = feed[i] > 0 ? log(feed[i]) : 0.0;
                P.ra32[i * n + c] = static_cast<float>(__builtin_sqrt(a > 0 ? a : 0.0));
                P.lz32[i * n + c] = static_cast<float>(lz);
                P.d32[i * n + c] = static_cast<float>(lz + ((info & 0x200) ? ref.lnphi[i] : 0.0));
            }
        }
        P.cstat[c] = S_INDETERMINATE;
        P.cell_u[c] = -1;
        P.cnt[0 * n + c] = prop;
        P.cnt[1 * n + c] = 0;
        P.cnt[2 * n + c] = 0;
        P.cnt[3 * n + c] = 0;
    }
}

// ------------------------------------------------------------------ TPD lanes
enum TpdMode : int { TM_START = 1, TM_BT = 2 };

template <int NC>
struct TpdLane {
    int64_t slot;   // c (initial) or u (review)
    int64_t cell;
    int mid;        // the cell's mixture in the launch's table
    int k, mode, bt, it;
    uint32_t evals;
    double p, t, tol;
    double z[NC];
    Activity<NC> ref;
    TpdPoint<NC> pt;
    double dir[NC], lw[NC], cand[NC];
    double slope, scale, alpha;
    uint32_t evals0;   // evaluations counted before an FP64 continuation
};

// stage 0: initial feed test; stage 1: common-tangent review of the selected pair.
// mode 1: FP64 continuation of the reduced-precision descents marked T_RESUME, from their
// stored point and iteration count; mode 2: the trials marked T_REDO, from their start
// (mixed precision).
template <int NC>
__device__ __forceinline__ bool tpd_begin(const Pipe<NC>& P, int stage, int k, int64_t slot, int mode, TpdLane<NC>& L) {
    const int64_t n = P.n;
    int64_t c;
    int32_t* st_arr = stage == 0 ? P.tst : P.fst;
    if (mode == 1 && (st_arr[k * n + slot] & 0xFF) != T_RESUME) { return false; }
    if (mode == 2 && (st_arr[k * n + slot] & 0xFF) != T_REDO) { return false; }
    const bool resume = mode == 1;
    if (stage == 0) {
        c = slot;
        if (P.flags[c] & CF_REJECTED) { return false; }
        const int32_t info = P.refinfo[c];
        const int active = active_count<NC>(info);
        const int ntr = (active > 1 ? active + 2 : 1) + hint_count<NC>(info);
        if (!(info & 0x200) || k >= ntr) { st_arr[k * n + slot] = -1; return false; }
    } else {
        const int sel = P.u_sel[slot];
        if (sel < 0 || !finite(P.u_ftol[slot])) { st_arr[k * n + slot] = -1; return false; }
        c = P.u_cell[slot];
        if (P.flags[c] & CF_REJECTED) { st_arr[k * n + slot] = -1; return false; }
        if (k >= review_trials<NC>(P.refinfo[c])) { st_arr[k * n + slot] = -1; return false; }
    }
    L.slot = slot;
    L.cell = c;
    L.mid = mix_at(P, c);
    L.k = k;
    L.p = P.p[c];
    L.t = P.t[c];
#pragma unroll
    for (int i = 0; i < NC; ++i) { L.z[i] = P.feed[i * n + c]; }
    const int active = active_count<NC>(P.refinfo[c]);
    const int gen = active > 1 ? active + 2 : 1;
    if (stage == 0) {
        L.tol = 1e-10;
#pragma unroll
        for (int i = 0; i < NC; ++i) { L.ref.lnphi[i] = P.refl[i * n + c]; }
        L.ref.branch = P.refinfo[c] & 0xFF;
        L.ref.smooth = (P.refinfo[c] >> 8) & 1;
    } else {
        const int sel = P.u_sel[slot];
        L.tol = P.u_ftol[slot];
        bool fin = true;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            const double common = P.a_v[((sel * 5 + 4) * NC + i) * n + slot];
            L.ref.lnphi[i] = L.z[i] > 0 ? common - log(L.z[i]) : 0.0;
            fin = fin && finite(L.ref.lnphi[i]);
        }
        L.ref.branch = 0;
        L.ref.smooth = true;
        if (!fin) { st_arr[k * n + slot] = T_PROPERTY_FAILURE; P.fev[k * n + slot] = 0; return false; }
    }
    if (resume) {
        const double* w = stage == 0 ? P.tw : P.fw;
#pragma unroll
        for (int i = 0; i < NC; ++i) { L.cand[i] = w[(k * NC + i) * n + slot]; }
        L.it = st_arr[k * n + slot] >> 16;
        L.evals = (stage == 0 ? P.tev : P.fev)[k * n + slot];
        L.evals0 = L.evals;
        L.mode = TM_START;
        return true;
    }
    if (k < gen) {
        trial_start<NC>(L.z, active, k, L.cand);
    } else if (stage == 0) {
        hint_start<NC>(P, c, k - gen, L.cand);
    } else if (k - gen < final_hint_count<NC>(P.refinfo[c])) {
        nth_hint<NC>(P.fhints, P.nfhint, n, c, k - gen, L.cand);
    } else {
        const int sel = P.u_sel[slot];
        // 0: liquid x, 1: vapor y of the selected pair, after the final-review hints
        const int q = k - gen - final_hint_count<NC>(P.refinfo[c]);
#pragma unroll
        for (int i = 0; i < NC; ++i) { L.cand[i] = P.a_v[((sel * 5 + q) * NC + i) * n + slot]; }
    }
    const double s = ksum<NC>(L.cand);
    bool lost = false;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        L.cand[i] /= s;
        lost = lost || (L.z[i] > 0 && L.cand[i] == 0);
    }
    L.evals = 0;
    L.evals0 = 0;
    L.it = 0;
    L.mode = TM_START;
    if (lost) {
        st_arr[k * n + slot] = T_UNREPRESENTABLE;
        (stage == 0 ? P.tev : P.fev)[k * n + slot] = 0;
        return false;
    }
    return true;
}

template <int NC>
__device__ __forceinline__ void tpd_finish(const Pipe<NC>& P, int stage, TpdLane<NC>& L, int status) {
    const int64_t n = P.n;
    const int64_t idx = L.k * n + L.slot;
    (stage == 0 ? P.tst : P.fst)[idx] = status | (L.pt.branch << 8);
    (stage == 0 ? P.tev : P.fev)[idx] = L.evals;
    if (status == T_NEGATIVE) {
        double* w = stage == 0 ? P.tw : P.fw;
#pragma unroll
        for (int i = 0; i < NC; ++i) { w[(L.k * NC + i) * n + L.slot] = L.pt.w[i]; }
    }
    atomicAdd(&P.cnt[0 * n + L.cell], L.evals - L.evals0);
    atomicAdd(&P.cnt[1 * n + L.cell], L.evals - L.evals0);
}

// Termination tests and the next log-space descent direction (pt_stability.hpp:528-561).
// Returns a trial status when the search ends, or -1 to start backtracking.
template <int NC>
__device__ __forceinline__ int tpd_check(TpdLane<NC>& L) {
    if (tpd_negative(L.pt.value, L.pt.guard, L.tol)) { return T_NEGATIVE; }
    if (!L.pt.smooth) { return T_NONSMOOTH; }
    if (L.pt.stat <= 1e-8) { return T_STATIONARY; }
    if (L.it >= 512) { return T_ITERATION_LIMIT; }
    L.scale = smax(1.0, L.pt.stat / 4.0);
    double slope = 0, sc = 0;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        L.dir[i] = 0;
        L.lw[i] = 0;
        if (L.z[i] == 0) { continue; }
        L.dir[i] = -L.pt.res[i] / L.scale;
        kadd((L.pt.w[i] * L.pt.res[i]) * L.dir[i], slope, sc);
        L.lw[i] = log(L.pt.w[i]);
    }
    if (!finite(slope) || !(slope < 0)) { return T_LINE_SEARCH_FAILED; }
    L.slope = slope;
    L.mode = TM_BT;
    L.bt = 0;
    L.alpha = 1;
    return -1;
}

// One step of an active TPD lane: one evaluation, or one rejected backtrack without
// evaluation. Returns false when the lane's trial has ended (recorded by tpd_finish).
template <int NC, bool kN>
__device__ __forceinline__ bool tpd_step(const Pipe<NC>& P, const Mix<NC>* smix, int stage, TpdLane<NC>& L) {
    // one evaluation (or one rejected backtrack without evaluation)
    bool do_eval = true;
    if (L.mode == TM_BT) {
        double largest = -kInf;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            if (L.z[i] > 0) {
                L.cand[i] = L.lw[i] + L.alpha * L.dir[i];
                largest = smax(largest, L.cand[i]);
            }
        }
#pragma unroll
        for (int i = 0; i < NC; ++i) { L.cand[i] = L.z[i] > 0 ? exp(L.cand[i] - largest) : 0.0; }
        const double s = ksum<NC>(L.cand);
        bool positive = finite(s) && s > 0;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            L.cand[i] /= s;
            positive = positive && (L.z[i] == 0 || L.cand[i] > 0);
        }
        do_eval = positive;
    }
    TpdPoint<NC> nx;
    int e = ISSUE_NONE;
    bool rejected = false;
    if (do_eval) {
        ++L.evals;
        Activity<NC> ph;
        e = stability_provider<NC, kN>(smix[L.mid], L.p, L.t, L.cand, ph);
        if (e == ISSUE_NONE) { e = tpd_point<NC>(L.cand, L.z, ph, L.ref, nx, rejected); }
        if (e == REJECTED || rejected) {
            atomicOr(&P.flags[L.cell], CF_REJECTED);
            tpd_finish<NC>(P, stage, L, T_PROPERTY_FAILURE);
            return false;
        }
    }
    int done = -1;
    if (L.mode == TM_START) {
        if (e != ISSUE_NONE) { done = T_PROPERTY_FAILURE; }
        else { L.pt = nx; done = tpd_check<NC>(L); }
    } else {
        bool ok = false;
        if (do_eval && e == ISSUE_NONE) {
            ok = tpd_negative(nx.value, nx.guard, L.tol);
            if (!ok) {
                // stability_accept_step (pt_stability.hpp:312-326)
                const double pred = -L.alpha * L.slope;
                const double gd = L.pt.guard + nx.guard;
                if (pred <= gd) { ok = nx.value <= L.pt.value + gd && L.alpha / L.scale > 0 && nx.stat < L.pt.stat; }
                else { ok = nx.value <= L.pt.value - 1e-4 * pred; }
            }
        }
        if (ok) {
            L.pt = nx;
            ++L.it;
            done = tpd_check<NC>(L);
        } else {
            ++L.bt;
            L.alpha *= 0.5;
            if (L.bt >= 32) { done = T_LINE_SEARCH_FAILED; }
        }
    }
    if (done >= 0) {
        tpd_finish<NC>(P, stage, L, done);
        return false;
    }
    return true;
}

template <int NC, bool kN>
__global__ __launch_bounds__(64) MPMC_TPD_ATTR void k_tpd(Pipe<NC> P, const double* __restrict__ mixc, int stage,
                                                              int resume) {
    Mix<NC>* const smix = stage_mixes<NC>(mixc, P.nmix);
    const int tk = stage == 0 ? NC + 2 + P.nhint : NC + 4 + P.nfhint;
    const int64_t nslot = stage == 0 ? P.n : static_cast<int64_t>(P.g[G_NU]);
    const int64_t total = tk * nslot;
    // resume 0: every trial; 1: FP64 continuation of T_RESUME trials; 2 and 3: FP64 redo of
    // T_REDO trials from their start (3: a second redo pass of the review, own counter)
    unsigned long long* counter =
        &P.g[stage == 0 ? (resume ? G_TASK_INIT_RESUME : G_TASK_INIT)
                        : (resume == 1 ? G_TASK_REVIEW_RESUME
                                       : (resume == 2 ? G_TASK_REVIEW_REDO : (resume == 3 ? G_TASK_REVIEW_REDO2 : G_TASK_REVIEW)))];
    const int mode = resume == 3 ? 2 : resume;
    TpdLane<NC> L;
    bool active = false, exhausted = false;
    for (;;) {
        const bool want = !active && !exhausted;
        const int64_t task = fetch_task(counter, want);
        if (want) {
            if (task >= total) { exhausted = true; }
            else {
                int k;
                int64_t slot;
                task_index<MPMC_TPD_ORDER>(task, nslot, static_cast<int>(total / nslot), k, slot);
                active = tpd_begin<NC>(P, stage, k, slot, mode, L);
            }
        }
        if (!__any(active)) {
            if (!__any(!exhausted)) { break; }
            continue;
        }
        if (!active) { continue; }
        active = tpd_step<NC, kN>(P, smix, stage, L);
    }
}

// ------------------------------------------------------------------ stability reduce
// The initial stability result of cell c from its trials (status, budget, fallback).
// Returns the cell's witness count (at most kWitnesses, the trials in wit) when it is
// unstable and continues on the staged path, else -1.
template <int NC>
__device__ __forceinline__ int init_reduce_cell(const Pipe<NC>& P, int64_t c, int (&wit)[kWitnesses]) {
    const int64_t n = P.n;
    if (P.flags[c] & CF_REJECTED) { return -1; }
    const int32_t info = P.refinfo[c];
    if (!(info & 0x200)) { P.cstat[c] = S_INDETERMINATE; return -1; }
    const int active = active_count<NC>(info);
    const int ntr = (active > 1 ? active + 2 : 1) + hint_count<NC>(info);
    uint64_t evals = 1;
    bool neg = false, all_stat = true, ambiguous = false;
    int nwit = 0;
    for (int k = 0; k < ntr; ++k) {
        const int st = P.tst[k * n + c] & 0xFF;
        ambiguous = ambiguous || st == T_AMBIG;
        evals += P.tev[k * n + c];
        if (st == T_NEGATIVE) {
            neg = true;
            if (nwit < kWitnesses) { wit[nwit] = k; }
            ++nwit;
        }
        all_stat = all_stat && st == T_STATIONARY;
    }
    if (evals > 100000 || ambiguous) { atomicOr(&P.flags[c], CF_FALLBACK); }
    const int status = neg ? S_UNSTABLE : (all_stat ? S_NO_INSTABILITY : S_INDETERMINATE);
    P.cstat[c] = status;
    if (status == S_UNSTABLE && !(P.flags[c] & CF_FALLBACK)) { return nwit < kWitnesses ? nwit : kWitnesses; }
    return -1;
}

// Wave-aggregated count by key: for every lane with key >= 0, atomicAdd(&counts[key], 1)
// with one atomic per distinct key in the wave. Returns the lane's previous count (the
// count before the wave plus the lane's rank among the wave's lanes of that key), or 0.
__device__ __forceinline__ unsigned long long count_keyed(unsigned long long* counts, int key) {
    const int lane = static_cast<int>(__lane_id());
    unsigned long long pending = __ballot(key >= 0);
    unsigned long long mine = 0;
    while (pending != 0ull) {
        const int leader = __ffsll(pending) - 1;
        const int k = __shfl(key, leader);
        const unsigned long long same = __ballot(key == k) & pending;
        unsigned long long base = 0;
        if (lane == leader) { base = atomicAdd(&counts[k], static_cast<unsigned long long>(__popcll(same))); }
        base = __shfl(base, leader);
        if ((same >> lane) & 1ull) { mine = base + static_cast<unsigned long long>(__popcll(same & ((1ull << lane) - 1ull))); }
        pending &= ~same;
    }
    return mine;
}

// The compact record of unstable cell c under index u: cell, witness count, witnesses.
template <int NC>
__device__ __forceinline__ void record_u(const Pipe<NC>& P, int64_t c, int64_t u, int nw, const int (&wit)[kWitnesses]) {
    const int64_t n = P.n;
    P.cell_u[c] = static_cast<int32_t>(u);
    P.u_cell[u] = static_cast<int32_t>(c);
    P.u_nwit[u] = nw;
    for (int j = 0; j < nw; ++j) { P.u_wit[j * n + u] = wit[j]; }
}

// Per cell: the stability result, and for the unstable cells that continue on the staged
// path a compact index u (one atomic per wave) with their witnesses; with MPMC_SPLIT_LPT the
// cells per witness count.
template <int NC>
__global__ __launch_bounds__(64) void k_init_reduce(Pipe<NC> P) {
    const int64_t n = P.n;
    for (int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < n;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        int wit[kWitnesses];
        const int nw = init_reduce_cell<NC>(P, c, wit);
        const int64_t u = fetch_task(&P.g[G_NU], nw >= 0);
        if (nw >= 0) { record_u<NC>(P, c, u, nw, wit); }
        if (MPMC_SPLIT_LPT) { (void)count_keyed(&P.g[G_BUCKET], nw); }
    }
}

// The split queue order of the unstable cells: by decreasing witness count (counting sort
// over the counts k_init_reduce gathered; the order within a count is arbitrary).
template <int NC>
__global__ __launch_bounds__(64) void k_order_u(Pipe<NC> P) {
    const int64_t nu = static_cast<int64_t>(P.g[G_NU]);
    for (int64_t u = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; u < nu;
         u += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int nw = P.u_nwit[u];
        unsigned long long base = 0;
        for (int b = nw + 1; b <= kWitnesses; ++b) { base += P.g[G_BUCKET + b]; }
        P.u_order[base + count_keyed(&P.g[G_CURSOR], nw)] = static_cast<int32_t>(u);
    }
}

// Decision-exact mixed precision: the trials an FP32 descent ended clearly negative are
// searched again in FP64 from their start in every cell that has one, so that the cell's
// witnesses (the split seeds) are the FP64 schedule's end points bit for bit.
template <int NC>
__global__ __launch_bounds__(64) void k_witness_redo(Pipe<NC> P) {
    const int64_t n = P.n;
    for (int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < n;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        if (P.flags[c] & CF_REJECTED) { continue; }
        const int32_t info = P.refinfo[c];
        if (!(info & 0x200)) { continue; }
        const int active = active_count<NC>(info);
        const int ntr = (active > 1 ? active + 2 : 1) + hint_count<NC>(info);
        for (int k = 0; k < ntr; ++k) {
            if ((P.tst[k * n + c] & 0xFF) == T_NEGATIVE) { P.tst[k * n + c] = T_REDO; }
        }
    }
}

// ------------------------------------------------------------------ split lanes
template <int NC>
struct SplitLane {
    int64_t u, cell;
    int j, mode, bt, it;
    uint32_t evals;
    double p, t;
    double z[NC], lk0[NC];
    SplitState<NC> pt;
    double scale, slope, alpha;
    Counters cnt;
};

// resume: FP64 continuation of the reduced-precision attempts marked A_RESUME, from their
// stored logK and iteration count (mixed precision).
template <int NC>
__device__ __forceinline__ bool split_begin(const Pipe<NC>& P, int j, int64_t u, bool resume, SplitLane<NC>& L) {
    const int64_t n = P.n;
    if (resume) {
        if (j >= 2 * P.u_nwit[u] || (P.a_st[j * n + u] & 0xFF) != A_RESUME) { return false; }
    } else if (j >= 2 * P.u_nwit[u]) {
        P.a_st[j * n + u] = -1;
        P.a_ev[j * n + u] = 0;
        return false;
    }
    const int64_t c = P.u_cell[u];
    L.u = u;
    L.cell = c;
    L.j = j;
    L.p = P.p[c];
    L.t = P.t[c];
#pragma unroll
    for (int i = 0; i < NC; ++i) { L.z[i] = P.feed[i * n + c]; }
    if (resume) {
#pragma unroll
        for (int i = 0; i < NC; ++i) { L.lk0[i] = P.a_v[((j * 5 + 4) * NC + i) * n + u]; }
        L.evals = 0;
        L.it = P.a_st[j * n + u] >> 16;
        L.mode = TM_START;
        L.cnt = Counters{0, 0, 0, 0};
        return true;
    }
    const int k = P.u_wit[(j / 2) * n + u];
    double w[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) { w[i] = P.tw[(k * NC + i) * n + c]; }
    L.evals = 0;
    L.it = 0;
    L.mode = TM_START;
    L.cnt = Counters{0, 0, 0, 0};
    if (!split_seed<NC>(L.z, w, (j % 2) == 0, L.lk0)) {
        P.a_st[j * n + u] = A_UNREPRESENTABLE_SEED;
        P.a_ev[j * n + u] = 0;
        return false;
    }
    return true;
}

template <int NC>
__device__ __forceinline__ void split_finish(const Pipe<NC>& P, SplitLane<NC>& L, int status) {
    const int64_t n = P.n;
    const int64_t idx = L.j * n + L.u;
    P.a_ev[idx] = L.evals;
    if (status == A_CONVERGED) {
        P.a_st[idx] = status | (L.pt.liq.branch << 8) | (L.pt.vap.branch << 16);
        const double s[6] = {L.pt.gibbs, L.pt.fr.beta, L.pt.zl, L.pt.zv, L.pt.norm, L.pt.guard};
#pragma unroll
        for (int q = 0; q < 6; ++q) { P.a_s[(L.j * 6 + q) * n + L.u] = s[q]; }
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            P.a_v[((L.j * 5 + 0) * NC + i) * n + L.u] = L.pt.fr.x[i];
            P.a_v[((L.j * 5 + 1) * NC + i) * n + L.u] = L.pt.fr.y[i];
            P.a_v[((L.j * 5 + 2) * NC + i) * n + L.u] = L.pt.liq.lnphi[i];
            P.a_v[((L.j * 5 + 3) * NC + i) * n + L.u] = L.pt.vap.lnphi[i];
            P.a_v[((L.j * 5 + 4) * NC + i) * n + L.u] = L.pt.common[i];
        }
    } else {
        P.a_st[idx] = status;
    }
    atomicAdd(&P.cnt[0 * n + L.cell], L.evals);
    atomicAdd(&P.cnt[2 * n + L.cell], L.cnt.rr);
}

// Convergence tests and the Gibbs descent slope of iterate_pt_split (pt_split.hpp:349-373).
template <int NC>
__device__ __forceinline__ int split_check(SplitLane<NC>& L, double fug_tol) {
    const SplitState<NC>& pt = L.pt;
    if (pt.norm <= fug_tol) {
        double contrast = 0;
#pragma unroll
        for (int i = 0; i < NC; ++i) { contrast = smax(contrast, __builtin_fabs(pt.lk[i])); }
        if (!(pt.fr.mass_abs <= 1e-12 && pt.fr.mass_rel <= 1e-10)) { return A_BALANCE_FAILURE; }
        if (pt.fr.beta <= 1e-10 || 1 - pt.fr.beta <= 1e-10) { return A_DISAPPEARANCE; }
        if (contrast <= 1e-7 || pt.zv - pt.zl <= 1e-8 * smax(pt.zv, pt.zl)) { return A_INDISTINGUISHABLE; }
        return A_CONVERGED;
    }
    if (L.it >= 512) { return A_ITERATION_LIMIT; }
    L.scale = smax(1.0, pt.norm / 2.0);
    double h = 0, hc = 0, f = 0, fc = 0, s2 = 0, s2c = 0;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        if (L.z[i] == 0) { continue; }
        const double x = pt.fr.x[i], y = pt.fr.y[i];
        const double c = (x / L.z[i]) * y;
        const double d = y - x;
        const double r = pt.r[i];
        kadd(d * (d / L.z[i]), h, hc);
        kadd(c * r, f, fc);
        kadd((c * r) * r, s2, s2c);
    }
    double slope;
    if (!(h > 0) || !finite(h) || !finite(f) || !finite(s2) || !finite(L.scale) || !(L.scale >= 1)) {
        slope = __builtin_nan("");
    } else {
        slope = -(pt.fr.beta * (1 - pt.fr.beta) * s2 + (f / h) * f) / L.scale;
    }
    if (!finite(slope) || !(slope < 0)) { return A_LINE_SEARCH_FAILED; }
    L.slope = slope;
    L.mode = TM_BT;
    L.bt = 0;
    L.alpha = 1;
    return -1;
}

// One step of an active split lane: one evaluation of a trial point. Returns false when
// the lane's attempt has ended (recorded by split_finish).
template <int NC, bool kN>
__device__ __forceinline__ bool split_step(const Pipe<NC>& P, const Mix<NC>* smix, SplitLane<NC>& L) {
    double lk[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) { lk[i] = L.mode == TM_START ? L.lk0[i] : L.pt.lk[i] + (L.alpha / L.scale) * L.pt.r[i]; }
    SplitState<NC> nx;
    bool rejected = false;
    const int e = split_evaluate<NC, kN>(smix[mix_at(P, L.cell)], L.p, L.t, L.z, lk, L.evals, 20000u, nx, L.cnt, rejected,
                                     L.mode == TM_START ? 0.5 : L.pt.fr.beta);
    if (rejected) {
        atomicOr(&P.flags[L.cell], CF_REJECTED);
        split_finish<NC>(P, L, A_PROPERTY_FAILURE);
        return false;
    }
    int done = -1;
    if (L.mode == TM_START) {
        if (e != -1) { done = e; }
        else { L.pt = nx; done = split_check<NC>(L, P.fug_tol); }
    } else if (e == -1) {
        const double pred = -L.alpha * L.slope;
        const double gd = L.pt.guard + nx.guard;
        bool ok;
        // split_accept_step (pt_split.hpp:232-244)
        if (pred > gd) { ok = nx.gibbs <= L.pt.gibbs - 1e-4 * pred; }
        else {
            ok = nx.gibbs <= L.pt.gibbs + gd &&
                 (nx.norm <= P.fug_tol || L.pt.norm - nx.norm > 1e-4 * (L.alpha / L.scale) * L.pt.norm);
        }
        if (ok) {
            L.pt = nx;
            ++L.it;
            done = split_check<NC>(L, P.fug_tol);
        } else {
            ++L.bt;
            L.alpha *= 0.5;
            if (L.bt >= 24) { done = A_LINE_SEARCH_FAILED; }
        }
    } else if (e == A_EVALUATION_LIMIT) {
        done = A_EVALUATION_LIMIT;
    } else {
        ++L.bt;
        L.alpha *= 0.5;
        if (L.bt >= 24) { done = A_LINE_SEARCH_FAILED; }
    }
    if (done >= 0) {
        split_finish<NC>(P, L, done);
        return false;
    }
    return true;
}

template <int NC, bool kN>
__global__ __launch_bounds__(64) MPMC_SPLIT_ATTR void k_split(Pipe<NC> P, const double* __restrict__ mixc, int resume) {
    Mix<NC>* const smix = stage_mixes<NC>(mixc, P.nmix);
    const int64_t nu = static_cast<int64_t>(P.g[G_NU]);
    const int64_t total = kAttempts * nu;
    SplitLane<NC> L;
    bool active = false, exhausted = false;
    for (;;) {
        const bool want = !active && !exhausted;
        const int64_t task = fetch_task(&P.g[resume ? G_TASK_SPLIT_RESUME : G_TASK_SPLIT], want);
        if (want) {
            if (task >= total) { exhausted = true; }
            else {
                int k;
                int64_t u;
                task_index<MPMC_SPLIT_ORDER>(task, nu, kAttempts, k, u);
                if (MPMC_SPLIT_LPT) { u = P.u_order[u]; }
                active = split_begin<NC>(P, k, u, resume != 0, L);
            }
        }
        if (!__any(active)) {
            if (!__any(!exhausted)) { break; }
            continue;
        }
        if (!active) { continue; }
        active = split_step<NC, kN>(P, smix, L);
    }
}

// Selection of the lowest-Gibbs converged attempt in attempt order, the feed Gibbs
// comparison and the review tolerance (pt_split.hpp:459-507).
template <int NC>
__device__ __forceinline__ void split_select_cell(const Pipe<NC>& P, int64_t u) {
    const int64_t n = P.n;
    {
        const int64_t c = P.u_cell[u];
        P.u_sel[u] = -1;
        P.u_v[u] = V_INDETERMINATE;
        P.u_ftol[u] = __builtin_nan("");
        P.u_dgbad[u] = 0;
        if (P.flags[c] & CF_REJECTED) { return; }
        const int na = 2 * P.u_nwit[u];
        uint64_t cum = 0;
        int best = -1;
        double best_g = 0;
        bool fallback = false;
        bool disappearance = false;
        for (int j = 0; j < na; ++j) {
            if (cum >= 20000) { fallback = true; break; }
            const uint32_t ev = P.a_ev[j * n + u];
            if (cum + ev > 20000) { fallback = true; }
            disappearance = disappearance || (P.a_st[j * n + u] & 0xFF) == A_DISAPPEARANCE;
            if ((P.a_st[j * n + u] & 0xFF) == A_CONVERGED) {
                const double g = P.a_s[(j * 6 + 0) * n + u];
                if (best < 0 || g < best_g) { best = j; best_g = g; }
            }
            cum += ev;
        }
        if (fallback) { atomicOr(&P.flags[c], CF_FALLBACK); return; }
        P.u_sel[u] = best;
        // bit 0: the pair's Gibbs energy exceeds the feed's; bit 1: a disappearance attempt
        P.u_dgbad[u] = disappearance ? 2 : 0;
        if (best < 0) { return; }
        double feed_g = 0, corr = 0, mag = 1;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            const double z = P.feed[i * n + c];
            if (z == 0) { continue; }
            const double lz = log(z);
            const double ph = P.refl[i * n + c];
            kadd(z * (lz + ph), feed_g, corr);
            mag += z * (__builtin_fabs(lz) + __builtin_fabs(ph));
        }
        const double dg = best_g - feed_g;
        const double guard = P.a_s[(best * 6 + 5) * n + u] + 256 * kEps * mag;
        const double ftol = 1e-10 + 0.5 * P.a_s[(best * 6 + 4) * n + u];
        if (!finite(dg) || !finite(guard) || !finite(ftol)) { return; }
        P.u_v[u] = -1;
        P.u_ftol[u] = ftol;
        P.u_dgbad[u] |= dg > guard ? 1 : 0;
    }
}

template <int NC>
__global__ __launch_bounds__(64) void k_split_select(Pipe<NC> P) {
    const int64_t nu = static_cast<int64_t>(P.g[G_NU]);
    for (int64_t u = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; u < nu;
         u += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        split_select_cell<NC>(P, u);
    }
}

template <int NC>
__device__ __forceinline__ void review_reduce_cell(const Pipe<NC>& P, int64_t u) {
    const int64_t n = P.n;
    {
        if (P.u_v[u] != -1) { return; }
        const int64_t c = P.u_cell[u];
        if (P.flags[c] & (CF_REJECTED | CF_FALLBACK)) { return; }
        const int ntr = review_trials<NC>(P.refinfo[c]);
        uint64_t evals = 0;
        bool neg = false, all_stat = true, ambiguous = false;
        for (int k = 0; k < ntr; ++k) {
            const int st = P.fst[k * n + u] & 0xFF;
            evals += P.fev[k * n + u];
            neg = neg || st == T_NEGATIVE;
            all_stat = all_stat && st == T_STATIONARY;
            ambiguous = ambiguous || st == T_AMBIG;
        }
        if (evals > 100000 || ambiguous) { atomicOr(&P.flags[c], CF_FALLBACK); return; }
        const int fin = neg ? S_UNSTABLE : (all_stat ? S_NO_INSTABILITY : S_INDETERMINATE);
        if (P.mixed == 1 && fin == S_UNSTABLE) {
            // mixed precision: a three-phase candidate's review is searched again in FP64 so
            // that the three-phase tail starts from the reference's witnesses
            for (int k = 0; k < ntr; ++k) { P.fst[k * n + u] = T_REDO; }
            return;
        }
        P.u_v[u] = fin == S_UNSTABLE ? V_UNSTABLE
                 : ((P.u_dgbad[u] & 1) ? V_INDETERMINATE : (fin == S_NO_INSTABILITY ? V_TWO : V_INDETERMINATE));
    }
}

template <int NC>
__global__ __launch_bounds__(64) void k_review_reduce(Pipe<NC> P) {
    const int64_t nu = static_cast<int64_t>(P.g[G_NU]);
    for (int64_t u = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; u < nu;
         u += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        review_reduce_cell<NC>(P, u);
    }
}

// ------------------------------------------------------------------ max3 tail
template <int NC, bool kN>
// One active lane per wave: the three-phase tail runs the serial searches through
// non-inlined calls, whose outcome came to depend on the other cells of the wave when
// several ran together (flash_kernels.hpp, serial kernels). Lane 0 takes one unstable cell
// at a time.
__global__ __launch_bounds__(64) void k_max3_tail(Pipe<NC> P, const double* __restrict__ mixc) {
    Mix<NC>* const smix = stage_mixes<NC>(mixc, P.nmix);
    if (threadIdx.x != 0) { return; }
    const int64_t n = P.n;
    const int64_t nu = static_cast<int64_t>(P.g[G_NU]);
    bool exhausted = false;
    while (!exhausted) {
        const int64_t u = static_cast<int64_t>(atomicAdd(&P.g[G_TASK_MAX3], 1ull));
        if (u >= nu) { exhausted = true; continue; }
        if (P.u_v[u] != V_UNSTABLE || P.max_phases != 3) { continue; }
        const int64_t c = P.u_cell[u];
        if (P.flags[c] & (CF_REJECTED | CF_FALLBACK)) { continue; }
        const int sel = P.u_sel[u];
        double z[NC], x[NC], y[NC];
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            z[i] = P.feed[i * n + c];
            x[i] = P.a_v[((sel * 5 + 0) * NC + i) * n + u];
            y[i] = P.a_v[((sel * 5 + 1) * NC + i) * n + u];
        }
        const double beta = P.a_s[(sel * 6 + 1) * n + u];
        const int ntr = review_trials<NC>(P.refinfo[c]);
        int nwit = 0;
        for (int k = 0; k < ntr; ++k) { nwit += (P.fst[k * n + u] & 0xFF) == T_NEGATIVE; }
        Counters cnt{0, 0, 0, 0};
        bool rejected = false;
        Max3Tail<NC> M;
        SuppliedStarts<NC> S;
        S.load(P, c);
        max3_tail<NC, kN>(smix[mix_at(P, c)], P.p[c], P.t[c], z, x, y, beta, nwit,
                      [&](int i, double* w, int& branch) {
                          int seen = -1, k = 0;
                          for (; k < ntr; ++k) {
                              if ((P.fst[k * n + u] & 0xFF) == T_NEGATIVE && ++seen == i) { break; }
                          }
#pragma unroll
                          for (int q = 0; q < NC; ++q) { w[q] = P.fw[(k * NC + q) * n + u]; }
                          branch = (P.fst[k * n + u] >> 8) & 0xFF;
                      },
                      P.fug_tol, M, cnt, rejected, S.start(0), S.sf0, S.start(1), S.sf1);
        atomicAdd(&P.cnt[0 * n + c], cnt.prop);
        atomicAdd(&P.cnt[1 * n + c], cnt.tpd);
        atomicAdd(&P.cnt[2 * n + c], cnt.rr);
        atomicAdd(&P.cnt[3 * n + c], cnt.rr3);
        if (rejected) { atomicOr(&P.flags[c], CF_REJECTED); continue; }
        P.m_st[u] = M.status | (M.use_neighbor ? 0x100 : 0) | (M.attempts > 0 ? 0x200 : 0) |
                    (M.start_consumed ? 0x400 : 0);
        if (M.status == M_THREE) {
#pragma unroll
            for (int ph = 0; ph < 3; ++ph) {
                P.m_s[ph * n + u] = M.three.frac[ph];
                P.m_s[(3 + ph) * n + u] = M.three.z[ph];
                P.m_br[ph * n + u] = M.three.act[ph].branch | 0x100;
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    P.m_v[(ph * NC + i) * n + u] = M.three.comp[ph][i];
                    P.m_v[((3 + ph) * NC + i) * n + u] = M.three.act[ph].lnphi[i];
                }
            }
        } else if (M.status == M_TWO && M.use_neighbor) {
            const SplitState<NC>& s = M.neighbor;
            P.m_s[0 * n + u] = 1.0 - s.fr.beta;
            P.m_s[1 * n + u] = s.fr.beta;
            P.m_s[3 * n + u] = s.zl;
            P.m_s[4 * n + u] = s.zv;
            P.m_br[0 * n + u] = s.liq.branch | 0x100;
            P.m_br[1 * n + u] = s.vap.branch | 0x100;
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                P.m_v[(0 * NC + i) * n + u] = s.fr.x[i];
                P.m_v[(1 * NC + i) * n + u] = s.fr.y[i];
                P.m_v[(3 * NC + i) * n + u] = s.liq.lnphi[i];
                P.m_v[(4 * NC + i) * n + u] = s.vap.lnphi[i];
            }
        }
    }
}

// ------------------------------------------------------------------ publication
template <int NC>
__device__ void put_record(const FlashOut& o, int64_t n, int64_t c, int set, int max3, int split, int np,
                           const double* beta, const double* zf, const int32_t* br, const double* x,
                           const double* lp, int flags = 0) {
    o.status[c] = set | (max3 << 4) | (split << 8) | flags;
    o.phase_count[c] = np;
    const double nan = __builtin_nan("");
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const bool have = k < np;
        o.beta[k * n + c] = have ? beta[k] : nan;
        o.zfactor[k * n + c] = have ? zf[k] : nan;
        o.branch[k * n + c] = have ? br[k] : -1;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            o.x[(k * NC + i) * n + c] = have ? x[k * NC + i] : nan;
            o.ln_phi[(k * NC + i) * n + c] = have ? lp[k * NC + i] : nan;
        }
    }
}

template <int NC>
__global__ __launch_bounds__(64) void k_publish(Pipe<NC> P) {
    const int64_t n = P.n;
    const bool max3 = P.max_phases == 3;
    for (int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < n;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int32_t flags = P.flags[c];
        if (flags & CF_FALLBACK) {
            const int64_t slot = static_cast<int64_t>(atomicAdd(&P.g[G_NFALLBACK], 1ull));
            P.fb_list[slot] = static_cast<int32_t>(c);
            continue;
        }
        if (P.out.counters) {
#pragma unroll
            for (int k = 0; k < 4; ++k) { P.out.counters[k * n + c] = P.cnt[k * n + c]; }
        }
        double beta[3], zf[3], x[3 * NC], lp[3 * NC];
        int32_t br[3];
        const double nan = __builtin_nan("");
        if (flags & CF_REJECTED) {
            put_record<NC>(P.out, n, c, 2, 0xF, 0xF, 0, beta, zf, br, x, lp);
            P.out.status[c] |= 0x2000;
            continue;
        }
        const int cs = P.cstat[c];
        const int m3 = max3 ? 0 : 0xF;
        if (cs == S_NO_INSTABILITY) {
            beta[0] = 1.0;
            zf[0] = nan;
            br[0] = P.refinfo[c] & 0x1FF;
#pragma unroll
            for (int i = 0; i < NC; ++i) { x[i] = P.feed[i * n + c]; lp[i] = P.refl[i * n + c]; }
            put_record<NC>(P.out, n, c, 0, max3 ? M_SINGLE : 0xF, V_SINGLE, 1, beta, zf, br, x, lp);
            continue;
        }
        if (cs != S_UNSTABLE) {
            put_record<NC>(P.out, n, c, 2, max3 ? M_INDETERMINATE : 0xF, V_INDETERMINATE, 0, beta, zf, br, x, lp);
            continue;
        }
        const int64_t u = P.cell_u[c];
        const int v = P.u_v[u];
        const int sel = P.u_sel[u];
        // MPMC_FLAG_SPLIT_DISAPPEARANCE, MPMC_FLAG_SPLIT_CANDIDATE, MPMC_FLAG_MAX3_ATTEMPTED
        const int fl = ((P.u_dgbad[u] & 2) ? 0x4000 : 0) | (sel >= 0 ? 0x8000 : 0) |
                       (max3 && v == V_UNSTABLE && (P.m_st[u] & 0x200) ? 0x10000 : 0) |
                       (max3 && v == V_UNSTABLE && (P.m_st[u] & 0x400) ? 0x20000 : 0);
        auto load_pair = [&]() {
            const int32_t st = P.a_st[sel * n + u];
            beta[1] = P.a_s[(sel * 6 + 1) * n + u];
            beta[0] = 1.0 - beta[1];
            zf[0] = P.a_s[(sel * 6 + 2) * n + u];
            zf[1] = P.a_s[(sel * 6 + 3) * n + u];
            br[0] = ((st >> 8) & 0xFF) | 0x100;
            br[1] = ((st >> 16) & 0xFF) | 0x100;
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                x[i] = P.a_v[((sel * 5 + 0) * NC + i) * n + u];
                x[NC + i] = P.a_v[((sel * 5 + 1) * NC + i) * n + u];
                lp[i] = P.a_v[((sel * 5 + 2) * NC + i) * n + u];
                lp[NC + i] = P.a_v[((sel * 5 + 3) * NC + i) * n + u];
            }
            return 2;
        };
        (void)m3;
        if (!max3) {
            const int set = v == V_TWO ? 0 : (v == V_UNSTABLE ? 1 : 2);
            const int np = sel >= 0 ? load_pair() : 0;
            put_record<NC>(P.out, n, c, set, 0xF, v, np, beta, zf, br, x, lp, fl);
            continue;
        }
        if (v == V_TWO) {
            put_record<NC>(P.out, n, c, 0, M_TWO, v, load_pair(), beta, zf, br, x, lp, fl);
            continue;
        }
        if (v != V_UNSTABLE) {
            put_record<NC>(P.out, n, c, 2, M_INDETERMINATE, v, 0, beta, zf, br, x, lp, fl);
            continue;
        }
        const int32_t ms = P.m_st[u];
        const int mstat = ms & 0xFF;
        if (mstat == M_THREE || (mstat == M_TWO && (ms & 0x100))) {
            const int np = mstat == M_THREE ? 3 : 2;
            for (int ph = 0; ph < np; ++ph) {
                beta[ph] = P.m_s[ph * n + u];
                zf[ph] = P.m_s[(3 + ph) * n + u];
                br[ph] = P.m_br[ph * n + u];
#pragma unroll
                for (int i = 0; i < NC; ++i) {
                    x[ph * NC + i] = P.m_v[(ph * NC + i) * n + u];
                    lp[ph * NC + i] = P.m_v[((3 + ph) * NC + i) * n + u];
                }
            }
            put_record<NC>(P.out, n, c, 0, mstat, v, np, beta, zf, br, x, lp, fl);
        } else if (mstat == M_HIGHER) {
            put_record<NC>(P.out, n, c, 1, M_HIGHER, v, load_pair(), beta, zf, br, x, lp, fl);
        } else {
            put_record<NC>(P.out, n, c, 2, mstat, v, 0, beta, zf, br, x, lp, fl);
        }
    }
}

// Cells whose searches a shared evaluation budget would have cut: the serial path.
template <int NC, bool kN>
__global__ __launch_bounds__(64) void k_fallback(Pipe<NC> P, const double* __restrict__ mixc) {
    Mix<NC>* const smix = stage_mixes<NC>(mixc, P.nmix);
    if (threadIdx.x != 0) { return; }   // one active lane per wave, as in k_max3_tail
    const int64_t n = P.n;
    const int64_t nf = static_cast<int64_t>(P.g[G_NFALLBACK]);
    bool exhausted = false;
    while (!exhausted) {
        const int64_t i = static_cast<int64_t>(atomicAdd(&P.g[G_TASK_FALLBACK], 1ull));
        if (i >= nf) { exhausted = true; continue; }
        const int64_t c = P.fb_list[i];
        double feed[NC];
#pragma unroll
        for (int q = 0; q < NC; ++q) { feed[q] = P.zin[q * n + c]; }
        Counters cnt{0, 0, 0, 0};
        bool rejected = false;
        PhaseView<NC> ph[3];
        double hw[kMaxHints][NC], fw[kMaxHints][NC];
        const double* hs[kMaxHints];
        const double* fs[kMaxHints];
        int nh = 0, nfh = 0;
        for (int h = 0; h < P.nhint; ++h) {
            if (__builtin_isnan(P.hints[(h * NC + 0) * n + c])) { continue; }
            for (int q = 0; q < NC; ++q) { hw[nh][q] = P.hints[(h * NC + q) * n + c]; }
            hs[nh] = hw[nh];
            ++nh;
        }
        for (int h = 0; h < P.nfhint; ++h) {
            if (__builtin_isnan(P.fhints[(h * NC + 0) * n + c])) { continue; }
            for (int q = 0; q < NC; ++q) { fw[nfh][q] = P.fhints[(h * NC + q) * n + c]; }
            fs[nfh] = fw[nfh];
            ++nfh;
        }
        if (P.max_phases == 2) {
            VleOut<NC> V;
            solve_vle<NC, kN>(smix[mix_at(P, c)], P.p[c], P.t[c], feed, P.fug_tol, hs, nh, V, cnt, rejected, fs, nfh);
            if (rejected) { write_cell<NC>(P.out, n, c, 2, 0xF, 0xF, 0, ph, cnt, 0x2000); }
            else { publish_vle<NC>(P.out, n, c, V, 0xF, cnt); }
        } else {
            Max3Out<NC> M;
            SuppliedStarts<NC> S;
            S.load(P, c);
            solve_max3<NC, kN>(smix[mix_at(P, c)], P.p[c], P.t[c], feed, P.fug_tol, M, cnt, rejected, hs, nh, fs,
                               nfh, S.start(0), S.sf0, S.start(1), S.sf1);
            if (rejected) { write_cell<NC>(P.out, n, c, 2, 0xF, 0xF, 0, ph, cnt, 0x2000); }
            else { publish_max3<NC>(P.out, n, c, M, cnt); }
        }
    }
}

} // namespace mpmc::dev

#include "mixed_pipeline.hpp"

namespace mpmc::dev {

// ------------------------------------------------------------------ layout
template <int NC>
__host__ size_t pipe_carve(Pipe<NC>& P, char* base, int64_t n) {
    size_t off = 0;
    auto take = [&](auto*& ptr, size_t count) {
        using T = std::remove_reference_t<decltype(*ptr)>;
        off = (off + 255) & ~size_t{255};
        ptr = reinterpret_cast<T*>(base ? base + off : nullptr);
        off += count * sizeof(T);
    };
    const size_t N = static_cast<size_t>(n);
    take(P.feed, NC * N);
    take(P.refl, NC * N);
    take(P.refinfo, N);
    take(P.flags, N);
    take(P.cstat, N);
    take(P.cell_u, N);
    take(P.cnt, 4 * N);
    take(P.tst, (NC + 2 + kMaxHints) * N);
    take(P.tev, (NC + 2 + kMaxHints) * N);
    take(P.tw, (NC + 2 + kMaxHints) * NC * N);
    take(P.u_cell, N);
    take(P.u_order, N);
    take(P.u_nwit, N);
    take(P.u_wit, kWitnesses * N);
    take(P.a_st, kAttempts * N);
    take(P.a_ev, kAttempts * N);
    take(P.a_s, kAttempts * 6 * N);
    take(P.a_v, kAttempts * 5 * NC * N);
    take(P.u_sel, N);
    take(P.u_v, N);
    take(P.u_ftol, N);
    take(P.u_dgbad, N);
    take(P.fst, (NC + 4 + kMaxHints) * N);
    take(P.fev, (NC + 4 + kMaxHints) * N);
    take(P.fw, (NC + 4 + kMaxHints) * NC * N);
    take(P.m_st, N);
    take(P.m_s, 6 * N);
    take(P.m_v, 6 * NC * N);
    take(P.m_br, 3 * N);
    take(P.g, G_SLOTS);
    take(P.fb_list, N);
    take(P.ra32, NC * N);
    take(P.lz32, NC * N);
    take(P.d32, NC * N);
    return (off + 255) & ~size_t{255};
}

} // namespace mpmc::dev

namespace mpmc::dev {

template <int NC>
size_t pipeline_bytes(int64_t n) {
    Pipe<NC> P{};
    return pipe_carve<NC>(P, nullptr, n);
}

// Persistent-grid size: resident workgroups per CU for `kernel` (with `lds` bytes of
// dynamic LDS) times the CU count.
template <typename K>
inline unsigned persistent_blocks(K kernel, size_t lds) {
    int dev = 0, cus = 0, per_cu = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64, lds) != hipSuccess || per_cu < 1) { per_cu = 1; }
    if (cus < 1) { cus = 1; }
    return static_cast<unsigned>(cus * per_cu);
}

// Grids of the persistent kernels for one LDS size (the one-mixture size is computed
// once; a multi-mixture table is sized per call).
struct PipeGrids {
    unsigned tpd, split, tail, fb, tpd32, split32;
    unsigned prep, publish, polish;   // caps of the grid-stride kernels with private memory
};

// Largest grid up to 16384 workgroups whose scratch reservation (private bytes per lane x
// 64 x workgroups) stays within kDispatchScratch (at least 64 workgroups).
template <typename K>
inline unsigned scratch_blocks(K kernel) {
    return capped_cell_blocks(int64_t{16384} * kBlock, kernel_lane_scratch(kernel));
}
template <int NC>
PipeGrids pipe_grids(size_t lds) {
    PipeGrids g{};
    g.tpd = persistent_blocks(k_tpd<NC, false>, lds);
    g.split = persistent_blocks(k_split<NC, false>, lds);
    // The three-phase tail and the serial fallback carry 8-22 KB of private memory per lane
    // and serve few cells; their grids are capped so that one dispatch reserves at most
    // ~100 MB of scratch (8192 and 1024 lanes), which keeps several launch streams of
    // such dispatches within the runtime's scratch limits.
    g.tail = std::min(persistent_blocks(k_max3_tail<NC, false>, lds), 256u);
    g.fb = std::min(persistent_blocks(k_fallback<NC, false>, lds), 64u);
    g.tpd32 = persistent_blocks(k_tpd32<NC>, lds);
    g.split32 = persistent_blocks(k_split32<NC>, lds);
    g.prep = scratch_blocks(k_prep<NC, false>);
    g.publish = scratch_blocks(k_publish<NC>);
    g.polish = scratch_blocks(k_polish<NC>);
    return g;
}

// The FP64 schedule.
template <int NC>
void launch_fp64(const Pipe<NC>& P, const double* mixc, int max_phases, const PipeGrids& b, size_t lds,
                 hipStream_t s) {
    const int64_t n = P.n;
    const unsigned cell_blocks = static_cast<unsigned>(std::min<int64_t>((n + 63) / 64, 16384));
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, P.g);
    // a multi-mixture table is staged in LDS by every workgroup: fewer, longer-lived
    // workgroups for the cell-parallel preparation
    const unsigned prep_blocks = std::min(cell_blocks, P.nmix > 1 ? std::min(b.prep, 2048u) : b.prep);
    hipLaunchKernelGGL((k_prep<NC, false>), dim3(prep_blocks), dim3(64), lds, s, P, mixc);
    hipLaunchKernelGGL((k_tpd<NC, false>), dim3(b.tpd), dim3(64), lds, s, P, mixc, 0, 0);
    hipLaunchKernelGGL(k_init_reduce<NC>, dim3(cell_blocks), dim3(64), 0, s, P);
    if (MPMC_SPLIT_LPT) { hipLaunchKernelGGL(k_order_u<NC>, dim3(cell_blocks), dim3(64), 0, s, P); }
    hipLaunchKernelGGL((k_split<NC, false>), dim3(b.split), dim3(64), lds, s, P, mixc, 0);
    hipLaunchKernelGGL(k_split_select<NC>, dim3(cell_blocks), dim3(64), 0, s, P);
    hipLaunchKernelGGL((k_tpd<NC, false>), dim3(b.tpd), dim3(64), lds, s, P, mixc, 1, 0);
    hipLaunchKernelGGL(k_review_reduce<NC>, dim3(cell_blocks), dim3(64), 0, s, P);
    if (max_phases == 3) { hipLaunchKernelGGL((k_max3_tail<NC, false>), dim3(b.tail), dim3(64), lds, s, P, mixc); }
    hipLaunchKernelGGL(k_publish<NC>, dim3(std::min(cell_blocks, b.publish)), dim3(64), 0, s, P);
    hipLaunchKernelGGL((k_fallback<NC, false>), dim3(b.fb), dim3(64), lds, s, P, mixc);
}

// Fast mixed precision (precision 2): reduced-precision searches, FP64 continuation of
// the searches they leave ambiguous, FP64 polish of the selected split, FP64 three-phase
// tail. Decisions can differ from FP64 where the reference's own split search fails.
template <int NC>
void launch_mixed(Pipe<NC> P, const double* mixc, int max_phases, const PipeGrids& b, size_t lds, hipStream_t s) {
    const int64_t n = P.n;
    const unsigned cell_blocks = static_cast<unsigned>(std::min<int64_t>((n + 63) / 64, 16384));
    const unsigned prep_blocks = std::min(cell_blocks, P.nmix > 1 ? std::min(b.prep, 2048u) : b.prep);
    P.mixed = 1;
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, P.g);
    hipLaunchKernelGGL((k_prep<NC, false>), dim3(prep_blocks), dim3(64), lds, s, P, mixc);
    hipLaunchKernelGGL(k_tpd32<NC>, dim3(b.tpd32), dim3(64), lds, s, P, mixc, 0);
    hipLaunchKernelGGL((k_tpd<NC, false>), dim3(b.tpd), dim3(64), lds, s, P, mixc, 0, 1);
    hipLaunchKernelGGL(k_init_reduce<NC>, dim3(cell_blocks), dim3(64), 0, s, P);
    if (MPMC_SPLIT_LPT) { hipLaunchKernelGGL(k_order_u<NC>, dim3(cell_blocks), dim3(64), 0, s, P); }
    hipLaunchKernelGGL(k_split32<NC>, dim3(b.split32), dim3(64), lds, s, P, mixc);
    hipLaunchKernelGGL(k_split_gate<NC>, dim3(cell_blocks), dim3(64), 0, s, P);
    hipLaunchKernelGGL((k_split<NC, false>), dim3(b.split), dim3(64), lds, s, P, mixc, 1);
    hipLaunchKernelGGL(k_polish<NC>, dim3(std::min(cell_blocks, b.polish)), dim3(64), lds, s, P, mixc);
    hipLaunchKernelGGL(k_tpd32<NC>, dim3(b.tpd32), dim3(64), lds, s, P, mixc, 1);
    hipLaunchKernelGGL((k_tpd<NC, false>), dim3(b.tpd), dim3(64), lds, s, P, mixc, 1, 1);
    hipLaunchKernelGGL(k_review_reduce<NC>, dim3(cell_blocks), dim3(64), 0, s, P);
    hipLaunchKernelGGL((k_tpd<NC, false>), dim3(b.tpd), dim3(64), lds, s, P, mixc, 1, 2);
    P.mixed = 2;
    hipLaunchKernelGGL(k_review_reduce<NC>, dim3(cell_blocks), dim3(64), 0, s, P);
    if (max_phases == 3) { hipLaunchKernelGGL((k_max3_tail<NC, false>), dim3(b.tail), dim3(64), lds, s, P, mixc); }
    hipLaunchKernelGGL(k_publish<NC>, dim3(std::min(cell_blocks, b.publish)), dim3(64), 0, s, P);
    hipLaunchKernelGGL((k_fallback<NC, false>), dim3(b.fb), dim3(64), lds, s, P, mixc);
}

// Decision-exact mixed precision (precision 1). FP32 descents screen every stability trial;
// each one whose outcome bears on a decision is searched in FP64 from its start: the
// trials FP32 cannot classify beyond rounding reach, the negative trials of every cell
// that has one (their end points seed the splits), every split attempt, and the whole
// review of a cell whose review finds instability. The trials FP32 classifies as
// stationary beyond rounding reach keep that status. Every search whose end point is
// used is therefore the FP64 schedule's own, and so is every decision built on them.
template <int NC>
void launch_mixed_exact(Pipe<NC> P, const double* mixc, int max_phases, const PipeGrids& b, size_t lds,
                        hipStream_t s) {
    const int64_t n = P.n;
    const unsigned cell_blocks = static_cast<unsigned>(std::min<int64_t>((n + 63) / 64, 16384));
    const unsigned prep_blocks = std::min(cell_blocks, P.nmix > 1 ? std::min(b.prep, 2048u) : b.prep);
    P.mixed = 1;
    P.exact = 1;
    hipLaunchKernelGGL(k_reset, dim3(1), dim3(64), 0, s, P.g);
    hipLaunchKernelGGL((k_prep<NC, false>), dim3(prep_blocks), dim3(64), lds, s, P, mixc);
    hipLaunchKernelGGL(k_tpd32<NC>, dim3(b.tpd32), dim3(64), lds, s, P, mixc, 0);
    hipLaunchKernelGGL(k_witness_redo<NC>, dim3(cell_blocks), dim3(64), 0, s, P);
    hipLaunchKernelGGL((k_tpd<NC, false>), dim3(b.tpd), dim3(64), lds, s, P, mixc, 0, 2);
    hipLaunchKernelGGL(k_init_reduce<NC>, dim3(cell_blocks), dim3(64), 0, s, P);
    if (MPMC_SPLIT_LPT) { hipLaunchKernelGGL(k_order_u<NC>, dim3(cell_blocks), dim3(64), 0, s, P); }
    hipLaunchKernelGGL((k_split<NC, false>), dim3(b.split), dim3(64), lds, s, P, mixc, 0);
    hipLaunchKernelGGL(k_split_select<NC>, dim3(cell_blocks), dim3(64), 0, s, P);
    hipLaunchKernelGGL(k_tpd32<NC>, dim3(b.tpd32), dim3(64), lds, s, P, mixc, 1);
    hipLaunchKernelGGL((k_tpd<NC, false>), dim3(b.tpd), dim3(64), lds, s, P, mixc, 1, 2);
    hipLaunchKernelGGL(k_review_reduce<NC>, dim3(cell_blocks), dim3(64), 0, s, P);
    hipLaunchKernelGGL((k_tpd<NC, false>), dim3(b.tpd), dim3(64), lds, s, P, mixc, 1, 3);
    P.mixed = 2;
    hipLaunchKernelGGL(k_review_reduce<NC>, dim3(cell_blocks), dim3(64), 0, s, P);
    if (max_phases == 3) { hipLaunchKernelGGL((k_max3_tail<NC, false>), dim3(b.tail), dim3(64), lds, s, P, mixc); }
    hipLaunchKernelGGL(k_publish<NC>, dim3(std::min(cell_blocks, b.publish)), dim3(64), 0, s, P);
    hipLaunchKernelGGL((k_fallback<NC, false>), dim3(b.fb), dim3(64), lds, s, P, mixc);
}

template <int NC>
hipError_t launch_pipeline_prec(const Pipe<NC>& P, const double* mixc, int max_phases, int precision,
                                hipStream_t s) {
    const size_t lds = static_cast<size_t>(P.nmix > 1 ? P.nmix : 1) * sizeof(Mix<NC>);
    static const PipeGrids one = pipe_grids<NC>(sizeof(Mix<NC>));
    const PipeGrids b = P.nmix > 1 ? pipe_grids<NC>(lds) : one;
    if (precision == 1) { launch_mixed_exact<NC>(P, mixc, max_phases, b, lds, s); }
    else if (precision == 2) { launch_mixed<NC>(P, mixc, max_phases, b, lds, s); }
    else { launch_fp64<NC>(P, mixc, max_phases, b, lds, s); }
    return hipGetLastError();
}

// precision 0: FP64 evaluations with the reference's bisection roots (bit-level
// parity); 1: decision-exact mixed precision (FP32 screening, every decision-bearing
// search in FP64); 2: fast mixed precision (reduced-precision searches, FP64 roots,
// decisions and polish, FP64 schedule for every cell the searches cannot settle).
//
// mixc: nmix consecutive Mix<NC> records; mix_of[n] (device) names each cell's record,
// nullptr when nmix == 1.
template <int NC>
hipError_t launch_pipeline(const double* mixc, int nmix, const int32_t* mix_of, void* workspace,
                           size_t workspace_bytes, int64_t n, const double* p, const double* t, const double* z,
                           int max_phases, double fug_tol, int precision, const FlashOut& out, hipStream_t s,
                           int nhint, const double* hints, int nfhint, const double* fhints, const double* tpz,
                           const double* tpf) {
    Pipe<NC> P{};
    if (nhint < 0 || nhint > kMaxHints || (nhint > 0 && !hints)) { return hipErrorInvalidValue; }
    if (nfhint < 0 || nfhint > kMaxHints || (nfhint > 0 && !fhints)) { return hipErrorInvalidValue; }
    if ((tpz == nullptr) != (tpf == nullptr) || (tpz && max_phases != 3)) { return hipErrorInvalidValue; }
    P.nhint = nhint;
    P.hints = hints;
    P.nfhint = nfhint;
    P.fhints = fhints;
    P.tpz = tpz;
    P.tpf = tpf;
    const size_t need = pipe_carve<NC>(P, static_cast<char*>(workspace), n);
    if (!workspace || need > workspace_bytes) { return hipErrorInvalidValue; }
    if (nmix < 1 || (nmix > 1 && !mix_of)) { return hipErrorInvalidValue; }
    P.mix_of = nmix > 1 ? mix_of : nullptr;
    P.nmix = nmix;
    P.n = n;
    P.max_phases = max_phases;
    P.fug_tol = fug_tol;
    P.p = p;
    P.t = t;
    P.zin = z;
    P.out = out;
    P.mixed = 0;
    P.exact = 0;
    return launch_pipeline_prec<NC>(P, mixc, max_phases, precision, s);
}

#define MPMC_DECLARE_PIPELINE(N)                                                                                  \
    extern template size_t pipeline_bytes<N>(int64_t);                                                            \
    extern template hipError_t launch_pipeline<N>(const double*, int, const int32_t*, void*, size_t, int64_t,     \
                                                  const double*, const double*, const double*, int, double, int,  \
                                                  const FlashOut&, hipStream_t, int, const double*, int,    \
                                                  const double*, const double*, const double*);
#define MPMC_INSTANTIATE_PIPELINE(N)                                                                              \
    template size_t pipeline_bytes<N>(int64_t);                                                                   \
    template hipError_t launch_pipeline<N>(const double*, int, const int32_t*, void*, size_t, int64_t,            \
                                           const double*, const double*, const double*, int, double, int,         \
                                           const FlashOut&, hipStream_t, int, const double*, int,           \
                                           const double*, const double*, const double*);

} // namespace mpmc::dev
