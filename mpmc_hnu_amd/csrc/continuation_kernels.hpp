// Ordered-path PT continuation (solve_pr76_pt_continuation, pr76_pt_continuation.hpp:194-310):
// the elementwise kernels between the per-point flashes of a batch of paths. Point k of
// every path is one max3 flash of the staged pipeline over the npath cells; these kernels
// turn point k-1's published record into point k's warm starts and, after the last point,
// reject every point of a path the reference would have rejected as a whole.
//
// Per-path buffers (n = npath, cell index fastest):
//   feed    [NC][n]        the path feed normalised once (pr76_pt_continuation.hpp:218-219)
//   hints   [6][NC][n]     initial stability starts (carried, then fallback), NaN past the list
//   fhints  [6][NC][n]     final-review starts (carried, then fallback)
//   tpz     [2][3][NC][n]  three-phase start slots: 0 carried state, 1 the fixed start
//   tpf     [2][2][n]      their phase-fraction seeds of phases 1, 2 (NaN: slot empty)
#pragma once

#include "flash_kernels.hpp"

namespace mpmc::dev {

// Point-k slice of a record batch laid out [npoint][...][n].
template <int NC>
__host__ __device__ inline FlashOut point_slice(const FlashOut& o, int64_t n, int k) {
    const int64_t c = static_cast<int64_t>(k) * n;
    return FlashOut{o.status + c,       o.phase_count + c,       o.beta + 3 * c,       o.zfactor + 3 * c,
                    o.branch + 3 * c,   o.x + 3 * NC * c,        o.ln_phi + 3 * NC * c,
                    o.counters ? o.counters + 4 * c : nullptr};
}

// Normalise each path's feed once. A feed the composition check rejects is kept as given, so
// that every point's flash rejects it (the reference rejects the path before any solve).
template <int NC>
__global__ __launch_bounds__(256) void k_cont_feed(int64_t n, const double* __restrict__ z,
                                                   double* __restrict__ feed) {
    for (int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < n;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        double w[NC];
#pragma unroll
        for (int i = 0; i < NC; ++i) { w[i] = z[i * n + c]; }
        bool rejected = false;
        const double sum = check_composition<NC>(w, rejected);
#pragma unroll
        for (int i = 0; i < NC; ++i) { feed[i * n + c] = rejected ? w[i] : w[i] / sum; }
    }
}

// Batch-wide options of the continuation (Pr76PtContinuationOptions,
// pr76_pt_continuation.hpp:27-34): carry bit 0 carries phase compositions, bit 1 the
// three-phase state; fallback initial / final stability starts (fi, ff); one fixed
// three-phase start (point_options.three_phase_starts) with its fraction seed.
struct ContStarts {
    int carry;
    int nfi, nff;
    int has_start;
    double fi[3][12], ff[3][12], st[3][12], sf[2];
};

// Appends composition v to the cell's start list H ([slots][NC][n]) unless it equals (==,
// component by component) an entry already there (pr76_pt_continuation_append_unique).
template <int NC>
__device__ inline void append_unique(double* H, int64_t n, int64_t c, int& count, const double* v) {
    bool dup = false;
    for (int e = 0; e < count; ++e) {
        bool same = true;
#pragma unroll
        for (int i = 0; i < NC; ++i) { same = same && H[(e * NC + i) * n + c] == v[i]; }
        dup = dup || same;
    }
    if (dup) { return; }
#pragma unroll
    for (int i = 0; i < NC; ++i) { H[(count * NC + i) * n + c] = v[i]; }
    ++count;
}

// Warm starts of point k from point k-1's record `prev` (have_prev false at point 0), as
// solve_pr76_pt_continuation builds them (pr76_pt_continuation.hpp:226-251) from the
// reference's result (:110-161): the accepted phase count comes from the max3 status code
// (single 0, two 1, three 2; anything else accepts nothing); a single phase carries the
// stability test's feed (the path feed normalised again), two phases the published liquid
// and vapour compositions (those of the accepted boundary neighbour where one was
// accepted), three phases the three compositions, and a three-phase point also the state
// (compositions, fractions of phases 1 and 2) for start slot 0. The initial starts are the
// carried compositions then the fallback initial starts, the final-review starts the carried
// compositions then the fallback final starts, each list without repeats; slots up to nh /
// nfh past the list are NaN. Start slot 1 holds the fixed three-phase start.
// info = incoming hint phase count | carried start count << 4 | carried three-phase start << 8.
template <int NC>
__global__ __launch_bounds__(256) void k_cont_hints(int64_t n, FlashOut prev, int have_prev,
                                                    const double* __restrict__ feed, ContStarts o, int nh, int nfh,
                                                    double* __restrict__ hints, double* __restrict__ fhints,
                                                    double* __restrict__ tpz, double* __restrict__ tpf,
                                                    int32_t* __restrict__ info) {
    const double nan = __builtin_nan("");
    for (int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < n;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int code = have_prev ? (prev.status[c] >> 4) & 0xF : 0xF;
        const int count = code <= 2 ? code + 1 : 0;
        double v0[NC], v1[NC], v2[NC];
#pragma unroll
        for (int i = 0; i < NC; ++i) { v0[i] = v1[i] = v2[i] = nan; }
        if (count == 1) {
            double w[NC];
#pragma unroll
            for (int i = 0; i < NC; ++i) { w[i] = feed[i * n + c]; }
            bool rejected = false;
            const double sum = check_composition<NC>(w, rejected);
#pragma unroll
            for (int i = 0; i < NC; ++i) { v0[i] = w[i] / sum; }
        } else if (count > 1) {
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                v0[i] = prev.x[(0 * NC + i) * n + c];
                v1[i] = prev.x[(1 * NC + i) * n + c];
                if (count == 3) { v2[i] = prev.x[(2 * NC + i) * n + c]; }
            }
        }
        const bool carry_comp = (o.carry & 1) != 0 && count > 0;
        for (int list = 0; list < 2; ++list) {
            double* H = list == 0 ? hints : fhints;
            const int slots = list == 0 ? nh : nfh;
            if (slots == 0) { continue; }
            int filled = 0;
            if (carry_comp) {
                append_unique<NC>(H, n, c, filled, v0);
                if (count >= 2) { append_unique<NC>(H, n, c, filled, v1); }
                if (count == 3) { append_unique<NC>(H, n, c, filled, v2); }
            }
            const int nfb = list == 0 ? o.nfi : o.nff;
            for (int h = 0; h < nfb; ++h) {
                double w[NC];
#pragma unroll
                for (int i = 0; i < NC; ++i) { w[i] = list == 0 ? o.fi[h][i] : o.ff[h][i]; }
                append_unique<NC>(H, n, c, filled, w);
            }
            for (int e = filled; e < slots; ++e) {
#pragma unroll
                for (int i = 0; i < NC; ++i) { H[(e * NC + i) * n + c] = nan; }
            }
        }
        const bool three = (o.carry & 2) != 0 && count == 3;
        if (tpz) {
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                tpz[(0 * NC + i) * n + c] = three ? v0[i] : nan;
                tpz[(1 * NC + i) * n + c] = three ? v1[i] : nan;
                tpz[(2 * NC + i) * n + c] = three ? v2[i] : nan;
                tpz[(3 * NC + i) * n + c] = o.has_start ? o.st[0][i] : nan;
                tpz[(4 * NC + i) * n + c] = o.has_start ? o.st[1][i] : nan;
                tpz[(5 * NC + i) * n + c] = o.has_start ? o.st[2][i] : nan;
            }
            tpf[0 * n + c] = three ? prev.beta[1 * n + c] : nan;
            tpf[1 * n + c] = three ? prev.beta[2 * n + c] : nan;
            tpf[2 * n + c] = o.has_start ? o.sf[0] : nan;
            tpf[3 * n + c] = o.has_start ? o.sf[1] : nan;
        }
        if (info) { info[c] = count | ((o.carry & 1) ? count << 4 : 0) | (three ? 0x100 : 0); }
    }
}

// After the last point: a path with a rejected point (status bit 0x2000) is rejected at every
// point, as the reference's continuation throws for the whole path.
template <int NC>
__global__ __launch_bounds__(256) void k_cont_reject(int64_t n, int npoint, FlashOut out,
                                                     int32_t* __restrict__ info) {
    for (int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < n;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        bool rejected = false;
        for (int k = 0; k < npoint; ++k) { rejected = rejected || (out.status[static_cast<int64_t>(k) * n + c] & 0x2000); }
        if (!rejected) { continue; }
        PhaseView<NC> ph[3];
        const Counters cnt{0, 0, 0, 0};
        for (int k = 0; k < npoint; ++k) {
            write_cell<NC>(point_slice<NC>(out, n, k), n, c, 2, 0xF, 0xF, 0, ph, cnt, 0x2000);
            if (info) { info[static_cast<int64_t>(k) * n + c] = 0; }
        }
    }
}

} // namespace mpmc::dev
