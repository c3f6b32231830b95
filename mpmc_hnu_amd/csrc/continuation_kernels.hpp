// Ordered-path PT continuation (solve_pr76_pt_continuation, pr76_pt_continuation.hpp:194-310):
// the elementwise kernels between the per-point flashes of a batch of paths. Point k of
// every path is one max3 flash of the staged pipeline over the npath cells; these kernels
// turn point k-1's published record into point k's warm starts and, after the last point,
// reject every point of a path the reference would have rejected as a whole.
//
// Per-path buffers (n = npath, cell index fastest):
//   feed   [NC][n]     the path feed normalised once (pr76_pt_continuation.hpp:218-219)
//   hints  [3][NC][n]  carried phase compositions, deduplicated, NaN in unused slots; passed
//                      as both the initial and the final-review stability starts
//   tpz    [3][NC][n]  carried three-phase compositions
//   tpf    [2][n]      carried phase-fraction seed of phases 1, 2 (NaN: no three-phase start)
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

// Warm starts of point k from point k-1's record `prev` (have_prev false at point 0), as
// pr76_pt_continuation_phase_compositions / _three_phase_start (pr76_pt_continuation.hpp:
// 110-161) read them from the reference's result: the accepted phase count comes from the
// max3 status code (single 0, two 1, three 2; anything else accepts nothing); a single phase
// carries the stability test's feed (the path feed normalised again), two phases carry the
// published liquid and vapour compositions (those of the accepted boundary neighbour where
// one was accepted), three phases carry the three compositions and the fractions of phases
// 1 and 2. info = incoming hint phase count | carried start count << 4 | carried
// three-phase start << 8.
template <int NC>
__global__ __launch_bounds__(256) void k_cont_hints(int64_t n, FlashOut prev, int have_prev,
                                                    const double* __restrict__ feed, int carry,
                                                    double* __restrict__ hints, double* __restrict__ tpz,
                                                    double* __restrict__ tpf, int32_t* __restrict__ info) {
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
        const bool carry_comp = (carry & 1) != 0;
        const bool three = (carry & 2) != 0 && count == 3;
        if (carry_comp) {
            // pr76_pt_continuation_append_unique: a composition equal (==, component by
            // component) to an earlier one is not added again
            bool same01 = true, same02 = true, same12 = true;
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                same01 = same01 && v1[i] == v0[i];
                same02 = same02 && v2[i] == v0[i];
                same12 = same12 && v2[i] == v1[i];
            }
            const bool use1 = count >= 2 && !same01;
            const bool use2 = count == 3 && !same02 && !same12;
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                const double h0 = count >= 1 ? v0[i] : nan;
                const double h1 = use1 ? v1[i] : (use2 ? v2[i] : nan);
                const double h2 = use1 && use2 ? v2[i] : nan;
                hints[(0 * NC + i) * n + c] = h0;
                hints[(1 * NC + i) * n + c] = h1;
                hints[(2 * NC + i) * n + c] = h2;
            }
        }
        if ((carry & 2) != 0) {
#pragma unroll
            for (int i = 0; i < NC; ++i) {
                tpz[(0 * NC + i) * n + c] = three ? v0[i] : nan;
                tpz[(1 * NC + i) * n + c] = three ? v1[i] : nan;
                tpz[(2 * NC + i) * n + c] = three ? v2[i] : nan;
            }
            tpf[c] = three ? prev.beta[1 * n + c] : nan;
            tpf[n + c] = three ? prev.beta[2 * n + c] : nan;
        }
        if (info) { info[c] = count | ((carry_comp ? count : 0) << 4) | (three ? 0x100 : 0); }
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
