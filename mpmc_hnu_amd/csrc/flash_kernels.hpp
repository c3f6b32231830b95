// gfx950 flash kernels: one cell per lane, 64-lane workgroups (one wavefront), the
// mixture constants staged in LDS per workgroup. The component count is a
// compile-time template argument so per-cell arrays stay in registers; each
// supported nc is instantiated in its own translation unit (flash_inst_*.hip).
#pragma once

#include "flash_device.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

namespace mpmc::dev {

struct FlashOut {
    int32_t* status;
    int32_t* phase_count;
    double* beta;
    double* zfactor;
    int32_t* branch;
    double* x;
    double* ln_phi;
    uint32_t* counters;
};

template <int NC>
__device__ __forceinline__ void load_mix(const double* __restrict__ src, Mix<NC>& dst) {
    double* d = reinterpret_cast<double*>(&dst);
    constexpr int n = 4 * NC + NC * NC + 1;
    for (int k = threadIdx.x; k < n; k += blockDim.x) { d[k] = src[k]; }
}

template <int NC>
struct PhaseView {
    double beta;
    const double* x;
    const Activity<NC>* act;
    double z;
};

template <int NC>
__device__ void write_cell(const FlashOut& o, int64_t N, int64_t c, int set, int max3, int split, int np,
                           const PhaseView<NC>* ph, const Counters& cnt, int flags) {
    o.status[c] = set | (max3 << 4) | (split << 8) | flags;
    o.phase_count[c] = np;
    const double nan = __builtin_nan("");
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const bool have = k < np;
        o.beta[k * N + c] = have ? ph[k].beta : nan;
        o.zfactor[k * N + c] = have ? ph[k].z : nan;
        o.branch[k * N + c] = have ? (ph[k].act->branch | (ph[k].act->smooth ? 0x100 : 0)) : -1;
#pragma unroll
        for (int i = 0; i < NC; ++i) {
            o.x[(k * NC + i) * N + c] = have ? ph[k].x[i] : nan;
            o.ln_phi[(k * NC + i) * N + c] = have ? ph[k].act->lnphi[i] : nan;
        }
    }
    if (o.counters) {
        o.counters[0 * N + c] = cnt.prop;
        o.counters[1 * N + c] = cnt.tpd;
        o.counters[2 * N + c] = cnt.rr;
        o.counters[3 * N + c] = cnt.rr3;
    }
}

template <int NC>
__device__ __forceinline__ int pair_view(const SplitState<NC>& s, PhaseView<NC>* ph) {
    ph[0] = {1.0 - s.fr.beta, s.fr.x, &s.liq, s.zl};
    ph[1] = {s.fr.beta, s.fr.y, &s.vap, s.zv};
    return 2;
}

// Status-word flags of the base split (MPMC_FLAG_SPLIT_DISAPPEARANCE, MPMC_FLAG_SPLIT_CANDIDATE):
// the facts of PtSplitResult that the reference's transition evidence reads
// (pr76_pt_flash_backend.hpp:42-95).
template <int NC>
__device__ __forceinline__ int split_flags(const VleOut<NC>& V) {
    return (V.any_disappearance ? 0x4000 : 0) | (V.has_candidate ? 0x8000 : 0);
}

// project_pt_vle_phase_set (pt_vle_phase_set.hpp:204-258)
template <int NC>
__device__ void publish_vle(const FlashOut& o, int64_t N, int64_t c, const VleOut<NC>& V, int max3_code,
                            const Counters& cnt, int extra_flags = 0) {
    PhaseView<NC> ph[3];
    const int fl = split_flags<NC>(V) | extra_flags;
    switch (V.status) {
    case V_SINGLE:
        if (!V.init.has_ref) { write_cell<NC>(o, N, c, 2, max3_code, V.status, 0, ph, cnt, fl); return; }
        ph[0] = {1.0, V.init.feed, &V.init.ref, __builtin_nan("")};
        write_cell<NC>(o, N, c, 0, max3_code, V.status, 1, ph, cnt, fl);
        return;
    case V_TWO:
        write_cell<NC>(o, N, c, 0, max3_code, V.status, pair_view<NC>(V.cand, ph), ph, cnt, fl);
        return;
    case V_UNSTABLE:
        write_cell<NC>(o, N, c, 1, max3_code, V.status, V.has_candidate ? pair_view<NC>(V.cand, ph) : 0, ph, cnt, fl);
        return;
    default:
        write_cell<NC>(o, N, c, 2, max3_code, V.status, V.has_candidate ? pair_view<NC>(V.cand, ph) : 0, ph, cnt, fl);
        return;
    }
}

// detail::project_max3_solution (pt_phase_set.hpp:90-150)
template <int NC>
__device__ void publish_max3(const FlashOut& o, int64_t N, int64_t c, const Max3Out<NC>& M, const Counters& cnt) {
    PhaseView<NC> ph[3];
    const int split = M.base.status;
    const int max3f = (M.tail.attempts > 0 ? 0x10000 : 0) | (M.tail.start_consumed ? 0x20000 : 0);
    const int fl = split_flags<NC>(M.base) | max3f;
    switch (M.status) {
    case M_SINGLE:
        publish_vle<NC>(o, N, c, M.base, M.status, cnt, max3f);
        return;
    case M_TWO:
        write_cell<NC>(o, N, c, 0, M.status, split, pair_view<NC>(M.tail.use_neighbor ? M.tail.neighbor : M.base.cand, ph),
                       ph, cnt, fl);
        return;
    case M_THREE:
#pragma unroll
        for (int k = 0; k < 3; ++k) { ph[k] = {M.tail.three.frac[k], M.tail.three.comp[k], &M.tail.three.act[k], M.tail.three.z[k]}; }
        write_cell<NC>(o, N, c, 0, M.status, split, 3, ph, cnt, fl);
        return;
    case M_HIGHER:
        write_cell<NC>(o, N, c, 1, M.status, split, M.base.has_candidate ? pair_view<NC>(M.base.cand, ph) : 0, ph, cnt,
                       fl);
        return;
    default:
        write_cell<NC>(o, N, c, 2, M.status, split, 0, ph, cnt, fl);
        return;
    }
}

// One active lane per wave. The serial searches run through non-inlined calls
// (solve_vle, solve_max3, max3_tail, stability, iterate_three) under divergent control
// flow. With several cells in one wave, the three-phase outcome of a cell on the
// sour-gas sweep came to depend on the other cells of its wave: about 5% of the cells
// that reach the three-phase stage changed outcome with the batch order or with the
// buffers of a previous call, while the same cells flashed one per call (one active lane)
// matched the oracle on every cell, and the host build of the same code is bit-identical
// to the oracle and clean under MemorySanitizer, AddressSanitizer and UBSan. So each
// wave runs one cell at a time; lane 0 does the work after the workgroup has staged the
// mixture, and the grid supplies the parallelism across cells.
template <int NC>
__global__ __launch_bounds__(64) void flash_vle_kernel(const double* __restrict__ mixc, int64_t ncell,
                                                       const double* __restrict__ p, const double* __restrict__ t,
                                                       const double* __restrict__ z, double fug_tol, FlashOut out) {
    __shared__ Mix<NC> smix;
    load_mix<NC>(mixc, smix);
    __syncthreads();
    if (threadIdx.x != 0) { return; }
    for (int64_t c = blockIdx.x; c < ncell; c += gridDim.x) {
        double feed[NC];
#pragma unroll
        for (int i = 0; i < NC; ++i) { feed[i] = z[i * ncell + c]; }
        Counters cnt{0, 0, 0, 0};
        bool rejected = false;
        VleOut<NC> V;
        solve_vle<NC>(smix, p[c], t[c], feed, fug_tol, nullptr, 0, V, cnt, rejected);
        if (rejected) {
            PhaseView<NC> ph[3];
            write_cell<NC>(out, ncell, c, 2, 0xF, 0xF, 0, ph, cnt, 0x2000);
            continue;
        }
        publish_vle<NC>(out, ncell, c, V, 0xF, cnt);
    }
}

template <int NC>
__global__ __launch_bounds__(64) void flash_max3_kernel(const double* __restrict__ mixc, int64_t ncell,
                                                        const double* __restrict__ p, const double* __restrict__ t,
                                                        const double* __restrict__ z, double fug_tol, FlashOut out) {
    __shared__ Mix<NC> smix;
    load_mix<NC>(mixc, smix);
    __syncthreads();
    if (threadIdx.x != 0) { return; }
    for (int64_t c = blockIdx.x; c < ncell; c += gridDim.x) {
        double feed[NC];
#pragma unroll
        for (int i = 0; i < NC; ++i) { feed[i] = z[i * ncell + c]; }
        Counters cnt{0, 0, 0, 0};
        bool rejected = false;
        Max3Out<NC> M;
        solve_max3<NC>(smix, p[c], t[c], feed, fug_tol, M, cnt, rejected);
        if (rejected) {
            PhaseView<NC> ph[3];
            write_cell<NC>(out, ncell, c, 2, 0xF, 0xF, 0, ph, cnt, 0x2000);
            continue;
        }
        publish_max3<NC>(out, ncell, c, M, cnt);
    }
}

template <int NC>
__global__ __launch_bounds__(64) void phase_kernel(const double* __restrict__ mixc, int64_t ncell,
                                                   const double* __restrict__ p, const double* __restrict__ t,
                                                   const double* __restrict__ x, int32_t* root_status,
                                                   int32_t* root_count, double* zroot, int32_t* root_flags,
                                                   double* ln_phi) {
    __shared__ Mix<NC> smix;
    load_mix<NC>(mixc, smix);
    __syncthreads();
    const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= ncell) { return; }
    double w[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) { w[i] = x[i * ncell + c]; }
    const double nan = __builtin_nan("");
    root_count[c] = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        zroot[k * ncell + c] = nan;
        root_flags[k * ncell + c] = 0;
#pragma unroll
        for (int i = 0; i < NC; ++i) { ln_phi[(k * NC + i) * ncell + c] = nan; }
    }
    PhaseCtx<NC> ctx;
    const int e = phase_prepare<NC>(smix, p[c], t[c], w, ctx);
    root_status[c] = e == REJECTED ? -1 : ctx.rs.status;
    if (e != ISSUE_NONE) { return; }
    root_count[c] = ctx.rs.count;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (k >= ctx.rs.count) { break; }
        const Root r = root_at(ctx.rs, k);
        zroot[k * ncell + c] = r.z;
        root_flags[k * ncell + c] = (r.slope_sign + 1) | (r.dvalid ? 0x10 : 0);
        double lp[NC];
        if (phase_lnphi<NC>(smix, p[c], t[c], ctx, k, lp, nullptr) == ISSUE_NONE) {
#pragma unroll
            for (int i = 0; i < NC; ++i) { ln_phi[(k * NC + i) * ncell + c] = lp[i]; }
        }
    }
}

// Batched Rachford-Rice at fixed log K (solve_rachford_rice, rachford_rice.hpp:70-147).
template <int NC>
__global__ __launch_bounds__(64) void rr_kernel(int64_t n, const double* __restrict__ z,
                                                const double* __restrict__ log_k, const double* __restrict__ guess,
                                                int32_t* status, double* beta, double* x, double* y,
                                                int32_t* iterations) {
    const int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= n) { return; }
    double zz[NC], lk[NC];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        zz[i] = z[i * n + c];
        lk[i] = log_k[i * n + c];
    }
    RRState<NC> r;
    r.beta = __builtin_nan("");
    Counters cnt{0, 0, 0, 0};
    bool rejected = false;
    const int st = rachford_rice<NC>(zz, lk, r, cnt, rejected, guess ? guess[c] : 0.5);
    status[c] = rejected ? -2 : st;
    beta[c] = r.beta;
    iterations[c] = static_cast<int32_t>(cnt.rr);
#pragma unroll
    for (int i = 0; i < NC; ++i) {
        x[i * n + c] = st == RR_INTERIOR ? r.x[i] : __builtin_nan("");
        y[i * n + c] = st == RR_INTERIOR ? r.y[i] : __builtin_nan("");
    }
}

} // namespace mpmc::dev


namespace mpmc::dev {

constexpr int kBlock = 64;

// Grid of a one-lane-per-cell kernel with `lane_bytes` of private memory per lane (one_lane:
// one active lane per wave, so one cell per workgroup): every cell's block, capped so that one dispatch's scratch (lane bytes x 64 x resident waves,
// backed on the dispatch's hardware queue) stays within kDispatchScratch; the kernels
// stride over the cells. At least 64 workgroups.
constexpr size_t kDispatchScratch = size_t(256) << 20;
inline unsigned capped_cell_blocks(int64_t ncell, size_t lane_bytes, bool one_lane = false) {
    const int64_t want = one_lane ? ncell : (ncell + kBlock - 1) / kBlock;
    const int64_t cap = lane_bytes == 0 ? want
                                        : std::max<int64_t>(64, static_cast<int64_t>(kDispatchScratch / (lane_bytes * 64)));
    return static_cast<unsigned>(std::max<int64_t>(1, std::min(want, cap)));
}

template <typename K>
inline size_t kernel_lane_scratch(K kernel) {
    hipFuncAttributes a{};
    if (hipFuncGetAttributes(&a, reinterpret_cast<const void*>(kernel)) != hipSuccess) { return 64 * 1024; }
    return a.localSizeBytes;
}

template <int NC>
hipError_t launch_flash(const double* constants, int64_t ncell, const double* p, const double* t, const double* z,
                        int max_phases, double fug_tol, const FlashOut& out, hipStream_t s) {
    static const size_t vle_lane = kernel_lane_scratch(flash_vle_kernel<NC>);
    static const size_t max3_lane = kernel_lane_scratch(flash_max3_kernel<NC>);
    if (max_phases == 2) {
        const dim3 grid(capped_cell_blocks(ncell, vle_lane, true));
        hipLaunchKernelGGL(flash_vle_kernel<NC>, grid, dim3(kBlock), 0, s, constants, ncell, p, t, z, fug_tol, out);
    } else {
        const dim3 grid(capped_cell_blocks(ncell, max3_lane, true));
        hipLaunchKernelGGL(flash_max3_kernel<NC>, grid, dim3(kBlock), 0, s, constants, ncell, p, t, z, fug_tol, out);
    }
    return hipGetLastError();
}

template <int NC>
hipError_t launch_phase(const double* constants, int64_t ncell, const double* p, const double* t, const double* x,
                        int32_t* rs, int32_t* rc, double* zr, int32_t* rf, double* lp, hipStream_t s) {
    const dim3 grid(static_cast<unsigned>((ncell + kBlock - 1) / kBlock));
    hipLaunchKernelGGL(phase_kernel<NC>, grid, dim3(kBlock), 0, s, constants, ncell, p, t, x, rs, rc, zr, rf, lp);
    return hipGetLastError();
}

template <int NC>
hipError_t launch_rr(int64_t n, const double* z, const double* lk, const double* guess, int32_t* st, double* beta,
                     double* x, double* y, int32_t* it, hipStream_t s) {
    const dim3 grid(static_cast<unsigned>((n + kBlock - 1) / kBlock));
    hipLaunchKernelGGL(rr_kernel<NC>, grid, dim3(kBlock), 0, s, n, z, lk, guess, st, beta, x, y, it);
    return hipGetLastError();
}

#define MPMC_FOR_EACH_NC(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12)

#define MPMC_DECLARE_LAUNCH(N)                                                                                   \
    extern template hipError_t launch_flash<N>(const double*, int64_t, const double*, const double*,            \
                                               const double*, int, double, const FlashOut&, hipStream_t);       \
    extern template hipError_t launch_phase<N>(const double*, int64_t, const double*, const double*,            \
                                               const double*, int32_t*, int32_t*, double*, int32_t*, double*,   \
                                               hipStream_t);                                                    \
    extern template hipError_t launch_rr<N>(int64_t, const double*, const double*, const double*, int32_t*,     \
                                            double*, double*, double*, int32_t*, hipStream_t);
#define MPMC_INSTANTIATE_LAUNCH(N)                                                                               \
    template hipError_t launch_flash<N>(const double*, int64_t, const double*, const double*, const double*,     \
                                        int, double, const FlashOut&, hipStream_t);                              \
    template hipError_t launch_phase<N>(const double*, int64_t, const double*, const double*, const double*,     \
                                        int32_t*, int32_t*, double*, int32_t*, double*, hipStream_t);    \
    template hipError_t launch_rr<N>(int64_t, const double*, const double*, const double*, int32_t*, double*,   \
                                     double*, double*, int32_t*, hipStream_t);

} // namespace mpmc::dev
