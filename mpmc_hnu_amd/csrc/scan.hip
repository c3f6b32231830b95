// Post-SNES phase-transition scan: the simulator-side batch caller of the flash
// (modules/flow/discretization/petsc/include/mpmc/flow_discretization_petsc/
//  post_snes_pt_flash_phase_transition_scanner.hpp).
//
// The reference walks the owned cells one by one (scan_post_snes_pt_flash_phase_transitions_3d,
// :303-423): it rebuilds each cell's PT feed from its current component inventory
// (make_source_snapshot, :160-200), solves one flash per cell through the backend and
// compares the accepted phase count with the cell's current one
// (scan_post_snes_pt_flash_source_cell_3d, :203-301). Here the whole batch is three
// cell-parallel steps on one stream: source (feed and validation), the staged flash of
// mpmc_flash_batch, decision. The reference stops at the first cell that is indeterminate
// or raises an error; the batch computes every cell and reduces the lowest such cell
// index, so the caller sees the same outcome (the summary) plus every cell's code.
#include "../../include/mpmc_hnu_amd.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace mpmc::api {
// The library's error record (mpmc_last_error), flash_api.hip.
int record_error(int code, const std::string& what);
} // namespace mpmc::api

namespace {

// Universal gas constant of the PR76 provider (the mixture's r_gas).
constexpr double kGasConstant = static_cast<double>(8.31446261815324L);
constexpr double kEps = 2.220446049250313e-16;

__device__ __forceinline__ bool finite(double v) { return __builtin_isfinite(v); }

// Sequential compensated sum (hi + lo) standing in for the reference's long double
// accumulation.
__device__ __forceinline__ void dd_add(double v, double& hi, double& lo) {
    const double s = hi + v;
    const double bb = s - hi;
    lo += (hi - (s - bb)) + (v - bb);
    hi = s;
}

// validate_source_snapshot (:76-115): positive pressure and temperature, a phase count in
// [1, 3], strictly positive finite fractions whose sum is within 4096 eps * nc of one.
__device__ bool source_valid(int nc, double p, double t, int32_t count, const double* z, int64_t n, int64_t c) {
    if (count < 1 || count > MPMC_MAX_PHASES || !finite(p) || !(p > 0) || !finite(t) || !(t > 0)) { return false; }
    double hi = 0, lo = 0;
    for (int i = 0; i < nc; ++i) {
        const double v = z[i * n + c];
        if (!finite(v) || !(v > 0)) { return false; }
        dd_add(v, hi, lo);
    }
    const double dev = (hi - 1.0) + lo;
    return finite(hi) && __builtin_fabs(dev) <= 4096.0 * kEps * static_cast<double>(nc);
}

// Source of a caller-given batch: a malformed source, or a phase count the backend does
// not support (max_phases), raises invalid_argument / returns PETSC_ERR_ARG_INCOMP in the
// reference's per-cell scan (:217-227), code 3.
__global__ void k_scan_source(int nc, int64_t n, const double* __restrict__ p, const double* __restrict__ t,
                              const double* __restrict__ z, const int32_t* __restrict__ count, int max_phases,
                              int32_t* __restrict__ code) {
    for (int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < n;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const bool ok = source_valid(nc, p[c], t[c], count[c], z, n, c) && count[c] <= max_phases;
        code[c] = ok ? MPMC_SCAN_SAME : MPMC_SCAN_ERROR;
    }
}

// make_source_snapshot (:160-200) of a two-phase cell: z_i = N_i / sum N from the current
// pore-volume inventory, at the cell's pressure q[0] and temperature q[1]. The reference
// divides in long double; here the quotient by the compensated total, corrected once by
// its fma residual. A cell whose state evaluation failed (code already 1 from the
// inventory pass), whose inventory is not finite and positive, or whose snapshot fails
// validation is indeterminate (the reference catches these and reports indeterminate),
// code 2; its feed is NaN.
__global__ void k_scan_source_2p(int nc, int64_t n, const double* __restrict__ q, const double* __restrict__ inv,
                                 double* __restrict__ feed, int32_t* __restrict__ code) {
    for (int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < n;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        bool ok = code[c] == 0;
        double hi = 0, lo = 0;
        for (int i = 0; i < nc; ++i) {
            const double v = inv[i * n + c];
            ok = ok && finite(v) && v > 0;
            dd_add(v, hi, lo);
        }
        ok = ok && finite(hi) && hi > 0;
        for (int i = 0; i < nc; ++i) {
            double zi = __builtin_nan("");
            if (ok) {
                const double v = inv[i * n + c];
                const double q0 = v / hi;
                const double r = __builtin_fma(-q0, hi, v) - q0 * lo;
                zi = q0 + r / hi;
            }
            feed[i * n + c] = zi;
        }
        ok = ok && source_valid(nc, q[c], q[n + c], 2, feed, n, c);
        code[c] = ok ? MPMC_SCAN_SAME : MPMC_SCAN_INDETERMINATE;
    }
}

// Decision (:236-292) on the flash record. A cell already marked (2, 3) keeps its code.
// The flash rejecting the feed of a validated source is the std::domain_error of the
// reference's stability and mixing checks, which the scanner catches as indeterminate
// (:292-299); no accepted phase set is indeterminate; an accepted set of the source's size
// is no transition. Otherwise the target phase molar densities are p / (Z R T) of the
// accepted phases, the density resolver of the reference's scanner test
// (post_snes_pt_flash_phase_transition_scanner_test.cpp:26-57): a density that is not
// finite and positive leaves the cell indeterminate. With the densities resolved, the
// candidate must carry the backend's transition evidence
// (make_phase_set_transition_candidate_from_flash, phase_set_transition_flash_adapter.hpp:
// 171-205, over project_pr76_pt_max3_transition_report, pr76_pt_flash_backend.hpp:94-174):
// a higher target always does (the base split or the three-phase solve challenged the
// source and resolved the target); a lower target only through a fresh phase-disappearance
// solve, which the PR76 backend reports for 3 -> 2 alone, when the three-phase tail
// resolved a two-phase neighbour (max3 two-phase after an unstable base split). Any other
// lower target is the adapter's std::invalid_argument: an error.
// zroot (density mode 1, else null): Z of every root of the cubic at the cell's feed
// (mpmc_phase_batch); a single accepted phase, whose record publishes no Z (as the
// reference's single-phase projection, pt_vle_phase_set.hpp:68-73), then takes the Z of
// its provider root (branch word of phase 0), and meets the evidence rule as any lower
// target does.
// stop[0] / stop[1]: lowest index of an indeterminate / error cell.
__global__ void k_scan_decide(int64_t n, const double* __restrict__ p, const double* __restrict__ t,
                              const int32_t* __restrict__ count, int32_t count_all,
                              const int32_t* __restrict__ status, const int32_t* __restrict__ phase_count,
                              const double* __restrict__ zfactor, const int32_t* __restrict__ branch,
                              const double* __restrict__ zroot, int32_t* __restrict__ code,
                              int32_t* __restrict__ target_count, double* __restrict__ density,
                              unsigned long long* __restrict__ stop) {
    for (int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < n;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int32_t src = count ? count[c] : count_all;
        int32_t k = code[c];
        int32_t target = 0;
        double rho[MPMC_MAX_PHASES];
        for (int ph = 0; ph < MPMC_MAX_PHASES; ++ph) { rho[ph] = __builtin_nan(""); }
        if (k == MPMC_SCAN_SAME) {
            const int32_t st = status[c];
            if (st & MPMC_FLAG_REJECTED) {
                k = MPMC_SCAN_INDETERMINATE;
            } else if (MPMC_STATUS_SET(st) != MPMC_SET_ACCEPTED) {
                k = MPMC_SCAN_INDETERMINATE;
            } else {
                target = phase_count[c];
                if (target != src) {
                    k = MPMC_SCAN_PROPOSAL;
                    for (int ph = 0; ph < MPMC_MAX_PHASES; ++ph) {
                        if (ph >= target) { continue; }
                        double zf = zfactor[ph * n + c];
                        if (zroot && target == 1 && !finite(zf)) {
                            const int root = MPMC_BRANCH_INDEX(branch[c]);
                            zf = root >= 0 && root < 3 ? zroot[root * n + c] : __builtin_nan("");
                        }
                        const double den = zf * kGasConstant * t[c];
                        const double r = p[c] / den;
                        if (!finite(zf) || !(zf > 0) || !finite(r) || !(r > 0)) { k = MPMC_SCAN_INDETERMINATE; }
                        rho[ph] = r;
                    }
                    const bool neighbour = src == 3 && target == 2 && MPMC_STATUS_MAX3(st) == MPMC_MAX3_TWO &&
                                           MPMC_STATUS_SPLIT(st) == MPMC_SPLIT_UNSTABLE;
                    if (k == MPMC_SCAN_PROPOSAL && target < src && !neighbour) { k = MPMC_SCAN_ERROR; }
                }
            }
        }
        code[c] = k;
        target_count[c] = target;
        for (int ph = 0; ph < MPMC_MAX_PHASES; ++ph) {
            density[ph * n + c] = k == MPMC_SCAN_PROPOSAL ? rho[ph] : __builtin_nan("");
        }
        // one atomic per wave: a wave's cells are consecutive in lane order, so its lowest
        // flagged lane holds its lowest flagged cell
        const int lane = static_cast<int>(__lane_id());
        const unsigned long long indet = __ballot(k == MPMC_SCAN_INDETERMINATE);
        const unsigned long long err = __ballot(k == MPMC_SCAN_ERROR);
        if (indet && lane == __builtin_ctzll(indet)) { atomicMin(&stop[0], static_cast<unsigned long long>(c)); }
        if (err && lane == __builtin_ctzll(err)) { atomicMin(&stop[1], static_cast<unsigned long long>(c)); }
    }
}

// Proposals the reference's cell loop has collected when it stops: those before the first
// indeterminate or error cell (:366-407), counted one atomic per wave into stop[2].
__global__ void k_scan_count(int64_t n, const int32_t* __restrict__ code, unsigned long long* __restrict__ stop) {
    const unsigned long long first = stop[0] < stop[1] ? stop[0] : stop[1];
    for (int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < n;
         c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const bool prop = code[c] == MPMC_SCAN_PROPOSAL && static_cast<unsigned long long>(c) < first;
        const unsigned long long b = __ballot(prop);
        if (b && static_cast<int>(__lane_id()) == __builtin_ctzll(b)) {
            atomicAdd(&stop[2], static_cast<unsigned long long>(__popcll(b)));
        }
    }
}

__global__ void k_scan_init(unsigned long long* stop) {
    if (threadIdx.x < 2) { stop[threadIdx.x] = ~0ull; }
    if (threadIdx.x == 2) { stop[2] = 0; }
}

int scan_fail(int code, const std::string& what) { return mpmc::api::record_error(code, what); }

unsigned grid_for(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return static_cast<unsigned>(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

// Flash, decision and the summary of the reference's stop-at-first rule.
int scan_finish(mpmc_mixture_t mixture, int64_t n, const double* p, const double* t, const double* z,
                const int32_t* count, int32_t count_all, const mpmc_flash_options* options,
                const mpmc_flash_outputs* out, const mpmc_scan_outputs* scan, mpmc_scan_summary* summary,
                int density_mode, hipStream_t s) {
    int rc = mpmc_flash_batch(mixture, n, p, t, z, options, out, s);
    if (rc != MPMC_OK) { return rc; }
    // density mode 1: every root of the cubic at the cell's feed (root status and count, Z
    // and flags of 3 roots, ln phi of 3 roots), in one device buffer
    char* phase = nullptr;
    const double* zroot = nullptr;
    if (density_mode == 1 && n > 0) {
        const int nc = mpmc_mixture_components(mixture);
        if (nc < 2) { return scan_fail(MPMC_E_INVALID_ARGUMENT, "transition_scan: bad mixture"); }
        const size_t bytes = static_cast<size_t>(n) * (2 * 4 + 3 * 8 + 3 * 4 + 3 * 8 * static_cast<size_t>(nc)) + 16;
        if (hipMalloc(reinterpret_cast<void**>(&phase), bytes) != hipSuccess) {
            return scan_fail(MPMC_E_NO_MEMORY, "transition_scan: device allocation failed");
        }
        auto* rs = reinterpret_cast<int32_t*>(phase);
        auto* rcnt = rs + n;
        auto* zr = reinterpret_cast<double*>(rcnt + n);  // 2n int32: 8-byte aligned
        auto* rf = reinterpret_cast<int32_t*>(zr + 3 * n);
        auto* lp = reinterpret_cast<double*>(rf + 3 * n + ((3 * n) & 1));
        rc = mpmc_phase_batch(mixture, n, p, t, z, rs, rcnt, zr, rf, lp, s);
        if (rc != MPMC_OK) {
            (void)hipStreamSynchronize(s);
            (void)hipFree(phase);
            return rc;
        }
        zroot = zr;
    }
    unsigned long long* stop = nullptr;
    hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&stop), 3 * sizeof(unsigned long long), s);
    if (e != hipSuccess) {
        if (phase) {
            (void)hipStreamSynchronize(s);
            (void)hipFree(phase);
        }
        return scan_fail(MPMC_E_NO_MEMORY, "transition_scan: device allocation failed");
    }
    unsigned long long host[3] = {~0ull, ~0ull, 0};
    hipLaunchKernelGGL(k_scan_init, dim3(1), dim3(64), 0, s, stop);
    hipLaunchKernelGGL(k_scan_decide, dim3(grid_for(n)), dim3(256), 0, s, n, p, t, count, count_all, out->status,
                       out->phase_count, out->zfactor, out->branch, zroot, scan->code, scan->target_count,
                       scan->target_density, stop);
    hipLaunchKernelGGL(k_scan_count, dim3(grid_for(n)), dim3(256), 0, s, n, scan->code, stop);
    e = hipGetLastError();
    if (e == hipSuccess) { e = hipMemcpyAsync(host, stop, sizeof(host), hipMemcpyDeviceToHost, s); }
    const hipError_t ef = hipFreeAsync(stop, s);
    if (e == hipSuccess) { e = ef; }
    const hipError_t es = hipStreamSynchronize(s);
    if (e == hipSuccess) { e = es; }
    if (phase) {
        const hipError_t ep = hipFree(phase);
        if (e == hipSuccess) { e = ep; }
    }
    if (e != hipSuccess) { return scan_fail(MPMC_E_DEVICE, std::string("transition_scan: ") + hipGetErrorString(e)); }
    // the reference returns at the first indeterminate cell (status indeterminate, the
    // collected proposals cleared) or at the first cell that raises an error (the error
    // code, status untouched, the proposals before that cell kept)
    const unsigned long long first = host[0] < host[1] ? host[0] : host[1];
    summary->first_stop = first == ~0ull ? -1 : static_cast<int64_t>(first);
    summary->status = host[0] < host[1] ? MPMC_SCAN_STATUS_INDETERMINATE : MPMC_SCAN_COMPLETE;
    summary->proposals = summary->status == MPMC_SCAN_STATUS_INDETERMINATE ? 0 : static_cast<int64_t>(host[2]);
    if (host[1] < host[0]) {
        return scan_fail(MPMC_E_INVALID_ARGUMENT,
                         "transition_scan: the first stopping cell has a malformed source or a rejected feed");
    }
    return MPMC_OK;
}

int scan_check(mpmc_mixture_t mixture, int64_t n, const mpmc_flash_outputs* out, const mpmc_scan_outputs* scan,
               const mpmc_scan_summary* summary, int density_mode) {
    if (!mixture || !out || !scan || !summary || n < 0) {
        return scan_fail(MPMC_E_INVALID_ARGUMENT, "transition_scan: null argument or negative cell count");
    }
    if (density_mode != 0 && density_mode != 1) {
        return scan_fail(MPMC_E_INVALID_ARGUMENT, "transition_scan: density_mode must be 0 or 1");
    }
    if (mpmc_mixture_components(mixture) < 2) {
        return scan_fail(MPMC_E_INVALID_ARGUMENT, "transition_scan: a PT source needs at least two components");
    }
    if (n > 0 && (!scan->code || !scan->target_count || !scan->target_density || !out->status ||
                  !out->phase_count || !out->zfactor || !out->branch)) {
        return scan_fail(MPMC_E_INVALID_ARGUMENT, "transition_scan: null output buffer");
    }
    return MPMC_OK;
}

} // namespace

extern "C" {

int mpmc_transition_scan(mpmc_mixture_t mixture, int64_t ncell, const double* p, const double* t, const double* z,
                         const int32_t* source_phase_count, const mpmc_flash_options* options,
                         int density_mode, const mpmc_flash_outputs* out, const mpmc_scan_outputs* scan,
                         mpmc_scan_summary* summary, void* stream) {
    int rc = scan_check(mixture, ncell, out, scan, summary, density_mode);
    if (rc != MPMC_OK) { return rc; }
    *summary = mpmc_scan_summary{MPMC_SCAN_COMPLETE, -1, 0};
    if (ncell == 0) { return MPMC_OK; }
    if (!p || !t || !z || !source_phase_count) { return scan_fail(MPMC_E_INVALID_ARGUMENT, "transition_scan: null input"); }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const int max_phases = options ? options->max_phases : MPMC_MAX_PHASES;
    hipLaunchKernelGGL(k_scan_source, dim3(grid_for(ncell)), dim3(256), 0, s, mpmc_mixture_components(mixture), ncell,
                       p, t, z, source_phase_count, max_phases, scan->code);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { return scan_fail(MPMC_E_DEVICE, std::string("transition_scan: ") + hipGetErrorString(e)); }
    return scan_finish(mixture, ncell, p, t, z, source_phase_count, 0, options, out, scan, summary, density_mode, s);
}

int mpmc_transition_scan_2p(mpmc_mixture_t mixture, mpmc_tpfa_grid_t grid, const mpmc_assembly_model* model,
                            const double* q, const int32_t* branch, double* inventory, double* feed,
                            const mpmc_flash_options* options, int density_mode, const mpmc_flash_outputs* out,
                            const mpmc_scan_outputs* scan, mpmc_scan_summary* summary, void* stream) {
    int64_t ncell = 0;
    int rc = mpmc_tpfa_grid_cells(grid, &ncell);
    if (rc != MPMC_OK) { return rc; }
    rc = scan_check(mixture, ncell, out, scan, summary, density_mode);
    if (rc != MPMC_OK) { return rc; }
    *summary = mpmc_scan_summary{MPMC_SCAN_COMPLETE, -1, 0};
    if (ncell == 0) { return MPMC_OK; }
    if (!q || !inventory || !feed) { return scan_fail(MPMC_E_INVALID_ARGUMENT, "transition_scan_2p: null buffer"); }
    const hipStream_t s = static_cast<hipStream_t>(stream);
    // state evaluation per cell: the inventory pass writes 0 or 1 (domain error) per cell
    rc = mpmc_inventory_2p(mixture, grid, model, q, branch, scan->code, inventory, nullptr, stream);
    if (rc != MPMC_OK) { return rc; }
    const int nc = mpmc_mixture_components(mixture);
    hipLaunchKernelGGL(k_scan_source_2p, dim3(grid_for(ncell)), dim3(256), 0, s, nc, ncell, q, inventory, feed,
                       scan->code);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) { return scan_fail(MPMC_E_DEVICE, std::string("transition_scan_2p: ") + hipGetErrorString(e)); }
    return scan_finish(mixture, ncell, q, q + ncell, feed, nullptr, 2, options, out, scan, summary, density_mode, s);
}

} // extern "C"
