/*
 * Reference-side drop-in for the PR76 PT flash backend, on the C ABI of this library.
 *
 * GpuPr76PtFlashBackend implements mpmc::flash::PtFlashBackend
 * (modules/flash/include/mpmc/flash/pt_flash_backend.hpp:187-195) with the capability,
 * publication and transition report of Pr76PtFlashBackend
 * (modules/flash/include/mpmc/flash/pr76_pt_flash_backend.hpp:166-282), and runs the flash
 * through mpmc_flash_batch. It is compiled inside the reference's build: it needs the
 * reference's headers (modules/flash/include, modules/thermodynamics/include,
 * modules/ad/include) on the include path, the HIP runtime headers, and links
 * libmpmc_hnu_amd.so and libamdhip64.so.
 *
 *   solve(request)        one request: a batch of one cell
 *   solve_batch(requests) many requests of the backend's mixture in one device call
 *
 * The result of a cell is rebuilt from its SoA flash record (rebuild_result), which is
 * also usable on records produced elsewhere (tests rebuild committed device records on
 * the CPU).
 */
#pragma once

#include "mpmc_hnu_amd.h"

#include <hip/hip_runtime_api.h>

#include <mpmc/flash/pr76_pt_flash_backend.hpp>

#include <cmath>
#include <cstdint>
#include <limits>
#include <span>
#include <stdexcept>
#include <string>
#include <vector>

namespace mpmc::hnu_amd {

// Status-word flags beyond the three enum fields (include/mpmc_hnu_amd.h).
inline constexpr int32_t kFlagRejected = MPMC_FLAG_REJECTED;
inline constexpr int32_t kFlagSplitDisappearance = MPMC_FLAG_SPLIT_DISAPPEARANCE;
inline constexpr int32_t kFlagSplitCandidate = MPMC_FLAG_SPLIT_CANDIDATE;
inline constexpr int32_t kFlagMax3Attempted = MPMC_FLAG_MAX3_ATTEMPTED;

// One cell of a host copy of the SoA flash record (mpmc_flash_outputs layout).
struct CellRecord {
    int32_t status{};
    int32_t phase_count{};
    double beta[3]{};
    double zfactor[3]{};
    int32_t branch[3]{};
    std::vector<double> x[3];
    std::vector<double> ln_phi[3];
};

// Cell c of SoA host arrays laid out as mpmc_flash_outputs (ncell cells, nc components).
inline CellRecord cell_record(int nc, int64_t ncell, int64_t c, const int32_t* status, const int32_t* phase_count,
                              const double* beta, const double* zfactor, const int32_t* branch, const double* x,
                              const double* ln_phi) {
    CellRecord r;
    r.status = status[c];
    r.phase_count = phase_count[c];
    for (int ph = 0; ph < 3; ++ph) {
        r.beta[ph] = beta[ph * ncell + c];
        r.zfactor[ph] = zfactor[ph * ncell + c];
        r.branch[ph] = branch[ph * ncell + c];
        r.x[ph].resize(static_cast<std::size_t>(nc));
        r.ln_phi[ph].resize(static_cast<std::size_t>(nc));
        for (int i = 0; i < nc; ++i) {
            r.x[ph][static_cast<std::size_t>(i)] = x[(ph * nc + i) * ncell + c];
            r.ln_phi[ph][static_cast<std::size_t>(i)] = ln_phi[(ph * nc + i) * ncell + c];
        }
    }
    return r;
}

// The capability Pr76PtFlashBackend::build_capability publishes for a parameter set
// (pr76_pt_flash_backend.hpp:242-268).
inline flash::PtFlashBackendCapability pr76_capability(const thermodynamics::PrParameterSet& parameters) {
    flash::PtFlashBackendCapability capability;
    capability.backend_id = std::string(flash::pr76_pt_flash_backend_id);
    capability.model_profile = std::string(thermodynamics::pr76_profile);
    capability.algorithm_profile = flash::pr76_pt_max3_convention;
    capability.publication_profile = flash::pr76_pt_max3_phase_set_publication_convention;
    capability.configuration_profile = std::string(flash::pr76_pt_flash_backend_configuration_profile);
    capability.dataset_id = parameters.dataset_id();
    capability.revision = parameters.revision();
    for (const auto& component : parameters.components().items()) { capability.component_ids.push_back(component.id); }
    capability.supported_phase_counts = {1U, 2U, 3U};
    using flash::PtPhaseTransitionSupport;
    capability.transition_capability.edges = {{1U, 2U, PtPhaseTransitionSupport::fresh_target_resolve, true},
                                              {2U, 1U, PtPhaseTransitionSupport::detection_only, true},
                                              {2U, 3U, PtPhaseTransitionSupport::fresh_target_resolve, true},
                                              {3U, 2U, PtPhaseTransitionSupport::fresh_target_resolve, true}};
    capability.performs_initial_stability_search = true;
    capability.performs_final_phase_set_review = true;
    capability.performs_boundary_neighbor_resolve = true;
    capability.global_stability_proven = false;
    return capability;
}

// The transition evidence of project_pr76_pt_max3_transition_report
// (pr76_pt_flash_backend.hpp:97-164), from the status word: base split status, max3
// status, the neighbour route (max3 two-phase after an unstable base), the split flags of an
// indeterminate base (a disappearance attempt; a converged candidate) and whether the
// three-phase stage ran any attempt.
inline flash::PtPhaseTransitionReport rebuild_transition_report(int32_t status) {
    using namespace flash;
    PtPhaseTransitionReport report;
    const std::string profile(pr76_pt_transition_evidence_profile);
    const auto add = [&](std::size_t from, std::optional<std::size_t> to, PtPhaseTransitionTrigger trigger,
                         PtPhaseTransitionResolution resolution, bool fresh, bool closed) {
        report.evidence.push_back({from, to, trigger, resolution, fresh, closed, profile, std::string()});
    };
    const auto split = static_cast<PtSplitStatus>((status >> 8) & 0xF);
    const auto max3 = static_cast<Pr76PtMax3Status>((status >> 4) & 0xF);
    if (split == PtSplitStatus::two_phase_no_instability_found) {
        add(1U, 2U, PtPhaseTransitionTrigger::initial_stability_witness, PtPhaseTransitionResolution::accepted_target,
            true, true);
        return report;
    }
    if (split == PtSplitStatus::indeterminate) {
        if (status & kFlagSplitDisappearance) {
            add(2U, 1U, PtPhaseTransitionTrigger::phase_disappearance,
                PtPhaseTransitionResolution::target_resolve_required, false, false);
        }
        if (status & kFlagSplitCandidate) {
            add(1U, 2U, PtPhaseTransitionTrigger::initial_stability_witness,
                PtPhaseTransitionResolution::candidate_not_accepted, true, false);
        }
        return report;
    }
    if (split != PtSplitStatus::phase_set_unstable) { return report; }
    add(1U, 2U, PtPhaseTransitionTrigger::initial_stability_witness, PtPhaseTransitionResolution::candidate_not_accepted,
        true, false);
    switch (max3) {
    case Pr76PtMax3Status::three_phase:
        add(2U, 3U, PtPhaseTransitionTrigger::final_phase_set_instability, PtPhaseTransitionResolution::accepted_target,
            true, true);
        break;
    case Pr76PtMax3Status::two_phase:
        add(2U, 3U, PtPhaseTransitionTrigger::final_phase_set_instability,
            PtPhaseTransitionResolution::candidate_not_accepted, true, false);
        add(3U, 2U, PtPhaseTransitionTrigger::phase_disappearance, PtPhaseTransitionResolution::accepted_target, true,
            true);
        break;
    case Pr76PtMax3Status::phase_boundary_unresolved:
        add(2U, 3U, PtPhaseTransitionTrigger::final_phase_set_instability,
            PtPhaseTransitionResolution::candidate_not_accepted, true, false);
        add(3U, 2U, PtPhaseTransitionTrigger::phase_disappearance, PtPhaseTransitionResolution::target_resolve_failed,
            true, false);
        break;
    case Pr76PtMax3Status::higher_phase_count_or_wrong_candidate:
        add(3U, std::nullopt, PtPhaseTransitionTrigger::final_phase_set_instability,
            PtPhaseTransitionResolution::broader_topology_required, true, false);
        break;
    case Pr76PtMax3Status::indeterminate:
        add(2U, 3U, PtPhaseTransitionTrigger::final_phase_set_instability, PtPhaseTransitionResolution::indeterminate,
            (status & kFlagMax3Attempted) != 0, false);
        break;
    case Pr76PtMax3Status::single_phase:
        break;
    }
    return report;
}

// The published phase set of one cell (project_pr76_pt_max3_phase_set,
// pr76_max3_phase_set.hpp:33-48 and pt_phase_set.hpp:90-150): status, the candidate phases
// in record order (fraction, composition, ln phi, provider branch and smooth flag, Z where
// published), the state and the normalised feed.
inline flash::PtPhaseSetResult rebuild_phase_set(const CellRecord& r, double pressure_pa, double temperature_k,
                                                 std::span<const double> feed) {
    using namespace flash;
    PtPhaseSetResult out;
    out.capability.maximum_phase_count = 3U;
    out.status = static_cast<PtPhaseSetStatus>(r.status & 0xF);
    out.pressure_pa = pressure_pa;
    out.temperature_k = temperature_k;
    const double sum = detail::stability_check_composition(feed);
    out.feed = detail::stability_normalize(feed, sum);
    out.global_stability_proven = false;
    if (r.phase_count > 0) {
        PtCandidatePhaseSet set;
        for (int ph = 0; ph < r.phase_count && ph < 3; ++ph) {
            PtCandidatePhase phase;
            phase.mole_phase_fraction = r.beta[ph];
            phase.composition = r.x[ph];
            phase.activity.ln_phi = r.ln_phi[ph];
            phase.activity.branch = static_cast<std::size_t>(r.branch[ph] & 0xFF);
            phase.activity.smooth = ((r.branch[ph] >> 8) & 1) != 0;
            if (std::isfinite(r.zfactor[ph])) { phase.compressibility_factor = r.zfactor[ph]; }
            set.phases.push_back(std::move(phase));
        }
        out.candidate_phase_set = std::move(set);
    }
    return out;
}

// The backend result of one cell (Pr76PtFlashBackend::solve, pr76_pt_flash_backend.hpp:190-237).
// A cell the flash rejected (MPMC_FLAG_REJECTED) is the reference's exception: the caller
// raises std::domain_error.
inline flash::PtFlashBackendResult rebuild_result(const flash::PtFlashBackendCapability& capability,
                                                  const CellRecord& r, double pressure_pa, double temperature_k,
                                                  std::span<const double> feed) {
    if (r.status & kFlagRejected) {
        throw std::domain_error("GPU PR76 max3 backend: the flash rejected the request's feed or state");
    }
    flash::PtFlashBackendResult result;
    result.capability = capability;
    result.solution = rebuild_phase_set(r, pressure_pa, temperature_k, feed);
    result.transition_report = rebuild_transition_report(r.status);
    result.provider_result_convention = flash::pr76_pt_max3_phase_set_publication_convention;
    result.morphology_resolved = false;
    if (!result.structurally_valid()) {
        result.solution.status = flash::PtPhaseSetStatus::indeterminate;
        result.solution.candidate_phase_set.reset();
        result.phase_metadata.clear();
        result.transition_report.evidence.clear();
        result.solution.diagnostic = "GPU PR76 max3 backend: generic result integrity guard failed";
    }
    return result;
}

// The backend. The mixture is uploaded once from the evaluator's parameter snapshot; each
// solve_batch packs the requests into SoA device buffers (kept and grown across calls), runs
// mpmc_flash_batch on the backend's stream and rebuilds every result from the record.
class GpuPr76PtFlashBackend final : public flash::PtFlashBackend {
public:
    explicit GpuPr76PtFlashBackend(const thermodynamics::PrParameterSet& parameters, double fugacity_tolerance = 1e-11,
                                   int precision = 0)
        : capability_(pr76_capability(parameters)), nc_(static_cast<int>(parameters.components().size())) {
        mpmc_flash_default_options(&options_);
        options_.fugacity_tolerance = fugacity_tolerance;
        options_.max_phases = 3;
        options_.precision = precision;
        check(mpmc_mixture_create(nc_, parameters.critical_temperatures_k().data(),
                                  parameters.critical_pressures_pa().data(), parameters.acentric_factors().data(),
                                  parameters.kij_matrix().data(), &mixture_),
              "mpmc_mixture_create");
    }
    ~GpuPr76PtFlashBackend() override {
        release();
        if (mixture_) { mpmc_mixture_destroy(mixture_); }
    }
    GpuPr76PtFlashBackend(const GpuPr76PtFlashBackend&) = delete;
    GpuPr76PtFlashBackend& operator=(const GpuPr76PtFlashBackend&) = delete;

    [[nodiscard]] const flash::PtFlashBackendCapability& capability() const noexcept override { return capability_; }

    [[nodiscard]] flash::PtFlashBackendResult solve(const flash::PtFlashRequest& request) override {
        return solve_batch(std::span<const flash::PtFlashRequest>(&request, 1)).front();
    }

    [[nodiscard]] std::vector<flash::PtFlashBackendResult> solve_batch(std::span<const flash::PtFlashRequest> requests) {
        const int64_t n = static_cast<int64_t>(requests.size());
        std::vector<flash::PtFlashBackendResult> results;
        if (n == 0) { return results; }
        const std::size_t un = static_cast<std::size_t>(n), unc = static_cast<std::size_t>(nc_);
        std::vector<double> p(un), t(un), z(unc * un);
        for (std::size_t c = 0; c < un; ++c) {
            const auto& q = requests[c];
            if (q.feed.size() != unc) { throw std::invalid_argument("GPU PR76 max3 backend: feed size mismatch"); }
            p[c] = q.pressure_pa;
            t[c] = q.temperature_k;
            for (std::size_t i = 0; i < unc; ++i) { z[i * un + c] = q.feed[i]; }
        }
        reserve(n);
        const std::size_t d8 = sizeof(double), i4 = sizeof(int32_t);
        hip(hipMemcpyAsync(dp_, p.data(), un * d8, hipMemcpyHostToDevice, stream_));
        hip(hipMemcpyAsync(dt_, t.data(), un * d8, hipMemcpyHostToDevice, stream_));
        hip(hipMemcpyAsync(dz_, z.data(), unc * un * d8, hipMemcpyHostToDevice, stream_));
        mpmc_flash_outputs out{ds_, dn_, db_, dzf_, dbr_, dx_, dl_, nullptr};
        check(mpmc_flash_batch(mixture_, n, dp_, dt_, dz_, &options_, &out, stream_), "mpmc_flash_batch");
        std::vector<int32_t> st(un), np(un), br(3 * un);
        std::vector<double> be(3 * un), zf(3 * un), x(3 * unc * un), lp(3 * unc * un);
        hip(hipMemcpyAsync(st.data(), ds_, un * i4, hipMemcpyDeviceToHost, stream_));
        hip(hipMemcpyAsync(np.data(), dn_, un * i4, hipMemcpyDeviceToHost, stream_));
        hip(hipMemcpyAsync(br.data(), dbr_, 3 * un * i4, hipMemcpyDeviceToHost, stream_));
        hip(hipMemcpyAsync(be.data(), db_, 3 * un * d8, hipMemcpyDeviceToHost, stream_));
        hip(hipMemcpyAsync(zf.data(), dzf_, 3 * un * d8, hipMemcpyDeviceToHost, stream_));
        hip(hipMemcpyAsync(x.data(), dx_, 3 * unc * un * d8, hipMemcpyDeviceToHost, stream_));
        hip(hipMemcpyAsync(lp.data(), dl_, 3 * unc * un * d8, hipMemcpyDeviceToHost, stream_));
        hip(hipStreamSynchronize(stream_));
        results.reserve(un);
        for (int64_t c = 0; c < n; ++c) {
            const CellRecord r = cell_record(nc_, n, c, st.data(), np.data(), be.data(), zf.data(), br.data(), x.data(),
                                             lp.data());
            const auto& q = requests[static_cast<std::size_t>(c)];
            results.push_back(rebuild_result(capability_, r, q.pressure_pa, q.temperature_k, q.feed));
        }
        return results;
    }

private:
    static void check(int rc, const char* what) {
        if (rc != MPMC_OK) {
            const char* detail = mpmc_last_error();
            throw std::runtime_error(std::string("GPU PR76 max3 backend: ") + what + " failed" +
                                     (detail ? std::string(": ") + detail : std::string()));
        }
    }
    static void hip(hipError_t e) {
        if (e != hipSuccess) { throw std::runtime_error(std::string("GPU PR76 max3 backend: ") + hipGetErrorString(e)); }
    }
    void release() {
        for (void* q : {static_cast<void*>(dp_), static_cast<void*>(dt_), static_cast<void*>(dz_),
                        static_cast<void*>(ds_), static_cast<void*>(dn_), static_cast<void*>(db_),
                        static_cast<void*>(dzf_), static_cast<void*>(dbr_), static_cast<void*>(dx_),
                        static_cast<void*>(dl_)}) {
            if (q) { (void)hipFree(q); }
        }
        dp_ = dt_ = dz_ = db_ = dzf_ = dx_ = dl_ = nullptr;
        ds_ = dn_ = dbr_ = nullptr;
        cap_ = 0;
    }
    template <typename T>
    void alloc(T*& ptr, std::size_t count) {
        hip(hipMalloc(reinterpret_cast<void**>(&ptr), count * sizeof(T)));
    }
    void reserve(int64_t n) {
        if (n <= cap_) { return; }
        release();
        const std::size_t un = static_cast<std::size_t>(n), unc = static_cast<std::size_t>(nc_);
        alloc(dp_, un);
        alloc(dt_, un);
        alloc(dz_, unc * un);
        alloc(ds_, un);
        alloc(dn_, un);
        alloc(db_, 3 * un);
        alloc(dzf_, 3 * un);
        alloc(dbr_, 3 * un);
        alloc(dx_, 3 * unc * un);
        alloc(dl_, 3 * unc * un);
        cap_ = n;
    }

    flash::PtFlashBackendCapability capability_;
    int nc_;
    mpmc_mixture_t mixture_{nullptr};
    mpmc_flash_options options_{};
    hipStream_t stream_{nullptr};
    int64_t cap_{0};
    double *dp_{nullptr}, *dt_{nullptr}, *dz_{nullptr}, *db_{nullptr}, *dzf_{nullptr}, *dx_{nullptr}, *dl_{nullptr};
    int32_t *ds_{nullptr}, *dn_{nullptr}, *dbr_{nullptr};
};

} // namespace mpmc::hnu_amd
