// Test infrastructure: checks the reference-side drop-in (include/mpmc_pr76_backend.hpp)
// against the reference's own Pr76PtFlashBackend on identical requests. Built by
// oracle/Makefile (target adapter) from the reference's headers; the product never links it.
//
//   adapter_check rebuild <cells.bin>   results rebuilt from the device flash records stored
//                                       in the file (no GPU), against the reference backend
//   adapter_check solve <cells.bin>     GpuPr76PtFlashBackend::solve_batch on the GPU, against
//                                       the reference backend run on the host
//
// cells.bin (little endian): int32 nc, int64 ncell, double tc[nc], pc[nc], omega[nc],
// kij[nc*nc], p[ncell], t[ncell], z[nc][ncell]; then, for rebuild, int32 status[ncell],
// int32 phase_count[ncell], double beta[3][ncell], zfactor[3][ncell], int32 branch[3][ncell],
// double x[3][nc][ncell], ln_phi[3][nc][ncell] (the mpmc_flash_outputs layout).
//
// A result matches when: both are structurally valid; the phase-set status, accepted phase
// count and candidate phase count are equal; per phase, the provider branch and smooth flag
// are equal, Z is published by both or neither, and fraction, composition and Z agree within
// 1e-9 relative and ln(phi) within 1e-9 absolute plus relative; the transition evidence
// agrees entry by entry (source, target, trigger, resolution, fresh attempt, closure); the
// capability and result conventions are equal; a rejected request is rejected by both.
#include "../include/mpmc_pr76_backend.hpp"

#include "ref_params.hpp"

#include <mpmc/flash/pr76_pt_flash_backend.hpp>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <optional>
#include <string>
#include <vector>

namespace fl = mpmc::flash;
namespace th = mpmc::thermodynamics;
namespace hn = mpmc::hnu_amd;

namespace {

struct Cells {
    int nc = 0;
    int64_t n = 0;
    std::vector<double> tc, pc, om, kij, p, t, z;
    std::vector<int32_t> status, count, branch;
    std::vector<double> beta, zf, x, lp;
    bool has_record = false;
};

template <typename T>
bool get(std::ifstream& f, std::vector<T>& v, std::size_t n) {
    v.resize(n);
    return static_cast<bool>(f.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(n * sizeof(T))));
}

bool load(const char* path, Cells& c, bool record) {
    std::ifstream f(path, std::ios::binary);
    if (!f.read(reinterpret_cast<char*>(&c.nc), 4) || !f.read(reinterpret_cast<char*>(&c.n), 8)) { return false; }
    const std::size_t nc = static_cast<std::size_t>(c.nc), n = static_cast<std::size_t>(c.n);
    bool ok = get(f, c.tc, nc) && get(f, c.pc, nc) && get(f, c.om, nc) && get(f, c.kij, nc * nc) && get(f, c.p, n) &&
              get(f, c.t, n) && get(f, c.z, nc * n);
    if (ok && record) {
        ok = get(f, c.status, n) && get(f, c.count, n) && get(f, c.beta, 3 * n) && get(f, c.zf, 3 * n) &&
             get(f, c.branch, 3 * n) && get(f, c.x, 3 * nc * n) && get(f, c.lp, 3 * nc * n);
        c.has_record = ok;
    }
    return ok;
}

bool close(double a, double b, double rel) {
    return std::fabs(a - b) <= rel * std::fmax(std::fabs(a), std::fabs(b));
}

// Empty string when the two results match (see the file header), else the first difference.
std::string compare(const fl::PtFlashBackendResult& ref, const fl::PtFlashBackendResult& gpu) {
    if (!ref.structurally_valid()) { return "reference result not structurally valid"; }
    if (!gpu.structurally_valid()) { return "rebuilt result not structurally valid"; }
    const auto& cr = ref.capability;
    const auto& cg = gpu.capability;
    if (cr.backend_id != cg.backend_id || cr.algorithm_profile != cg.algorithm_profile ||
        cr.publication_profile != cg.publication_profile || cr.configuration_profile != cg.configuration_profile ||
        cr.dataset_id != cg.dataset_id || cr.revision != cg.revision || cr.component_ids != cg.component_ids ||
        cr.supported_phase_counts != cg.supported_phase_counts ||
        cr.transition_capability.edges.size() != cg.transition_capability.edges.size()) {
        return "capability";
    }
    if (ref.provider_result_convention != gpu.provider_result_convention) { return "result convention"; }
    const auto& sr = ref.solution;
    const auto& sg = gpu.solution;
    if (sr.status != sg.status) { return "phase-set status"; }
    if (sr.accepted_phase_count() != sg.accepted_phase_count()) { return "accepted phase count"; }
    if (sr.capability.maximum_phase_count != sg.capability.maximum_phase_count) { return "maximum phase count"; }
    if (sr.pressure_pa != sg.pressure_pa || sr.temperature_k != sg.temperature_k) { return "state"; }
    if (sr.feed != sg.feed) { return "normalised feed"; }
    if (sr.candidate_phase_set.has_value() != sg.candidate_phase_set.has_value()) { return "candidate presence"; }
    if (sr.candidate_phase_set) {
        const auto& pr = sr.candidate_phase_set->phases;
        const auto& pg = sg.candidate_phase_set->phases;
        if (pr.size() != pg.size()) { return "candidate phase count"; }
        for (std::size_t k = 0; k < pr.size(); ++k) {
            if (pr[k].activity.branch != pg[k].activity.branch || pr[k].activity.smooth != pg[k].activity.smooth) {
                return "branch label of phase " + std::to_string(k);
            }
            if (!close(pr[k].mole_phase_fraction, pg[k].mole_phase_fraction, 1e-9)) {
                return "fraction of phase " + std::to_string(k);
            }
            if (pr[k].compressibility_factor.has_value() != pg[k].compressibility_factor.has_value() ||
                (pr[k].compressibility_factor &&
                 !close(*pr[k].compressibility_factor, *pg[k].compressibility_factor, 1e-9))) {
                return "Z of phase " + std::to_string(k);
            }
            for (std::size_t i = 0; i < pr[k].composition.size(); ++i) {
                if (!close(pr[k].composition[i], pg[k].composition[i], 1e-9)) {
                    return "composition of phase " + std::to_string(k);
                }
                const double a = pr[k].activity.ln_phi[i], b = pg[k].activity.ln_phi[i];
                if (std::fabs(a - b) > 1e-9 * (1.0 + std::fabs(a))) { return "ln phi of phase " + std::to_string(k); }
            }
        }
    }
    const auto& er = ref.transition_report.evidence;
    const auto& eg = gpu.transition_report.evidence;
    if (er.size() != eg.size()) { return "evidence count"; }
    for (std::size_t k = 0; k < er.size(); ++k) {
        if (er[k].source_phase_count != eg[k].source_phase_count ||
            er[k].target_phase_count != eg[k].target_phase_count || er[k].trigger != eg[k].trigger ||
            er[k].resolution != eg[k].resolution ||
            er[k].fresh_target_solve_attempted != eg[k].fresh_target_solve_attempted ||
            er[k].target_topology_closed != eg[k].target_topology_closed ||
            er[k].provider_evidence_profile != eg[k].provider_evidence_profile) {
            return "evidence entry " + std::to_string(k);
        }
    }
    return {};
}

} // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: adapter_check rebuild|solve cells.bin\n");
        return 2;
    }
    const bool rebuild = std::strcmp(argv[1], "rebuild") == 0;
    Cells c;
    if (!load(argv[2], c, rebuild)) {
        std::fprintf(stderr, "cannot read %s\n", argv[2]);
        return 2;
    }
    const auto parameters = mpmc_oracle::make_parameters(c.nc, c.tc.data(), c.pc.data(), c.om.data(), c.kij.data());
    const auto model = th::Pr76Phase<double>::from_parameters(parameters);
    fl::Pr76VleEvaluator evaluator(model);
    fl::Pr76PtFlashBackendOptions options;
    options.split.iteration.fugacity_tolerance = 1e-11;
    fl::Pr76PtFlashBackend reference(evaluator, options);
    const std::size_t nc = static_cast<std::size_t>(c.nc), n = static_cast<std::size_t>(c.n);
    std::vector<fl::PtFlashRequest> requests(n);
    for (std::size_t k = 0; k < n; ++k) {
        requests[k].pressure_pa = c.p[k];
        requests[k].temperature_k = c.t[k];
        requests[k].feed.resize(nc);
        for (std::size_t i = 0; i < nc; ++i) { requests[k].feed[i] = c.z[i * n + k]; }
    }
    std::vector<std::optional<fl::PtFlashBackendResult>> gpu(n);
    std::vector<bool> gpu_rejected(n, false);
    const auto capability = hn::pr76_capability(parameters);
    if (rebuild) {
        if (!c.has_record) {
            std::fprintf(stderr, "no record in %s\n", argv[2]);
            return 2;
        }
        for (std::size_t k = 0; k < n; ++k) {
            const auto rec = hn::cell_record(c.nc, c.n, static_cast<int64_t>(k), c.status.data(), c.count.data(),
                                             c.beta.data(), c.zf.data(), c.branch.data(), c.x.data(), c.lp.data());
            try {
                gpu[k] = hn::rebuild_result(capability, rec, c.p[k], c.t[k], requests[k].feed);
            } catch (const std::domain_error&) { gpu_rejected[k] = true; }
        }
    } else {
        hn::GpuPr76PtFlashBackend backend(parameters, 1e-11, 0);
        std::vector<fl::PtFlashRequest> ok;
        const auto results = backend.solve_batch(requests);
        for (std::size_t k = 0; k < n; ++k) { gpu[k] = results[k]; }
    }
    int mismatches = 0, valid = 0, accepted = 0, three = 0, rejected = 0;
    for (std::size_t k = 0; k < n; ++k) {
        std::optional<fl::PtFlashBackendResult> ref;
        try { ref = reference.solve(requests[k]); } catch (const std::exception&) {}
        if (!ref || gpu_rejected[k]) {
            const bool same = !ref && gpu_rejected[k];
            rejected += same;
            if (!same) {
                ++mismatches;
                if (mismatches <= 10) { std::printf("cell %zu: rejected by only one side\n", k); }
            }
            continue;
        }
        const std::string d = compare(*ref, *gpu[k]);
        valid += gpu[k]->structurally_valid();
        accepted += gpu[k]->accepted_phase_set() != nullptr;
        three += gpu[k]->solution.accepted_phase_count() == 3;
        if (!d.empty()) {
            ++mismatches;
            if (mismatches <= 10) {
                std::printf("cell %zu: %s (status word 0x%x)\n", k, d.c_str(), c.has_record ? c.status[k] : -1);
            }
        }
    }
    std::printf("adapter_check %s: cells=%zu structurally_valid=%d accepted=%d three_phase=%d rejected=%d "
                "mismatches=%d\n", argv[1], n, valid, accepted, three, rejected, mismatches);
    return mismatches == 0 ? 0 : 1;
}
