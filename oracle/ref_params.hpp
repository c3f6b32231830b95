// Test infrastructure: the reference's PR76 parameter snapshot for a synthetic mixture,
// shared by the reference drivers under oracle/ (compiled with the reference's headers).
#pragma once

#include <mpmc/thermodynamics/pr_parameters.hpp>

#include <optional>
#include <string>
#include <vector>

namespace mpmc_oracle {

namespace th = mpmc::thermodynamics;

inline th::Provenance synthetic(const std::string& locator) {
    return {th::SourceKind::synthetic_test, "mpmc-hnu-amd://oracle", "v1", locator,
            "Synthetic mixture for parity vectors", "generated", "none"};
}

inline th::SourcedScalar datum(double value, th::Unit unit, const std::string& locator) {
    return {value, unit, synthetic(locator), "SI", "identity"};
}

inline th::PrParameterSet make_parameters(int nc, const double* tc, const double* pc,
                                   const double* omega, const double* kij) {
    std::vector<th::Component> catalog;
    th::PrParameterInput input;
    input.model_id = std::string(th::pr76_profile);
    input.dataset_id = "mpmc-hnu-amd-synthetic";
    input.revision = "v1";
    input.applicability = {std::nullopt, std::nullopt, synthetic("bounds")};
    std::vector<std::string> order;
    for (int i = 0; i < nc; ++i) {
        const std::string id = "c" + std::to_string(i);
        order.push_back(id);
        catalog.push_back({id, id, th::ComponentKind::pure, synthetic(id), {}});
        input.pure.push_back({id, datum(tc[i], th::Unit::kelvin, id + ".tc"),
                              datum(pc[i], th::Unit::pascal, id + ".pc"),
                              datum(omega[i], th::Unit::dimensionless, id + ".omega")});
        for (int j = 0; j < i; ++j) {
            input.binary.push_back({id, "c" + std::to_string(j),
                                    datum(kij[i * nc + j], th::Unit::dimensionless, id + ".kij")});
        }
    }
    return th::PrParameterSet::create(catalog, order, input, th::DataPolicy::allow_synthetic_tests);
}

} // namespace mpmc_oracle
