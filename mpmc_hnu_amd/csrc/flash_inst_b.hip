// Explicit gfx950 instantiations of the flash kernels for nc in {4,5,6}.
#include "flash_pipeline.hpp"

namespace mpmc::dev {
MPMC_INSTANTIATE_LAUNCH(4)
MPMC_INSTANTIATE_PIPELINE(4)
MPMC_INSTANTIATE_LAUNCH(5)
MPMC_INSTANTIATE_PIPELINE(5)
MPMC_INSTANTIATE_LAUNCH(6)
MPMC_INSTANTIATE_PIPELINE(6)
} // namespace mpmc::dev
