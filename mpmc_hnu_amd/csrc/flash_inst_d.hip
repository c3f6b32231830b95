// Explicit gfx950 instantiations of the flash kernels for nc in {10,11,12}.
#include "flash_pipeline.hpp"

namespace mpmc::dev {
MPMC_INSTANTIATE_LAUNCH(10)
MPMC_INSTANTIATE_PIPELINE(10)
MPMC_INSTANTIATE_LAUNCH(11)
MPMC_INSTANTIATE_PIPELINE(11)
MPMC_INSTANTIATE_LAUNCH(12)
MPMC_INSTANTIATE_PIPELINE(12)
} // namespace mpmc::dev
