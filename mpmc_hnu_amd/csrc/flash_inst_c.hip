// Explicit gfx950 instantiations of the flash kernels for nc in {7,8,9}.
#include "flash_pipeline.hpp"

namespace mpmc::dev {
MPMC_INSTANTIATE_LAUNCH(7)
MPMC_INSTANTIATE_PIPELINE(7)
MPMC_INSTANTIATE_LAUNCH(8)
MPMC_INSTANTIATE_PIPELINE(8)
MPMC_INSTANTIATE_LAUNCH(9)
MPMC_INSTANTIATE_PIPELINE(9)
} // namespace mpmc::dev
