// Explicit gfx950 instantiations of the flash kernels for nc in {1,2,3}.
#include "flash_pipeline.hpp"

namespace mpmc::dev {
MPMC_INSTANTIATE_LAUNCH(1)
MPMC_INSTANTIATE_PIPELINE(1)
MPMC_INSTANTIATE_LAUNCH(2)
MPMC_INSTANTIATE_PIPELINE(2)
MPMC_INSTANTIATE_LAUNCH(3)
MPMC_INSTANTIATE_PIPELINE(3)
} // namespace mpmc::dev
