// Explicit gfx950 instantiations of the flash kernels for nc in {1,2,3}.
#include "flash_kernels.hpp"

namespace mpmc::dev {
MPMC_INSTANTIATE_LAUNCH(1)
MPMC_INSTANTIATE_LAUNCH(2)
MPMC_INSTANTIATE_LAUNCH(3)
} // namespace mpmc::dev
