// Explicit gfx950 instantiations of the closure kernels for nc in {8, 9, 10, 11, 12}.
#include "closure_device.hpp"

namespace mpmc::dev {
MPMC_INSTANTIATE_CLOSURE(8)
MPMC_INSTANTIATE_CLOSURE(9)
MPMC_INSTANTIATE_CLOSURE(10)
MPMC_INSTANTIATE_CLOSURE(11)
MPMC_INSTANTIATE_CLOSURE(12)
} // namespace mpmc::dev
