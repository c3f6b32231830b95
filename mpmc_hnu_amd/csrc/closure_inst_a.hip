// Explicit gfx950 instantiations of the closure kernels for nc in {2, 3, 4, 5, 6, 7}.
#include "closure_device.hpp"

namespace mpmc::dev {
MPMC_INSTANTIATE_CLOSURE(2)
MPMC_INSTANTIATE_CLOSURE(3)
MPMC_INSTANTIATE_CLOSURE(4)
MPMC_INSTANTIATE_CLOSURE(5)
MPMC_INSTANTIATE_CLOSURE(6)
MPMC_INSTANTIATE_CLOSURE(7)
} // namespace mpmc::dev
