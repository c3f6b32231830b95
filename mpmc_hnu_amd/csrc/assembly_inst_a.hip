// Explicit gfx950 instantiations of the two-phase assembly kernels for nc in {2, 3, 4, 5, 6, 7}.
#include "assembly_device.hpp"

namespace mpmc::dev {
MPMC_INSTANTIATE_ASSEMBLY(2)
MPMC_INSTANTIATE_ASSEMBLY(3)
MPMC_INSTANTIATE_ASSEMBLY(4)
MPMC_INSTANTIATE_ASSEMBLY(5)
MPMC_INSTANTIATE_ASSEMBLY(6)
MPMC_INSTANTIATE_ASSEMBLY(7)
} // namespace mpmc::dev
