// Explicit gfx950 instantiations of the two-phase assembly kernels for nc in {8, 9, 10, 11, 12}.
#include "assembly_device.hpp"

namespace mpmc::dev {
MPMC_INSTANTIATE_ASSEMBLY(8)
MPMC_INSTANTIATE_ASSEMBLY(9)
MPMC_INSTANTIATE_ASSEMBLY(10)
MPMC_INSTANTIATE_ASSEMBLY(11)
MPMC_INSTANTIATE_ASSEMBLY(12)
} // namespace mpmc::dev
