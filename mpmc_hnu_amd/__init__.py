"""mpmc_hnu_amd — MI355X (gfx950) implementation of MPMC_HNU's per-cell PR76
compositional kernel: PT flash (TPD stability, cubic roots, Rachford-Rice, SSI, up to
three phases), behind a C ABI (include/mpmc_hnu_amd.h).

Submodules:
  native     ctypes binding of the C ABI (no fallback path)
  flash      batched flash / phase evaluation on device tensors
  workloads  seeded synthetic inputs
  build      hipcc build recipe
"""
from . import native, workloads  # noqa: F401

__all__ = ["native", "workloads"]
