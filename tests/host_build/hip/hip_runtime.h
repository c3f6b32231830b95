// Stand-in for <hip/hip_runtime.h> when the device headers of this library
// (mpmc_hnu_amd/csrc/pr76_device.hpp, flash_device.hpp) are compiled as plain host C++ by
// the CPU tests: the HIP function attributes become no-ops and the math functions are
// the C library's. Only the per-lane search code is built this way; no kernel is.
#pragma once
#include <cmath>
#define __device__
#define __host__
#define __forceinline__ inline __attribute__((always_inline))
#define __noinline__ __attribute__((noinline))
using std::exp;
using std::expm1;
using std::log;
using std::log1p;
using std::sqrt;
