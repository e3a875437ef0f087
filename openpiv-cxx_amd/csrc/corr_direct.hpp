// corr_direct.hpp -- launch interface of the BF16-split MFMA direct
// correlation kernel (16 x 16 windows).
#pragma once

#include <hip/hip_runtime.h>

#include "corr_fft.hpp"

namespace opiv {

// same parameters and outputs as launch_corr_fft (single pair, one task per
// window); the circular variant only
hipError_t launch_corr_direct(const CorrParams& p, int win, int dtype, bool planes, hipStream_t stream);

// 32 x 32 windows of 8- or 16-bit frames (corr_direct32.hip)
hipError_t launch_corr_direct32(const CorrParams& p, int dtype, bool planes, hipStream_t stream);

} // namespace opiv
