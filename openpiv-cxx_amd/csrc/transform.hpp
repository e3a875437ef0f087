// transform.hpp -- launch interface of the complex 2-D DFT kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace opiv {

// n stacked win x win windows of interleaved complex doubles (device memory,
// win 8 to 128); in == out is allowed, other overlaps are not
hipError_t launch_fft_transform(const double* in, double* out, int win, size_t n, bool forward, hipStream_t stream);

// the 128 x 128 form (corr_fft128.hip)
hipError_t launch_fft_transform128(const double* in, double* out, size_t n, bool forward, hipStream_t stream);

} // namespace opiv
