// transform.hpp -- launch interface of the complex 2-D DFT kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace opiv {

// n stacked win x win windows of interleaved complex doubles (device memory);
// in and out may not overlap
hipError_t launch_fft_transform(const double* in, double* out, int win, size_t n, bool forward, hipStream_t stream);

} // namespace opiv
