// background.hpp -- launch interface of the background-subtraction kernels
// (examples/average_subtract/main.cpp:98-159).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace opiv {

struct BgParams {
    const void* const* frames;     // device table of n frame pointers (device memory)
    uint16_t* const* out;          // device table of n output pointers (w * h, row-major)
    int n;
    int w, h;
    size_t pitch;                  // frame elements per row
    double* mean;                  // w * h, row-major, device
    unsigned long long* range;     // 2 per frame: ordered keys of min and max
    bool vec;                      // runs of 8 pixels as single loads and stores: w and pitch
                                   // multiples of 8, every frame, output and mean 16-byte aligned
};

// mean over frames 1..n-1, per-frame range of frame - mean, 16-bit scaling
hipError_t launch_average_subtract(const BgParams& p, int dtype, hipStream_t stream);

} // namespace opiv
