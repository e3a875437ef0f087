// corr_fft.hpp -- launch interface of the fused FFT correlation kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "openpiv_amd.h"

namespace opiv {

constexpr int kCorrThreads = 256;

// One launch processes windows [first, first + n_windows) of a cartesian
// grid (grid.impl.h:44-58): window g sits at
// (x_start + (g % nx) * step_x, y_start + (g / nx) * step_y).
struct CorrParams {
    const void* a;
    const void* b;
    size_t pitch;            // elements per frame row
    int32_t x_start, y_start;
    int32_t step_x, step_y;
    int32_t nx;
    int32_t first;
    int32_t n_windows;
    uint32_t img_h;          // for the y flip of main.cpp:230
    int32_t limit_search;
    int32_t real_compat;
    float tie_tol;           // relative FP32 tie tolerance for OPIV_FLAG_NEAR_TIE
    opiv_window_result* out; // fused mode: results indexed by local window
    double* planes;          // plane mode: swapped planes indexed by local window
};

hipError_t launch_corr_fft(const CorrParams& p, int win, int dtype, bool planes, hipStream_t stream);

} // namespace opiv
