// corr_fft.hpp -- launch interface of the fused FFT correlation kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "geom.hpp"
#include "openpiv_amd.h"

namespace opiv {

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
    uint32_t nx_mul;         // fastdiv_mul(nx)
    int32_t first;
    int32_t n_windows;
    uint32_t img_h;          // for the y flip of main.cpp:230
    int32_t limit_search;
    int32_t real_compat;
    float tie_tol;           // relative FP32 tie tolerance for OPIV_FLAG_NEAR_TIE
    opiv_window_result* out; // fused mode: results indexed by local window
    double* planes;          // plane mode: swapped planes indexed by local window
    const BatchPair* pairs;  // batch mode (fused only): pair table, else nullptr
    int32_t npairs;
};

// Default shape: 256 threads per block, one block per 4 wave tasks.
hipError_t launch_corr_fft(const CorrParams& p, int win, int dtype, bool planes, hipStream_t stream);

// Explicit shape: nt threads per block (256; batch launches also 64) and at
// most `blocks` blocks (0 = one per task group; fewer makes the launch
// persistent).  The shape never changes the arithmetic of a window.
hipError_t launch_corr_fft_shape(const CorrParams& p, int win, int dtype, bool planes, int nt, int blocks,
                                 hipStream_t stream);

// Every pass of a batch in one persistent launch (corr_fft_multi.hip):
// pass k is a batch-mode CorrParams (pair table) of window size win[k] in
// [8, 64]; the grid's waves work through the passes in order, each XCD
// taking a contiguous share of every pass.  blocks: the persistent grid
// (two 256-thread blocks fit per CU).
struct MultiPassParams {
    CorrParams pass[4];
    int32_t win[4];
    int32_t npasses;
};
hipError_t launch_corr_fft_multipass(const MultiPassParams& mp, int dtype, int blocks, hipStream_t stream);

// 128 x 128 windows (corr_fft128.hip): one workgroup of 256 threads per
// window, single pair or batch (p.pairs), fused or plane mode
hipError_t launch_corr_fft128(const CorrParams& p, int dtype, bool planes, hipStream_t stream);

} // namespace opiv
