// adjoint.hpp -- launch interface of the displacement adjoint kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "openpiv_amd.h"

namespace opiv {

// per-window tap weights of the VJP: d(cot . (u, v))/dC_k for the five
// stencil taps k = l, c, r, up, down (zero for windows without a fit), and
// the peak shift in correlation coordinates, sx | sy << 16
struct __attribute__((aligned(16))) AdjCoef {
    float f[5];
    int32_t shift;
    int32_t pad_[2];
};

struct AdjParams {
    const void* a;
    const void* b;
    size_t pitch;                    // elements per frame row
    int32_t x_start, y_start;
    int32_t step_x, step_y;
    int32_t nx, ny;
    int32_t n_windows;
    int32_t img_h;                   // frame rows (the VJP writes gpitch x img_h gradients)
    const opiv_window_result* fwd;   // forward results (integer peaks, flags), device
    const double* cot;               // VJP: (du, dv) cotangent per window, device
    AdjCoef* coef;                   // VJP: per-window tap weights (32 B each, scratch), device
    float* grad_a;                   // VJP: frame-sized gradients (every element written), device
    float* grad_b;
    size_t gpitch;                   // elements per gradient row
    double sigma;                    // variance: pixel noise
    double* var;                     // variance: (var u, var v) per window, device
};

// variance = false: grad_a/grad_b = sum_i cot_i . d(u_i, v_i)/d(frames)
//                   (two launches: tap weights per window, then a gather per pixel)
// variance = true:  var_i = sigma^2 * |d(u_i, v_i)/d(frames)|^2
hipError_t launch_adjoint(const AdjParams& p, int win, int dtype, bool variance, hipStream_t stream);

} // namespace opiv
