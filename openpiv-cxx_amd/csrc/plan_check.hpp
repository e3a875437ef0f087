// plan_check.hpp -- descriptor checks of plans and batch pairs, with no HIP
// dependency: every rule that opiv_plan_create and opiv_batch_create apply to
// their arguments before they touch the device, and the grid they derive
// (core::generate_cartesian_grid, grid.impl.h:10-61).
#pragma once

#include <cstddef>
#include <cstdint>

#include "geom.hpp"
#include "openpiv_amd.h"

namespace opiv::detail {

// What a plan descriptor resolves to: the row pitch, the variant after
// OPIV_AUTO, and the window grid.
struct PlanGeom {
    size_t pitch = 0;
    int32_t variant = OPIV_CIRCULAR_FFT_FP32;
    uint32_t off_x = 0, off_y = 0;
    uint32_t nx = 0, ny = 0;
    int32_t x_start = 0, y_start = 0;
};

// algos::FFT's constructor check (fft.h:76-84) and the engine's window range
int check_window(uint32_t win);

// OPIV_AUTO's pick for a window size and pixel type
int32_t auto_variant(int win, int dtype);

// every check of opiv_plan_create; fills g on success
int check_plan_desc(const opiv_plan_desc* d, PlanGeom* g);

// pair i of pass `pass` of a batch (window win, overlap, pixel type dtype):
// its grid and table record, first and out_first left to the caller
int check_batch_pair(const opiv_pair_desc& pr, size_t i, uint32_t pass, uint32_t win, double overlap, int dtype,
                     BatchPair* e);

} // namespace opiv::detail
