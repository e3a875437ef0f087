// geom.hpp -- window geometry shared by the host code and the kernels, with
// no HIP dependency: the batch pair record, how many windows a wave task of
// the FFT kernel takes, the limits that keep its 32-bit frame offsets in
// range, and the window sizes each correlation path implements.
#pragma once

#include <cstddef>
#include <cstdint>

namespace opiv {

// square window sides the engine correlates (powers of two, fft.h:76-84)
constexpr uint32_t kMinWin = 8;
constexpr uint32_t kMaxWin = 128;

// window sizes and pixel types (0 u8, 1 u16, 2 f32, 3 f64) the direct
// (MFMA) path implements: 16 x 16 for every pixel type, 32 x 32 for 8- and
// 16-bit frames (int8 byte planes)
inline bool direct_supported(int win, int dtype) { return win == 16 || (win == 32 && (dtype == 0 || dtype == 1)); }

// One frame pair of a batch pass (device table): its frames and its grid.
// The pass's windows are numbered pair by pair, pair k owning
// [first, first + nx * ny); each pair's range starts on a multiple of the
// kernel's task size (batch_align), so every window shares its packed
// inverse FFT with the same partner as in a single-pair launch.  Results of
// pair k start at out_first.
struct BatchPair {
    const void* a;
    const void* b;
    size_t pitch;
    int32_t x_start, y_start;
    int32_t step_x, step_y;
    int32_t nx, ny;
    uint32_t img_h;
    int32_t first;
    int32_t out_first;
    int32_t pad_;
};

// windows per wave task of the kernel for window size win and pixel type
// dtype (0 u8, 1 u16, 2 f32, 3 f64): two rounds of 64 / win lane groups, one
// round for f64 frames at 64 x 64; one window per workgroup at 128 x 128
constexpr int batch_align(int win, int dtype)
{
    return win >= 128 ? 1 : (win == 64 && dtype == 3 ? 1 : 2) * (64 / win);
}

// Largest frame row pitch in bytes.  The FFT kernel addresses a wave task's
// pixels as 32-bit byte offsets from the task's first window: a task spans
// at most 129 rows of pitch, plus 63 more for a window's last row.
constexpr size_t kMaxPitchBytes = size_t(1) << 24;

// Whether every wave task of a grid (nx x ny windows of side win, row step
// step_y, rows pitch_bytes apart) stays within 32-bit byte offsets of its
// first window: a task's windows are consecutive in grid order, so they span
// at most ceil((task - 1) / nx) grid rows, plus one window height.
inline bool task_span_fits(int win, int dtype, uint64_t nx, uint64_t ny, uint64_t step_y, uint64_t pitch_bytes)
{
    const uint64_t tw = (uint64_t)batch_align(win, dtype);
    if (nx == 0 || ny == 0) return true;
    uint64_t diy = (tw - 1 + nx - 1) / nx;
    if (diy > ny - 1) diy = ny - 1;
    return (diy * step_y + (uint64_t)win) * pitch_bytes < (uint64_t(1) << 32);
}

} // namespace opiv
