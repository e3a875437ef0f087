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
    uint32_t nx_mul;         // fastdiv_mul(nx)
};

// Division by a grid width without a divide: the round-up method of
// Granlund and Montgomery (PLDI 1994, fig. 4.1).  For 1 <= d < 2^32 with
// l = ceil(log2 d) and m = floor(2^32 (2^l - d) / d) + 1, every 32-bit n has
// n / d = (t + ((n - t) >> min(l, 1))) >> max(l - 1, 0), t = (m n) >> 32.
// constexpr, so the kernels use the same code (t is one multiply-high).
constexpr uint32_t ceil_log2_u32(uint32_t d)
{
    uint32_t l = 0;
    while (l < 32 && (uint64_t(1) << l) < d) ++l;
    return l;
}
// (d = 0 gives 0: no grid has zero columns, and the host never divides by it)
constexpr uint32_t fastdiv_mul(uint32_t d)
{
    return d == 0 ? 0u : (uint32_t)(((uint64_t(1) << 32) * ((uint64_t(1) << ceil_log2_u32(d)) - d)) / d + 1);
}
constexpr uint32_t fastdiv(uint32_t n, uint32_t l, uint32_t m)
{
    const uint32_t t = (uint32_t)(((uint64_t)m * n) >> 32);
    return (t + ((n - t) >> (l < 1 ? l : 1))) >> (l > 0 ? l - 1 : 0);
}
static_assert(fastdiv(0, ceil_log2_u32(1), fastdiv_mul(1)) == 0 &&
              fastdiv(4294967295u, ceil_log2_u32(1), fastdiv_mul(1)) == 4294967295u &&
              fastdiv(65024, ceil_log2_u32(255), fastdiv_mul(255)) == 254 &&
              fastdiv(65025, ceil_log2_u32(255), fastdiv_mul(255)) == 255 &&
              fastdiv(4294967295u, ceil_log2_u32(7), fastdiv_mul(7)) == 613566756u &&
              fastdiv(4294967295u, ceil_log2_u32(2147483649u), fastdiv_mul(2147483649u)) == 1 &&
              fastdiv(1023, ceil_log2_u32(1024), fastdiv_mul(1024)) == 0 &&
              fastdiv(1024, ceil_log2_u32(1024), fastdiv_mul(1024)) == 1,
              "round-up division by a multiplier");

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
