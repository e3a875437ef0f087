// plan_check.cpp -- the grid (core::generate_cartesian_grid,
// grid.impl.h:10-61) and the descriptor checks of plans and batch pairs.
// Host-only C++ with no HIP dependency, so that it also builds and runs
// under the address and undefined-behaviour sanitizers on the CPU.
#include "plan_check.hpp"

#include "status.hpp"

using namespace opiv::detail;

extern "C" {

int opiv_grid_offsets(uint32_t win_w, uint32_t win_h, double overlap, uint32_t* off_x, uint32_t* off_y)
{
    if (!off_x || !off_y) return fail(OPIV_E_INVALID, "opiv_grid_offsets: null output pointer");
    // grid.impl.h:14-21 (a NaN fails both comparisons: rejected too)
    if (!(overlap >= 0.0 && overlap <= 1.0)) return fail(OPIV_E_GRID, "offsets must be between 0.0 and 1.0");
    *off_x = (uint32_t)(win_w * overlap);
    *off_y = (uint32_t)(win_h * overlap);
    return OPIV_OK;
}

int opiv_grid(uint32_t img_w, uint32_t img_h, uint32_t win_w, uint32_t win_h,
              uint32_t off_x, uint32_t off_y, uint32_t* nx, uint32_t* ny, int32_t* xy)
{
    if (!nx || !ny) return fail(OPIV_E_INVALID, "opiv_grid: null output pointer");
    // grid.impl.h:29-40 (size::area() is a uint32 product)
    if ((uint32_t)(img_w * img_h) == 0) return fail(OPIV_E_GRID, "image size must be non-zero");
    if ((uint32_t)(win_w * win_h) == 0) return fail(OPIV_E_GRID, "interrogation size must be non-zero");
    if (off_x == 0 || off_y == 0) return fail(OPIV_E_GRID, "offsets must be non-zero");
    if (win_w > img_w || win_h > img_h) return fail(OPIV_E_GRID, "interrogation size is bigger than image");
    // grid.impl.h:44-58
    const uint32_t x_count = 1 + (img_w - win_w) / off_x;
    const uint32_t y_count = 1 + (img_h - win_h) / off_y;
    const uint32_t x_start = (img_w - win_w - off_x * (x_count - 1)) / 2;
    const uint32_t y_start = (img_h - win_h - off_y * (y_count - 1)) / 2;
    *nx = x_count;
    *ny = y_count;
    if (xy) {
        size_t k = 0;
        for (uint32_t y = 0; y < y_count; ++y)
            for (uint32_t x = 0; x < x_count; ++x, ++k) {
                xy[2 * k] = (int32_t)(x_start + x * off_x);
                xy[2 * k + 1] = (int32_t)(y_start + y * off_y);
            }
    }
    return OPIV_OK;
}

} // extern "C"

namespace opiv::detail {

int check_window(uint32_t win)
{
    // fft.h:82-83
    if (!is_pow2(win)) return fail(OPIV_E_NOT_POW2, "dimensions must be power of 2: [%u,%u]", win, win);
    if (win < kMinWin || win > kMaxWin)
        return fail(OPIV_E_UNSUPPORTED, "window %u outside the supported range [%u, %u]", win, kMinWin, kMaxWin);
    return OPIV_OK;
}

// OPIV_AUTO: the faster path as measured on MI355X at 4096 x 4096, step W/2:
// the direct kernel for 8-bit frames at 16 x 16 (exact int8 MFMA, 65.7 us
// against the FFT kernel's 74.1 us), the FFT kernel otherwise (16 x 16 u16:
// 76 against 173 us, f32: 82 against 254 us; 32 x 32 u8: 79 against 87 us,
// u16: 81 against 220 us).
int32_t auto_variant(int win, int dtype)
{
    return win == 16 && dtype == OPIV_U8 ? OPIV_DIRECT_BF16_SPLIT : OPIV_CIRCULAR_FFT_FP32;
}

int check_plan_desc(const opiv_plan_desc* d, PlanGeom* g)
{
    if (!d || !g) return fail(OPIV_E_INVALID, "opiv_plan_create: null argument");
    const uint32_t win = d->win;
    if (int rc = check_window(win)) return rc;
    const size_t es = dtype_size(d->dtype);
    if (es == 0) return fail(OPIV_E_INVALID, "unknown dtype %d", d->dtype);
    if (d->variant < OPIV_CIRCULAR_FFT_FP32 || d->variant > OPIV_AUTO)
        return fail(OPIV_E_UNSUPPORTED, "unknown variant %d", d->variant);
    if (d->variant == OPIV_DIRECT_BF16_SPLIT && !direct_supported((int)win, d->dtype))
        return fail(OPIV_E_UNSUPPORTED, "the direct path does not implement %ux%u windows of this pixel type", win,
                    win);
    const size_t pitch = d->pitch ? d->pitch : d->img_w;
    if (pitch < d->img_w) return fail(OPIV_E_INVALID, "pitch %zu smaller than width %u", pitch, d->img_w);
    if (pitch > kMaxPitchBytes / es)
        return fail(OPIV_E_UNSUPPORTED, "row pitch of %zu pixels exceeds %zu bytes", pitch, kMaxPitchBytes);
    PlanGeom r;
    r.pitch = pitch;
    r.variant = d->variant == OPIV_AUTO ? auto_variant((int)win, d->dtype) : d->variant;
    if (d->step_x && d->step_y) {
        r.off_x = d->step_x;
        r.off_y = d->step_y;
    } else if (int rc = opiv_grid_offsets(win, win, d->overlap, &r.off_x, &r.off_y)) {
        return rc;
    }
    if (int rc = opiv_grid(d->img_w, d->img_h, win, win, r.off_x, r.off_y, &r.nx, &r.ny, nullptr)) return rc;
    if ((uint64_t)r.nx * r.ny > 0x7fffffffull) return fail(OPIV_E_UNSUPPORTED, "too many windows");
    if (!task_span_fits((int)win, d->dtype, r.nx, r.ny, r.off_y, pitch * es))
        return fail(OPIV_E_UNSUPPORTED, "grid step %u with %zu-byte rows spans more than 4 GiB per wave task",
                    r.off_y, pitch * es);
    r.x_start = (int32_t)((d->img_w - win - r.off_x * (r.nx - 1)) / 2);
    r.y_start = (int32_t)((d->img_h - win - r.off_y * (r.ny - 1)) / 2);
    *g = r;
    return OPIV_OK;
}

int check_batch_pair(const opiv_pair_desc& pr, size_t i, uint32_t pass, uint32_t win, double overlap, int dtype,
                     BatchPair* e)
{
    const size_t es = dtype_size(dtype);
    if (es == 0) return fail(OPIV_E_INVALID, "unknown dtype %d", dtype);
    if (!pr.a || !pr.b) return fail(OPIV_E_INVALID, "pair %zu: null frame pointer", i);
    if (((uintptr_t)pr.a | (uintptr_t)pr.b) % es != 0)
        return fail(OPIV_E_INVALID, "pair %zu: frame pointers must be aligned to their %zu-byte pixels", i, es);
    const size_t pitch = pr.pitch ? pr.pitch : pr.img_w;
    if (pitch < pr.img_w) return fail(OPIV_E_INVALID, "pair %zu: pitch smaller than width", i);
    if (pitch > kMaxPitchBytes / es)
        return fail(OPIV_E_UNSUPPORTED, "pair %zu: row pitch of %zu pixels exceeds %zu bytes", i, pitch,
                    kMaxPitchBytes);
    uint32_t ox, oy, nx, ny;
    if (int rc = opiv_grid_offsets(win, win, overlap, &ox, &oy)) return rc;
    if (int rc = opiv_grid(pr.img_w, pr.img_h, win, win, ox, oy, &nx, &ny, nullptr)) {
        const std::string why = g_last_error;
        return fail(rc, "pair %zu, pass %u: %s", i, pass, why.c_str());
    }
    if ((uint64_t)nx * ny > 0x7fffffffull) return fail(OPIV_E_UNSUPPORTED, "pair %zu, pass %u: too many windows", i, pass);
    if (!task_span_fits((int)win, dtype, nx, ny, oy, pitch * es))
        return fail(OPIV_E_UNSUPPORTED, "pair %zu, pass %u: rows of %zu bytes span more than 4 GiB per wave task", i,
                    pass, pitch * es);
    e->a = pr.a;
    e->b = pr.b;
    e->pitch = pitch;
    e->x_start = (int32_t)((pr.img_w - win - ox * (nx - 1)) / 2);
    e->y_start = (int32_t)((pr.img_h - win - oy * (ny - 1)) / 2);
    e->step_x = (int32_t)ox;
    e->step_y = (int32_t)oy;
    e->nx = (int32_t)nx;
    e->ny = (int32_t)ny;
    e->img_h = pr.img_h;
    e->nx_mul = fastdiv_mul(nx);
    return OPIV_OK;
}

} // namespace opiv::detail
