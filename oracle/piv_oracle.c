/*
 * piv_oracle.c -- FP64 CPU restatement of the openpiv-cxx window-correlation
 * path.  TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Build with -O2 -ffp-contract=off and no -march (baseline x86-64 has no FMA),
 * so that every operation rounds exactly as the reference's FP64 code does.
 * Each function names the reference file:line it restates.
 */
#define _GNU_SOURCE
#include "piv_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------ */
/* complex arithmetic: openpiv/core/pixel_types.h:149-198             */
/* ------------------------------------------------------------------ */

static inline oref_c c_add(oref_c x, oref_c y) { oref_c r = { x.re + y.re, x.im + y.im }; return r; }
static inline oref_c c_sub(oref_c x, oref_c y) { oref_c r = { x.re - y.re, x.im - y.im }; return r; }
/* pixel_types.h:159-163: (a*c - b*d, b*c + a*d) */
static inline oref_c c_mul(oref_c x, oref_c y)
{
    const double a = x.re, b = x.im, c = y.re, d = y.im;
    oref_c r = { a * c - b * d, b * c + a * d };
    return r;
}
/* pixel_types.h:196 */
static inline oref_c c_conj(oref_c x) { oref_c r = { x.re, -x.im }; return r; }

/* ------------------------------------------------------------------ */
/* grid: openpiv/core/detail/grid.impl.h                              */
/* ------------------------------------------------------------------ */

int oref_grid_offsets_from_fraction(uint32_t win_w, uint32_t win_h, double fraction,
                                    uint32_t* off_x, uint32_t* off_y)
{
    /* grid.impl.h:14-15 */
    if (fraction < 0.0 || fraction > 1.0)
        return -1;
    /* grid.impl.h:20-21: truncating conversion of win * fraction */
    *off_x = (uint32_t)(win_w * fraction);
    *off_y = (uint32_t)(win_h * fraction);
    return 0;
}

int oref_grid(uint32_t img_w, uint32_t img_h, uint32_t win_w, uint32_t win_h,
              uint32_t off_x, uint32_t off_y, uint32_t* nx, uint32_t* ny, int32_t* xy)
{
    /* grid.impl.h:29-40 (size::area() is a uint32 product) */
    if ((uint32_t)(img_w * img_h) == 0) return -1;
    if ((uint32_t)(win_w * win_h) == 0) return -2;
    if (off_x == 0 || off_y == 0) return -3;
    if (win_w > img_w || win_h > img_h) return -4;

    /* grid.impl.h:44-48 */
    const uint32_t x_count = 1 + (img_w - win_w) / off_x;
    const uint32_t y_count = 1 + (img_h - win_h) / off_y;
    const uint32_t x_start = (img_w - win_w - (off_x * (x_count - 1))) / 2;
    const uint32_t y_start = (img_h - win_h - (off_y * (y_count - 1))) / 2;
    *nx = x_count;
    *ny = y_count;
    if (xy) {
        /* grid.impl.h:51-58: y outer, x inner */
        size_t k = 0;
        for (size_t y = 0; y < y_count; ++y)
            for (size_t x = 0; x < x_count; ++x) {
                xy[2 * k + 0] = (int32_t)(x_start + x * off_x);
                xy[2 * k + 1] = (int32_t)(y_start + y * off_y);
                ++k;
            }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* FFT: openpiv/algos/fft.h                                           */
/* ------------------------------------------------------------------ */

#define OREF_MAX_LOG2 16

struct oref_fft {
    uint32_t w, h;
    /* twiddles per n = N, N/2, ... > 2 (fft.h:274-294), indexed by log2(n) */
    oref_c* tw[2][OREF_MAX_LOG2 + 1];
};

static int is_pow2(uint64_t v) { return v && (v == (v & (~v + 1))); } /* util.h:67-70 */

static int ilog2(uint64_t v) { int l = 0; while ((1ull << l) < v) ++l; return l; }

oref_fft* oref_fft_create(uint32_t w, uint32_t h)
{
    /* fft.h:82-83 */
    if (!(is_pow2(w) && is_pow2(h)))
        return NULL;
    oref_fft* f = (oref_fft*)calloc(1, sizeof(oref_fft));
    f->w = w;
    f->h = h;
    for (int d = 0; d < 2; ++d) {
        /* fft.h:277-291: n from maximal_size(size).width() down while n > 2 */
        size_t n = w > h ? w : h;
        do {
            const double scaling = (d == OREF_FORWARD) ? -1.0 : 1.0;
            oref_c* t = (oref_c*)malloc(sizeof(oref_c) * n);
            for (size_t i = 0; i < n; ++i) {
                /* fft.h:285: theta = (scaling * M_PI * i)/n with int i, size_t n */
                const double theta = (scaling * M_PI * (double)(int)i) / (double)n;
                t[i].re = cos(theta);
                t[i].im = sin(theta);
            }
            f->tw[d][ilog2(n)] = t;
            n /= 2;
        } while (n > 2);
    }
    return f;
}

void oref_fft_destroy(oref_fft* f)
{
    if (!f) return;
    for (int d = 0; d < 2; ++d)
        for (int l = 0; l <= OREF_MAX_LOG2; ++l)
            free(f->tw[d][l]);
    free(f);
}

/* fft.h:245-263: recursive radix-2 decimation in time, ping-pong buffers */
static void fft_inner(oref_c* in, oref_c* out, const oref_c* scaling, size_t n, size_t step)
{
    if (step < n) {
        const size_t doublestep = 2 * step;
        fft_inner(out, in, scaling, n, doublestep);
        fft_inner(out + step, in + step, scaling, n, doublestep);
        for (size_t i = 0; i < n; i += doublestep) {
            const oref_c e = out[i];
            const oref_c o = c_mul(out[i + step], scaling[i]);
            in[i / 2] = c_add(e, o);
            in[(i + n) / 2] = c_sub(e, o);
        }
    }
}

/* fft.h:265-272 */
static void fft_line(const oref_fft* f, oref_c* line, size_t n, int dir, oref_c* buffer)
{
    memcpy(buffer, line, n * sizeof(oref_c));
    fft_inner(line, buffer, f->tw[dir][ilog2(n)], n, 1);
}

/* image_utils.impl.h:317-348: out[w][h] = in[h][w]; in is iw wide, ih high */
static void transpose_c(const oref_c* in, uint32_t iw, uint32_t ih, oref_c* out)
{
    for (uint32_t h = 0; h < ih; ++h)
        for (uint32_t w = 0; w < iw; ++w)
            out[(size_t)w * ih + h] = in[(size_t)h * iw + w];
}

/* fft.h:86-119 (square images: the reference's temp buffer is resized to
 * the input size before the transpose, fft.h:102, so only w == h works) */
static void transform_inplace(const oref_fft* f, oref_c* out, int dir, oref_c* temp, oref_c* buffer)
{
    const uint32_t W = f->w, H = f->h;
    for (uint32_t h = 0; h < H; ++h)
        fft_line(f, out + (size_t)h * W, W, dir, buffer);
    transpose_c(out, W, H, temp);
    for (uint32_t h = 0; h < W; ++h)
        fft_line(f, temp + (size_t)h * H, H, dir, buffer);
    transpose_c(temp, H, W, out);
}

void oref_fft_transform(const oref_fft* f, const oref_c* in, oref_c* out, int dir)
{
    const size_t n = (size_t)f->w * f->h;
    oref_c* temp = (oref_c*)malloc(sizeof(oref_c) * n);
    oref_c* buffer = (oref_c*)malloc(sizeof(oref_c) * (f->w > f->h ? f->w : f->h));
    if (out != in)
        memcpy(out, in, n * sizeof(oref_c));
    transform_inplace(f, out, dir, temp, buffer);
    free(buffer);
    free(temp);
}

/* working storage for one window (the reference's thread-local cache,
 * fft.h:37-73) */
typedef struct {
    oref_c* za;
    oref_c* zb;
    oref_c* temp;
    oref_c* buffer;
} work_t;

static void work_init(work_t* wk, uint32_t W, uint32_t H)
{
    const size_t n = (size_t)W * H;
    wk->za = (oref_c*)malloc(sizeof(oref_c) * n);
    wk->zb = (oref_c*)malloc(sizeof(oref_c) * n);
    wk->temp = (oref_c*)malloc(sizeof(oref_c) * n);
    wk->buffer = (oref_c*)malloc(sizeof(oref_c) * (W > H ? W : H));
}

static void work_free(work_t* wk)
{
    free(wk->za); free(wk->zb); free(wk->temp); free(wk->buffer);
}

/* fft.h:131-186 */
static void transform_real_w(const oref_fft* f, const double* a, const double* b,
                             oref_c* out_a, oref_c* out_b, work_t* wk)
{
    const uint32_t width = f->w, height = f->h;
    const size_t n = (size_t)width * height;
    oref_c* t = wk->za;
    /* join_from_channels: image_utils.impl.h:288-315 */
    for (size_t i = 0; i < n; ++i) { t[i].re = a[i]; t[i].im = b[i]; }
    transform_inplace(f, t, OREF_FORWARD, wk->temp, wk->buffer);

    /* fresh cf_image is zero-initialised (image.h:51-53) */
    memset(out_a, 0, n * sizeof(oref_c));
    memset(out_b, 0, n * sizeof(oref_c));
    const oref_c half = { 0.5, 0.0 };   /* 0.5 * complex -> complex{0.5} * c (pixel_types.h:217-220) */
    for (uint32_t h = 1; h < height / 2; ++h) {
        for (uint32_t w = 1; w < width; ++w) {
            const oref_c t1 = t[(size_t)h * width + w];
            const oref_c t2 = t[(size_t)(height - h) * width + (width - w)];
            const oref_c av = c_mul(half, c_add(t1, c_conj(t2)));
            out_a[(size_t)h * width + w] = av;
            out_a[(size_t)(height - h) * width + (width - w)] = c_conj(av);

            oref_c bv = c_mul(half, c_sub(t1, c_conj(t2)));
            const oref_c bt = { bv.im, -1.0 * bv.re };
            bv = bt;
            out_b[(size_t)h * width + w] = bv;
            out_b[(size_t)(height - h) * width + (width - w)] = c_conj(bv);
        }
    }
}

void oref_fft_transform_real(const oref_fft* f, const double* a, const double* b,
                             oref_c* out_a, oref_c* out_b)
{
    work_t wk;
    work_init(&wk, f->w, f->h);
    transform_real_w(f, a, b, out_a, out_b, &wk);
    work_free(&wk);
}

/* image_utils.impl.h:364-386 */
void oref_swap_quadrants(double* im, uint32_t width, uint32_t height)
{
    for (uint32_t h = 0; h < height; ++h) {
        double* i = im + (size_t)h * width;
        double* o = im + (size_t)((h + height / 2) % height) * width;
        for (uint32_t w = 0; w < width / 2; ++w) {
            const uint32_t ow = (w + width / 2) % width;
            const double tmp = i[w];
            i[w] = o[ow];
            o[ow] = tmp;
        }
    }
}

/* fft.h:195-206 with the workspace supplied */
static void cross_correlate_w(const oref_fft* f, const double* a, const double* b, double* out, work_t* wk)
{
    const size_t n = (size_t)f->w * f->h;
    /* image.h:136-145 + pixel_types.h:432-439: g -> complex{g, 0} */
    for (size_t i = 0; i < n; ++i) { wk->za[i].re = a[i]; wk->za[i].im = 0.0; }
    transform_inplace(f, wk->za, OREF_FORWARD, wk->temp, wk->buffer);
    for (size_t i = 0; i < n; ++i) { wk->zb[i].re = b[i]; wk->zb[i].im = 0.0; }
    transform_inplace(f, wk->zb, OREF_FORWARD, wk->temp, wk->buffer);
    /* fft.h:201: a_fft = b_fft * conj(a_fft) */
    for (size_t i = 0; i < n; ++i)
        wk->za[i] = c_mul(wk->zb[i], c_conj(wk->za[i]));
    transform_inplace(f, wk->za, OREF_REVERSE, wk->temp, wk->buffer);
    /* fft.h:202: real(); fft.h:203: swap_quadrants */
    for (size_t i = 0; i < n; ++i) out[i] = wk->za[i].re;
    oref_swap_quadrants(out, f->w, f->h);
}

/* fft.h:218-227 */
static void cross_correlate_real_w(const oref_fft* f, const double* a, const double* b, double* out, work_t* wk)
{
    const size_t n = (size_t)f->w * f->h;
    oref_c* fa = (oref_c*)malloc(sizeof(oref_c) * n);
    oref_c* fb = (oref_c*)malloc(sizeof(oref_c) * n);
    transform_real_w(f, a, b, fa, fb, wk);
    for (size_t i = 0; i < n; ++i)
        fa[i] = c_mul(fb[i], c_conj(fa[i]));
    transform_inplace(f, fa, OREF_REVERSE, wk->temp, wk->buffer);
    for (size_t i = 0; i < n; ++i) out[i] = fa[i].re;
    oref_swap_quadrants(out, f->w, f->h);
    free(fb);
    free(fa);
}

void oref_cross_correlate(const oref_fft* f, const double* a, const double* b, double* out)
{
    work_t wk;
    work_init(&wk, f->w, f->h);
    cross_correlate_w(f, a, b, out, &wk);
    work_free(&wk);
}

void oref_cross_correlate_real(const oref_fft* f, const double* a, const double* b, double* out)
{
    work_t wk;
    work_init(&wk, f->w, f->h);
    cross_correlate_real_w(f, a, b, out, &wk);
    work_free(&wk);
}

/* fft.h:233-242 */
void oref_auto_correlate(const oref_fft* f, const double* a, double* out)
{
    const size_t n = (size_t)f->w * f->h;
    work_t wk;
    work_init(&wk, f->w, f->h);
    for (size_t i = 0; i < n; ++i) { wk.za[i].re = a[i]; wk.za[i].im = 0.0; }
    transform_inplace(f, wk.za, OREF_FORWARD, wk.temp, wk.buffer);
    /* abs_sqr: (c * conj(c)).real (pixel_types.h:198), stored as complex{v} */
    for (size_t i = 0; i < n; ++i) {
        const double v = c_mul(wk.za[i], c_conj(wk.za[i])).re;
        wk.za[i].re = v;
        wk.za[i].im = 0.0;
    }
    transform_inplace(f, wk.za, OREF_REVERSE, wk.temp, wk.buffer);
    for (size_t i = 0; i < n; ++i) out[i] = wk.za[i].re;
    oref_swap_quadrants(out, f->w, f->h);
    work_free(&wk);
}

/* ------------------------------------------------------------------ */
/* peaks + fit: openpiv/core/detail/image_utils.impl.h                */
/* ------------------------------------------------------------------ */

size_t oref_find_peaks(const double* plane, uint32_t w, uint32_t h,
                       uint32_t rx, uint32_t ry, uint32_t rw, uint32_t rh,
                       uint16_t num_peaks, uint32_t radius,
                       oref_peak* out, int* exact_tie)
{
    (void)h;
    size_t cap = 64, n = 0;
    oref_peak* cand = (oref_peak*)malloc(sizeof(oref_peak) * cap);
    /* image_utils.impl.h:48-64: rows [r, H-2r), columns [r, W-r), strict
     * comparison against the four direct neighbours */
    for (uint32_t y = radius; y < rh - 2 * radius; ++y) {
        const double* above = plane + (size_t)(ry + y - 1) * w + rx;
        const double* line = plane + (size_t)(ry + y) * w + rx;
        const double* below = plane + (size_t)(ry + y + 1) * w + rx;
        for (uint32_t x = radius; x < rw - radius; ++x) {
            if (line[x - 1] < line[x] && line[x + 1] < line[x] && above[x] < line[x] && below[x] < line[x]) {
                if (n == cap) {
                    cap *= 2;
                    cand = (oref_peak*)realloc(cand, sizeof(oref_peak) * cap);
                }
                /* midpoint of the (2r+1)^2 view at (x-r, y-r): rect.cpp:63-66 */
                cand[n].x = (int32_t)(rx + x - radius + (2 * radius + 1) / 2);
                cand[n].y = (int32_t)(ry + y - radius + (2 * radius + 1) / 2);
                cand[n].value = line[x];
                ++n;
            }
        }
    }
    /* image_utils.impl.h:67-71: sort descending by centre value, keep
     * num_peaks.  Insertion sort is stable: equal values keep scan order. */
    for (size_t i = 1; i < n; ++i) {
        const oref_peak v = cand[i];
        size_t j = i;
        while (j > 0 && cand[j - 1].value < v.value) { cand[j] = cand[j - 1]; --j; }
        cand[j] = v;
    }
    const size_t keep = n < num_peaks ? n : num_peaks;
    if (exact_tie)
        *exact_tie = (n >= 2 && num_peaks >= 2 && cand[0].value == cand[1].value);
    memcpy(out, cand, keep * sizeof(oref_peak));
    free(cand);
    return keep;
}

/* image_utils.impl.h:86-94 */
static double gauss3(double l, double c, double r)
{
    const double num = log(l) - log(r);
    const double den = 2.0 * (log(l) + log(r) - 2.0 * log(c));
    if (den == 0.0)
        return 0.0;
    return num / den;
}

/* image_utils.impl.h:81-101: x from ({0,1},{1,1},{2,1}), y from
 * ({1,0},{1,1},{1,2}) of the 3x3 view centred on (cx, cy) */
void oref_fit_simple_gaussian(const double* plane, uint32_t w, int32_t cx, int32_t cy,
                              double* fx, double* fy)
{
    const double* row = plane + (size_t)cy * w;
    const double c = row[cx];
    *fx = (double)cx;
    *fy = (double)cy;
    *fx += gauss3(row[cx - 1], c, row[cx + 1]);
    *fy += gauss3(plane[(size_t)(cy - 1) * w + cx], c, plane[(size_t)(cy + 1) * w + cx]);
}

/* rect.cpp:109-126 */
void oref_dilate(uint32_t w, uint32_t h, double d, int32_t* x, int32_t* y,
                 uint32_t* rw, uint32_t* rh)
{
    const int32_t iw = (int32_t)w, ih = (int32_t)h;
    const int32_t lw = (int32_t)lround(d * iw);
    const int32_t lh = (int32_t)lround(d * ih);
    *x = (iw - lw) / 2;
    *y = (ih - lh) / 2;
    *rw = (uint32_t)lw;
    *rh = (uint32_t)lh;
}

/* ------------------------------------------------------------------ */
/* per-window processor: examples/process/main.cpp:189-237            */
/* ------------------------------------------------------------------ */

static void process_window_w(const oref_fft* f, const double* wa, const double* wb, uint32_t win,
                             int32_t blx, int32_t bly, uint32_t img_h, int variant,
                             int limit_search, double* plane, oref_vector* res, work_t* wk)
{
    memset(res, 0, sizeof(*res));
    if (variant == OREF_CORR_REAL)
        cross_correlate_real_w(f, wa, wb, plane, wk);
    else
        cross_correlate_w(f, wa, wb, plane, wk);

    /* main.cpp:204-211 */
    uint32_t rx = 0, ry = 0, rw = win, rh = win;
    if (limit_search) {
        int32_t dx, dy;
        oref_dilate(win, win, 0.5, &dx, &dy, &rw, &rh);
        rx = (uint32_t)dx;
        ry = (uint32_t)dy;
    }
    oref_peak peaks[2];
    int tie = 0;
    const size_t np = oref_find_peaks(plane, win, win, rx, ry, rw, rh, 2, 1, peaks, &tie);
    if (np != 2) {
        /* main.cpp:214-218: the default (all-zero) record stays */
        res->flags |= OREF_FLAG_FEWER_THAN_2_PEAKS;
        return;
    }
    if (tie) res->flags |= OREF_FLAG_EXACT_TIE;

    double plx, ply;
    oref_fit_simple_gaussian(plane, win, peaks[0].x, peaks[0].y, &plx, &ply);
    if (!isfinite(plx) || !isfinite(ply)) res->flags |= OREF_FLAG_NONFINITE_FIT;

    /* main.cpp:221-227: midpoint and bottom-left are integer points */
    const int32_t midx = blx + (int32_t)(win / 2);
    const int32_t midy = bly + (int32_t)(win / 2);
    res->x = (double)midx;
    res->y = (double)midy;
    res->u = (double)midx - ((double)blx + plx);
    res->v = (double)midy - ((double)bly + ply);
    /* main.cpp:230 */
    res->y = (double)img_h - res->y;
    /* main.cpp:233-234 */
    if (peaks[1].value > 0)
        res->sn = peaks[0].value / peaks[1].value;
    res->px = peaks[0].x;
    res->py = peaks[0].y;
    res->peak0 = peaks[0].value;
    res->peak1 = peaks[1].value;
}

void oref_process_window(const oref_fft* f, const double* wa, const double* wb, uint32_t win,
                         int32_t blx, int32_t bly, uint32_t img_h, int variant,
                         int limit_search, double* plane_scratch, oref_vector* res)
{
    work_t wk;
    work_init(&wk, win, win);
    process_window_w(f, wa, wb, win, blx, bly, img_h, variant, limit_search, plane_scratch, res, &wk);
    work_free(&wk);
}

/* image_utils.impl.h:394-419 (extract), widening the pixel type to double
 * as the reference's gf_image conversion does */
static void extract_window(const void* img, int dtype, size_t pitch, int32_t x0, int32_t y0,
                           uint32_t win, double* dst)
{
    for (uint32_t y = 0; y < win; ++y) {
        const size_t base = (size_t)(y0 + (int32_t)y) * pitch + (size_t)x0;
        double* d = dst + (size_t)y * win;
        switch (dtype) {
        case OREF_U8:  { const uint8_t* s = (const uint8_t*)img + base;  for (uint32_t x = 0; x < win; ++x) d[x] = (double)s[x]; } break;
        case OREF_U16: { const uint16_t* s = (const uint16_t*)img + base; for (uint32_t x = 0; x < win; ++x) d[x] = (double)s[x]; } break;
        case OREF_F32: { const float* s = (const float*)img + base;     for (uint32_t x = 0; x < win; ++x) d[x] = (double)s[x]; } break;
        default:       { const double* s = (const double*)img + base;   for (uint32_t x = 0; x < win; ++x) d[x] = s[x]; } break;
        }
    }
}

typedef struct {
    const oref_job* job;
    const oref_fft* fft;
    const int32_t* xy;
    size_t first, end;
    atomic_size_t* next;
    oref_vector* out;
} pool_arg;

static void* pool_worker(void* p)
{
    pool_arg* a = (pool_arg*)p;
    const oref_job* job = a->job;
    const uint32_t W = job->win;
    double* wa = (double*)malloc(sizeof(double) * W * W);
    double* wb = (double*)malloc(sizeof(double) * W * W);
    double* plane = (double*)malloc(sizeof(double) * W * W);
    work_t wk;
    work_init(&wk, W, W);
    const size_t chunk = 8;
    for (;;) {
        const size_t s = atomic_fetch_add(a->next, chunk);
        if (s >= a->end) break;
        const size_t e = s + chunk < a->end ? s + chunk : a->end;
        for (size_t i = s; i < e; ++i) {
            const int32_t x0 = a->xy[2 * i], y0 = a->xy[2 * i + 1];
            extract_window(job->a, job->dtype, job->pitch, x0, y0, W, wa);
            extract_window(job->b, job->dtype, job->pitch, x0, y0, W, wb);
            process_window_w(a->fft, wa, wb, W, x0, y0, job->img_h, job->variant,
                             job->limit_search, plane, &a->out[i - a->first], &wk);
        }
    }
    work_free(&wk);
    free(plane); free(wb); free(wa);
    return NULL;
}

long oref_process_pair(const oref_job* job, oref_vector* out)
{
    uint32_t nx, ny;
    int rc = oref_grid(job->img_w, job->img_h, job->win, job->win, job->off_x, job->off_y, &nx, &ny, NULL);
    if (rc) return rc;
    oref_fft* f = oref_fft_create(job->win, job->win);
    if (!f) return -10;
    const size_t total = (size_t)nx * ny;
    int32_t* xy = (int32_t*)malloc(sizeof(int32_t) * 2 * total);
    oref_grid(job->img_w, job->img_h, job->win, job->win, job->off_x, job->off_y, &nx, &ny, xy);
    size_t first = job->first < total ? job->first : total;
    size_t end = job->count ? first + job->count : total;
    if (end > total) end = total;

    atomic_size_t next;
    atomic_init(&next, first);
    pool_arg arg = { job, f, xy, first, end, &next, out };
    int nt = job->nthreads < 1 ? 1 : job->nthreads;
    if (nt == 1) {
        pool_worker(&arg);
    } else {
        pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * nt);
        for (int t = 0; t < nt; ++t) pthread_create(&th[t], NULL, pool_worker, &arg);
        for (int t = 0; t < nt; ++t) pthread_join(th[t], NULL);
        free(th);
    }
    free(xy);
    oref_fft_destroy(f);
    return (long)(end - first);
}

/* ------------------------------------------------------------------ */
/* adjoint of the per-window vector (no reference counterpart)        */
/* ------------------------------------------------------------------ */

/* d(num/den)/d(l, c, r) for num = ln l - ln r, den = 2 (ln l + ln r - 2 ln c) */
static void gauss3_grad(double l, double c, double r, double* gl, double* gc, double* gr)
{
    const double L = log(l), R = log(r), Cc = log(c);
    const double num = L - R, den = 2.0 * (L + R - 2.0 * Cc);
    if (den == 0.0) { *gl = *gc = *gr = 0.0; return; }
    const double d2 = den * den;
    *gl = (den - 2.0 * num) / d2 / l;
    *gr = (-den - 2.0 * num) / d2 / r;
    *gc = 4.0 * num / d2 / c;
}

int oref_window_jacobian(const oref_fft* f, const double* wa, const double* wb, uint32_t win,
                         int limit_search, double* ju_a, double* ju_b, double* jv_a, double* jv_b)
{
    const uint32_t W = win;
    const size_t n = (size_t)W * W;
    memset(ju_a, 0, n * sizeof(double));
    memset(ju_b, 0, n * sizeof(double));
    memset(jv_a, 0, n * sizeof(double));
    memset(jv_b, 0, n * sizeof(double));
    double* plane = (double*)malloc(sizeof(double) * n);
    oref_cross_correlate(f, wa, wb, plane);
    uint32_t rx = 0, ry = 0, rw = W, rh = W;
    if (limit_search) {
        int32_t dx, dy;
        oref_dilate(W, W, 0.5, &dx, &dy, &rw, &rh);
        rx = (uint32_t)dx;
        ry = (uint32_t)dy;
    }
    oref_peak peaks[2];
    const size_t np = oref_find_peaks(plane, W, W, rx, ry, rw, rh, 2, 1, peaks, NULL);
    if (np != 2) { free(plane); return -1; }
    const int32_t px = peaks[0].x, py = peaks[0].y;
    const double l = plane[(size_t)py * W + px - 1], c = plane[(size_t)py * W + px],
                 r = plane[(size_t)py * W + px + 1], t = plane[(size_t)(py - 1) * W + px],
                 bt = plane[(size_t)(py + 1) * W + px];
    free(plane);
    double fx, fy;
    {
        /* the forward fit must be finite for the derivative to exist */
        const double dx = (log(l) - log(r)), dxd = 2.0 * (log(l) + log(r) - 2.0 * log(c));
        const double dy = (log(t) - log(bt)), dyd = 2.0 * (log(t) + log(bt) - 2.0 * log(c));
        fx = dxd == 0.0 ? 0.0 : dx / dxd;
        fy = dyd == 0.0 ? 0.0 : dy / dyd;
        if (!isfinite(fx) || !isfinite(fy)) return -1;
    }
    /* u = W/2 - px - dx: du/dP = -d(dx)/dP; the plane holds W*W times the
     * correlation, a scale that cancels in the logarithmic ratios */
    double gl, gc_x, gr, gt, gc_y, gb;
    gauss3_grad(l, c, r, &gl, &gc_x, &gr);
    gauss3_grad(t, c, bt, &gt, &gc_y, &gb);
    const double scale = (double)n;
    /* shifts: swapped plane (h, w) holds C[(h + W/2) % W][(w + W/2) % W] */
    const uint32_t sy = ((uint32_t)py + W / 2) % W, sx = ((uint32_t)px + W / 2) % W;
    struct { uint32_t sy, sx; double gu, gv; } tap[5] = {
        { sy, (sx + W - 1) % W, -gl * scale, 0.0 },
        { sy, sx, -gc_x * scale, -gc_y * scale },
        { sy, (sx + 1) % W, -gr * scale, 0.0 },
        { (sy + W - 1) % W, sx, 0.0, -gt * scale },
        { (sy + 1) % W, sx, 0.0, -gb * scale },
    };
    for (int k = 0; k < 5; ++k) {
        for (uint32_t y = 0; y < W; ++y) {
            for (uint32_t x = 0; x < W; ++x) {
                const double bv = wb[(size_t)((y + tap[k].sy) % W) * W + (x + tap[k].sx) % W];
                const double av = wa[(size_t)((y + W - tap[k].sy) % W) * W + (x + W - tap[k].sx) % W];
                ju_a[(size_t)y * W + x] += tap[k].gu * bv;
                jv_a[(size_t)y * W + x] += tap[k].gv * bv;
                ju_b[(size_t)y * W + x] += tap[k].gu * av;
                jv_b[(size_t)y * W + x] += tap[k].gv * av;
            }
        }
    }
    return 0;
}

static long adjoint_pair(const oref_job* job, const double* cot, double* grad_a, double* grad_b,
                         double sigma, double* var)
{
    uint32_t nx, ny;
    int rc = oref_grid(job->img_w, job->img_h, job->win, job->win, job->off_x, job->off_y, &nx, &ny, NULL);
    if (rc) return rc;
    oref_fft* f = oref_fft_create(job->win, job->win);
    if (!f) return -10;
    const size_t total = (size_t)nx * ny;
    const uint32_t W = job->win;
    const size_t n = (size_t)W * W;
    int32_t* xy = (int32_t*)malloc(sizeof(int32_t) * 2 * total);
    oref_grid(job->img_w, job->img_h, W, W, job->off_x, job->off_y, &nx, &ny, xy);
    double* wa = (double*)malloc(sizeof(double) * n);
    double* wb = (double*)malloc(sizeof(double) * n);
    double* j = (double*)malloc(sizeof(double) * 4 * n);
    if (grad_a) memset(grad_a, 0, sizeof(double) * job->img_w * job->img_h);
    if (grad_b) memset(grad_b, 0, sizeof(double) * job->img_w * job->img_h);
    for (size_t i = 0; i < total; ++i) {
        const int32_t x0 = xy[2 * i], y0 = xy[2 * i + 1];
        extract_window(job->a, job->dtype, job->pitch, x0, y0, W, wa);
        extract_window(job->b, job->dtype, job->pitch, x0, y0, W, wb);
        const int ok = oref_window_jacobian(f, wa, wb, W, job->limit_search, j, j + n, j + 2 * n, j + 3 * n) == 0;
        if (var) {
            double su = 0.0, sv = 0.0;
            if (ok)
                for (size_t k = 0; k < n; ++k) {
                    su += j[k] * j[k] + j[n + k] * j[n + k];
                    sv += j[2 * n + k] * j[2 * n + k] + j[3 * n + k] * j[3 * n + k];
                }
            var[2 * i] = sigma * sigma * su;
            var[2 * i + 1] = sigma * sigma * sv;
        }
        if (grad_a && ok) {
            const double cu = cot[2 * i], cv = cot[2 * i + 1];
            for (uint32_t y = 0; y < W; ++y)
                for (uint32_t x = 0; x < W; ++x) {
                    const size_t g = (size_t)(y0 + (int32_t)y) * job->img_w + (size_t)(x0 + (int32_t)x);
                    const size_t k = (size_t)y * W + x;
                    grad_a[g] += cu * j[k] + cv * j[2 * n + k];
                    grad_b[g] += cu * j[n + k] + cv * j[3 * n + k];
                }
        }
    }
    free(j); free(wb); free(wa); free(xy);
    oref_fft_destroy(f);
    return (long)total;
}

long oref_vjp_pair(const oref_job* job, const double* cot, double* grad_a, double* grad_b)
{
    return adjoint_pair(job, cot, grad_a, grad_b, 0.0, NULL);
}

long oref_variance_pair(const oref_job* job, double sigma, double* var)
{
    return adjoint_pair(job, NULL, NULL, NULL, sigma, var);
}


/* ---- background subtraction (examples/average_subtract/main.cpp:98-159) ---- */

int oref_average_subtract(const double* frames, size_t n, uint32_t w, uint32_t h, double* mean, uint16_t* out)
{
    if (n < 2) return -1;
    const size_t np = (size_t)w * h;
    /* main.cpp:98-122: avg starts at zero, frames 1..n-1 added in order */
    for (size_t p = 0; p < np; ++p) mean[p] = 0.0;
    size_t processed = 0;
    for (size_t i = 1; i < n; ++i) {
        const double* f = frames + i * np;
        for (size_t p = 0; p < np; ++p) mean[p] = mean[p] + f[p];
        ++processed;
    }
    /* main.cpp:131: avg / g_f{processed} */
    const double div = (double)processed;
    for (size_t p = 0; p < np; ++p) mean[p] = mean[p] / div;
    double* d = (double*)malloc(np * sizeof(double));
    if (!d) return -1;
    for (size_t i = 0; i < n; ++i) {
        const double* f = frames + i * np;
        for (size_t p = 0; p < np; ++p) d[p] = f[p] - mean[p];          /* main.cpp:146 */
        /* stats.h:22-31 */
        double mn = d[0], mx = d[0];
        for (size_t p = 1; p < np; ++p) {
            mn = d[p] < mn ? d[p] : mn;
            mx = d[p] > mx ? d[p] : mx;
        }
        /* pnm_image_loader.cpp:369-391 */
        double range = mx - mn;
        if (mx == mn) range = 1.0;
        uint16_t* o = out + i * np;
        for (size_t p = 0; p < np; ++p) o[p] = (uint16_t)(65535.0 * ((d[p] - mn) / range));
    }
    free(d);
    return 0;
}
