/*
 * piv_oracle.h -- FP64 CPU restatement of the openpiv-cxx window-correlation
 * path.  TEST INFRASTRUCTURE ONLY: this is the checker that the HIP engine is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links it.
 *
 * Every function follows the reference operation by operation so that, built
 * with -ffp-contract=off on baseline x86-64 (no FMA), its results are
 * bit-identical to the reference's FP64 path.  Citations are to files of the
 * reference (ErichZimmer/openpiv-cxx, version 0.2.0).
 */
#ifndef PIV_ORACLE_H
#define PIV_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } oref_c;

enum { OREF_FORWARD = 0, OREF_REVERSE = 1 };

/* correlator variants (examples/process/main.cpp:153-178) */
enum {
    OREF_CORR_COMPLEX = 0,   /* algos::FFT::cross_correlate (also pocket, pocket_real: same math) */
    OREF_CORR_REAL    = 1    /* algos::FFT::cross_correlate_real (zeroed-bin unpack quirk) */
};

/* pixel types accepted by oref_process_pair (all widened to double, as the
 * reference's loaders widen to gf_image) */
enum { OREF_U8 = 0, OREF_U16 = 1, OREF_F32 = 2, OREF_F64 = 3 };

/* per-window flags */
enum {
    OREF_FLAG_FEWER_THAN_2_PEAKS = 1u << 0,  /* main.cpp:214-218: vector left zero */
    OREF_FLAG_NONFINITE_FIT      = 1u << 1,  /* log of a value <= 0 in the fit */
    OREF_FLAG_EXACT_TIE          = 1u << 3   /* top-2 peak values equal: std::sort order unpinned */
};

/* ---- grid: openpiv/core/detail/grid.impl.h ---- */

/* grid.impl.h:10-22 -- step = (uint32)(win * fraction); 0 on success,
 * -1 when fraction is outside [0,1] */
int oref_grid_offsets_from_fraction(uint32_t win_w, uint32_t win_h, double fraction,
                                    uint32_t* off_x, uint32_t* off_y);

/* grid.impl.h:24-61 -- writes count rects (bottom-left x,y pairs) in
 * row-major y-outer order; returns 0, or -1..-4 for the reference's four
 * error conditions (zero image, zero window, zero offset, window > image).
 * When xy is NULL only *nx, *ny are computed. */
int oref_grid(uint32_t img_w, uint32_t img_h, uint32_t win_w, uint32_t win_h,
              uint32_t off_x, uint32_t off_y, uint32_t* nx, uint32_t* ny, int32_t* xy);

/* ---- FFT: openpiv/algos/fft.h ---- */

typedef struct oref_fft oref_fft;

/* fft.h:76-84 + 274-294; returns NULL if a dimension is not a power of 2 */
oref_fft* oref_fft_create(uint32_t w, uint32_t h);
void oref_fft_destroy(oref_fft* f);

/* fft.h:86-119 -- un-normalised 2-D DFT of a w*h complex image (row-major) */
void oref_fft_transform(const oref_fft* f, const oref_c* in, oref_c* out, int dir);

/* fft.h:121-186 -- two-real-image transform with the reference's unpack */
void oref_fft_transform_real(const oref_fft* f, const double* a, const double* b,
                             oref_c* out_a, oref_c* out_b);

/* fft.h:188-206 -- real(IFFT(FFT(b) * conj(FFT(a)))), quadrant-swapped */
void oref_cross_correlate(const oref_fft* f, const double* a, const double* b, double* out);

/* fft.h:208-227 */
void oref_cross_correlate_real(const oref_fft* f, const double* a, const double* b, double* out);

/* fft.h:229-242 (returns the real part; the reference returns it as complex
 * with zero imaginary parts) */
void oref_auto_correlate(const oref_fft* f, const double* a, double* out);

/* ---- image utils: openpiv/core/detail/image_utils.impl.h ---- */

/* :364-386 */
void oref_swap_quadrants(double* im, uint32_t w, uint32_t h);

typedef struct { int32_t x, y; double value; } oref_peak;

/* :37-74 over the sub-region (rx, ry, rw, rh) of a w-wide plane (the region
 * is the whole plane, or an image_view such as the limit-search centre).
 * Peak coordinates are plane coordinates of the 3x3 view midpoint.  Equal
 * values keep scan order; *exact_tie is set when the kept top-2 tie. */
size_t oref_find_peaks(const double* plane, uint32_t w, uint32_t h,
                       uint32_t rx, uint32_t ry, uint32_t rw, uint32_t rh,
                       uint16_t num_peaks, uint32_t radius,
                       oref_peak* out, int* exact_tie);

/* :76-101 -- fit around plane centre (cx, cy) */
void oref_fit_simple_gaussian(const double* plane, uint32_t w, int32_t cx, int32_t cy,
                              double* fx, double* fy);

/* rect.cpp:109-126 rect::dilate(double) applied to a rect at origin 0 */
void oref_dilate(uint32_t w, uint32_t h, double d, int32_t* x, int32_t* y,
                 uint32_t* rw, uint32_t* rh);

/* ---- the per-window processor of examples/process/main.cpp:189-237 ---- */

typedef struct {
    double x, y;          /* window midpoint, y flipped (main.cpp:226-230) */
    double u, v;          /* displacement vector (main.cpp:227) */
    double sn;            /* peak0 / peak1 when peak1 > 0 (main.cpp:233-234) */
    int32_t px, py;       /* integer peak (quadrant-swapped plane coords) */
    double peak0, peak1;  /* the two peak values */
    uint32_t flags;
    uint32_t pad_;
} oref_vector;

typedef struct {
    const void* a;        /* frame A */
    const void* b;        /* frame B */
    int dtype;            /* OREF_U8 .. OREF_F64 */
    uint32_t img_w, img_h;
    size_t pitch;         /* elements per row */
    uint32_t win;         /* square window side (power of 2) */
    uint32_t off_x, off_y;/* grid offsets (steps) */
    int variant;          /* OREF_CORR_COMPLEX / OREF_CORR_REAL */
    int limit_search;     /* main.cpp:204-211 */
    int nthreads;         /* <=1: single thread */
    size_t first, count;  /* window range to process (count 0 = all) */
} oref_job;

/* Returns number of windows written to out (grid order), or <0 on error. */
long oref_process_pair(const oref_job* job, oref_vector* out);

/* one window from already-extracted W*W double windows */
void oref_process_window(const oref_fft* f, const double* wa, const double* wb, uint32_t win,
                         int32_t blx, int32_t bly, uint32_t img_h, int variant,
                         int limit_search, double* plane_scratch, oref_vector* res);

/* ---- adjoint (no reference counterpart; pinned by FP64 central finite
 * differences of oref_process_window) ----
 *
 * Jacobian of one window's vector (u, v) (main.cpp:227) with respect to the
 * window's pixels of both frames, in FP64: ju_a[y*W+x] = du/da[y][x], etc.
 * u = W/2 - px - dx with dx = (ln l - ln r) / (2 (ln l + ln r - 2 ln c))
 * (image_utils.impl.h:86-94), and each stencil value is the circular
 * correlation C[s] = sum_n a[n] b[n+s] at a fixed shift s, so
 * dC[s]/da[n] = b[n+s] and dC[s]/db[m] = a[m-s].  The integer peak is held
 * fixed (the Jacobian of the argmax is zero almost everywhere).  Returns 0,
 * or -1 when the window has fewer than two peaks or a non-finite fit. */
int oref_window_jacobian(const oref_fft* f, const double* wa, const double* wb, uint32_t win,
                         int limit_search, double* ju_a, double* ju_b, double* jv_a, double* jv_b);

/* sum over windows of cot[2i] * d(u_i) + cot[2i+1] * d(v_i), accumulated
 * into frame-sized gradients (FP64, row-major, pitch = img_w) */
long oref_vjp_pair(const oref_job* job, const double* cot, double* grad_a, double* grad_b);

/* per-window variance of (u, v) under i.i.d. pixel noise sigma on both
 * frames: var[2i] = sigma^2 * sum(ju_a^2 + ju_b^2), var[2i+1] likewise
 * (0 for windows without a valid fit) */
long oref_variance_pair(const oref_job* job, double sigma, double* var);

/* ---- background subtraction: examples/average_subtract/main.cpp:98-159 ----
 * frames: n row-major w*h images (pixel values as loaded, FP64).
 * mean = (frames[1] + ... + frames[n-1]) / (n - 1): the reference sizes its
 * accumulator from the first input and adds every later one, in input order
 * (main.cpp:98-122).  For every frame i, d = frames[i] - mean (main.cpp:146)
 * and out[i] = the 16-bit image the PNM writer saves for d
 * (pnm_image_loader.cpp:367-392): (uint16)(65535 * ((d - min) / range)),
 * min/max by algos::find_image_range (stats.h:15-34), range = max - min, or 1
 * when max == min.  Returns -1 when n < 2 (the reference's "failed to load
 * any files"), else 0. */
int oref_average_subtract(const double* frames, size_t n, uint32_t w, uint32_t h, double* mean, uint16_t* out);

#ifdef __cplusplus
}
#endif
#endif
