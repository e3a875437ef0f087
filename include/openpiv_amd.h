/*
 * openpiv_amd.h -- C ABI of the MI355X (gfx950) PIV window-correlation engine.
 *
 * This is the drop-in boundary for the hot path of openpiv-cxx
 * (ErichZimmer/openpiv-cxx 0.2.0): per-window FFT circular cross-correlation,
 * find_peaks(.., 2, 1) and the 3-point Gaussian fit, run over the windows of
 * generate_cartesian_grid, as examples/process/main.cpp:189-302 drives them.
 * The reference exposes that path as C++ templates; each entry point below
 * names the reference interface it replaces.
 *
 * Conventions: every function returns an opiv_status (0 = OPIV_OK); no C++
 * exception crosses this boundary.  On failure opiv_last_error() returns the
 * message of the failing call on the calling thread; messages that mirror a
 * reference exception carry the reference's text so a C++ wrapper can rethrow
 * it verbatim.  Frame pointers may be host or device memory.
 */
#ifndef OPENPIV_AMD_H
#define OPENPIV_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum opiv_status {
    OPIV_OK = 0,
    OPIV_E_INVALID = 1,      /* bad argument (null pointer, zero size, bad enum) */
    OPIV_E_NOT_POW2 = 2,     /* fft.h:82-83 "dimensions must be power of 2: ..." */
    OPIV_E_SIZE = 3,         /* fft.h:94-98 "image size is different from expected: ..." */
    OPIV_E_GRID = 4,         /* grid.impl.h:14-40 grid errors */
    OPIV_E_HIP = 5,          /* a HIP runtime call failed */
    OPIV_E_UNSUPPORTED = 6,  /* configuration outside the engine's range */
    OPIV_E_NO_DEVICE = 7     /* no usable gfx950 device */
} opiv_status;

typedef enum opiv_dtype {   /* pixel types of the frames (loaders deliver g_8 / g_16) */
    OPIV_U8 = 0,
    OPIV_U16 = 1,
    OPIV_F32 = 2,
    OPIV_F64 = 3             /* core::gf_image; widened pixels are narrowed to FP32 on load */
} opiv_dtype;

typedef enum opiv_variant {
    /* exact circular cross-correlation: algos::FFT::cross_correlate
     * (fft.h:188-206); pocket / pocket_real (pocket_fft.h:236-275) compute
     * the same correlation.  FP32 arithmetic, FP64 Gaussian fit. */
    OPIV_CIRCULAR_FFT_FP32 = 0,
    /* algos::FFT::cross_correlate_real (fft.h:208-227) including its
     * unpack (fft.h:166-183) that leaves spectral row 0, row H/2 and
     * column 0 zero */
    OPIV_REAL_COMPAT = 1,
    /* the circular correlation computed directly (no FFT) as an im2col GEMM
     * on MFMA.  16x16 windows: u8 frames exactly on int8 MFMA (pixels
     * offset by 128), u16 exact bf16 products, f32/f64 a three-term bf16
     * split, FP32 accumulation.  32x32 windows of u8 and u16 frames: exactly
     * on int8 MFMA (16-bit pixels as two offset byte planes, four int8
     * products).  Other sizes and types: OPIV_E_UNSUPPORTED. */
    OPIV_DIRECT_BF16_SPLIT = 2,
    /* the faster of OPIV_CIRCULAR_FFT_FP32 and OPIV_DIRECT_BF16_SPLIT for
     * the plan's window size and pixel type as measured on MI355X at
     * 4096x4096 (same semantics): the direct path for 8-bit frames at 16x16
     * (68.5 against 74.9 us), the FFT path otherwise (8-bit at 32x32: 81.6
     * against 88.0 us) */
    OPIV_AUTO = 3
} opiv_variant;

typedef enum opiv_flag {
    OPIV_FLAG_FEWER_THAN_2_PEAKS = 1u << 0, /* main.cpp:214-218: record left all-zero */
    OPIV_FLAG_NONFINITE_FIT = 1u << 1,      /* a stencil value <= 0 made the fit NaN/inf */
    OPIV_FLAG_NEAR_TIE = 1u << 2,           /* top-2 (or peak vs neighbour) closer than the FP32 tolerance */
    OPIV_FLAG_EXACT_TIE = 1u << 3           /* top-2 values equal (reference order is std::sort's) */
} opiv_flag;

/* One vector of examples/process (main.cpp:144-150 point_vector: xy, vxy,
 * sn) plus the integer peak, both peak heights and flags. */
typedef struct opiv_window_result {
    double x, y;       /* window midpoint; y = image height - mid_y (main.cpp:226-230) */
    double u, v;       /* mid - (bottom_left + fitted peak) (main.cpp:227) */
    double sn;         /* peak0 / peak1 if peak1 > 0, else 0 (main.cpp:233-234) */
    double peak0;      /* correlation value at the integer peak */
    double peak1;      /* second highest local maximum */
    int32_t px, py;    /* integer peak in the quadrant-swapped plane */
    uint32_t flags;    /* opiv_flag bits */
    uint32_t reserved;
} opiv_window_result;

typedef struct opiv_ctx opiv_ctx;
typedef struct opiv_plan opiv_plan;

/* Plan for one frame size / window size / grid; the batched replacement of
 * `static algos::FFT fft{ia}` + the grid of main.cpp:137-138. */
typedef struct opiv_plan_desc {
    uint32_t img_w, img_h;   /* frame size in pixels */
    size_t pitch;            /* elements per frame row (0 = img_w) */
    int32_t dtype;           /* opiv_dtype */
    uint32_t win;            /* square window side, power of 2 in [8, 128] */
    uint32_t step_x, step_y; /* grid offsets in pixels (grid.impl.h:24-61); 0 = use overlap */
    double overlap;          /* process -o: step = (uint32)(win * overlap) (grid.impl.h:10-22) */
    int32_t variant;         /* opiv_variant */
    int32_t limit_search;    /* process -l: search the dilate(0.5) centre (main.cpp:204-211) */
} opiv_plan_desc;

/* Message of the last failing call on this thread ("" if none). */
const char* opiv_last_error(void);

/* Library version string. */
const char* opiv_version(void);

/* Context: one device, one HIP stream.  A context must not be used from two
 * host threads at once without external locking. */
int opiv_ctx_create(int device, opiv_ctx** out);
int opiv_ctx_destroy(opiv_ctx* ctx);
int opiv_ctx_synchronize(opiv_ctx* ctx);
/* The hipStream_t every launch of this context is enqueued on. */
void* opiv_ctx_stream(opiv_ctx* ctx);

/* core::generate_cartesian_grid (grid.impl.h:10-61), host only.  Writes
 * bottom-left (x, y) pairs in the reference's y-outer order when xy != NULL. */
int opiv_grid_offsets(uint32_t win_w, uint32_t win_h, double overlap,
                      uint32_t* off_x, uint32_t* off_y);
int opiv_grid(uint32_t img_w, uint32_t img_h, uint32_t win_w, uint32_t win_h,
              uint32_t off_x, uint32_t off_y, uint32_t* nx, uint32_t* ny, int32_t* xy);

/* Plan creation checks the window the way algos::FFT's constructor does
 * (fft.h:76-84) and builds the grid (main.cpp:137-138). */
int opiv_plan_create(opiv_ctx* ctx, const opiv_plan_desc* desc, opiv_plan** out);
int opiv_plan_destroy(opiv_plan* plan);
int opiv_plan_window_count(const opiv_plan* plan, size_t* count);
int opiv_plan_grid(const opiv_plan* plan, int32_t* xy);

/* The whole processing loop of examples/process/main.cpp:241-302 for one
 * frame pair: every window's correlation, find_peaks(.., 2, 1),
 * fit_simple_gaussian and vector, in grid order.  a, b, out may be host or
 * device pointers; n must equal the plan's window count.  Device frames (here
 * and in every entry point below) must be aligned to their pixel size
 * (OPIV_E_INVALID otherwise).  Returns when the results are in `out`. */
int opiv_run_pair(opiv_plan* plan, const void* a, const void* b,
                  opiv_window_result* out, size_t n);

/* Same, device pointers only, enqueued on the context stream without
 * synchronising (the caller synchronises). */
int opiv_run_pair_async(opiv_plan* plan, const void* d_a, const void* d_b,
                        opiv_window_result* d_out, size_t n);

/* correlator_t of main.cpp:153 for a batch of n windows:
 * FFT::cross_correlate (fft.h:188-206) or cross_correlate_real
 * (fft.h:208-227) of stacked win*win FP64 windows a[i], b[i] (win 8 to
 * 128); writes the quadrant-swapped real planes (FP64, computed in FP32).
 * Host or device pointers. */
int opiv_correlate_windows(opiv_ctx* ctx, uint32_t win, size_t n, int32_t variant,
                           const double* a, const double* b, double* planes);

/* FFT::auto_correlate (fft.h:229-242) for n stacked windows (real part of
 * the returned plane). */
int opiv_auto_correlate_windows(opiv_ctx* ctx, uint32_t win, size_t n,
                                const double* a, double* planes);

/* algos::direction (fft_common.h:9-12) */
typedef enum opiv_direction {
    OPIV_FORWARD = 0,        /* exp(-i 2 pi k n / N) */
    OPIV_REVERSE = 1         /* exp(+i 2 pi k n / N), un-normalised */
} opiv_direction;

/* FFT::transform (fft.h:86-119): un-normalised 2-D DFT of n stacked
 * win*win complex windows (interleaved re, im FP64, row-major; win 8 to
 * 128), computed in FP32.  in == out is allowed.  Host or device pointers. */
int opiv_fft_transform_windows(opiv_ctx* ctx, uint32_t win, size_t n, int32_t direction,
                               const double* in, double* out);

/* ---- frame ingest: the image loaders ----
 *
 * loaders/image_loader.cpp:23-57 (format from the first bytes),
 * pnm_image_loader.cpp:279-310 and tiff_image_loader.cpp:160-236 (read
 * without libtiff): PNM P5 (grey) and P6 (RGB); TIFF strips of one or three
 * samples of 8 or 16 bits, uncompressed, Deflate, LZW or PackBits (the
 * horizontal predictor with LZW and Deflate), every page.  RGB becomes grey
 * by the reference's integer rule (218 r + 732 g + 74 b) >> 10
 * (pixel_types.h:397-400).  Frames are delivered in their own depth
 * (OPIV_U8 / OPIV_U16; 8-bit P6 as OPIV_U16, since the reference reads its
 * samples through a signed char) -- no widening.  Errors keep the
 * reference's texts ("failed to open ...", "failed to find loader for ...",
 * "image index out of range ..."). */
typedef struct opiv_image_info {
    uint32_t width, height;
    int32_t dtype;           /* OPIV_U8 or OPIV_U16 */
    uint32_t count;          /* images in the file (TIFF pages; 1 for PNM) */
} opiv_image_info;

int opiv_image_probe(const char* path, opiv_image_info* info);

/* image `index` of path into dst (pitch elements per row, 0 = width) in the
 * file's pixel type; dst host or device memory (device: copied on ctx's
 * stream, ctx required; host: ctx may be NULL).  Returns when written. */
int opiv_image_read(opiv_ctx* ctx, const char* path, uint32_t index, void* dst, size_t pitch);

/* ---- background subtraction: examples/average_subtract/main.cpp:98-159 ----
 *
 * For n >= 2 frames of one size and pixel type (frames[i]: host or device,
 * pitch elements per row, 0 = img_w): mean = (frame 1 + ... + frame n-1) /
 * (n - 1) -- the reference sizes its accumulator from the first input and
 * averages every later one (main.cpp:98-131) -- and for every frame i the
 * 16-bit image the reference's PNM writer saves for frame_i - mean
 * (pnm_image_loader.cpp:367-392: (uint16)(65535 * ((d - min) / range)),
 * range = max - min, or 1 when equal).  out[i]: img_w*img_h uint16,
 * row-major, host or device; mean: img_w*img_h doubles or NULL.  FP64
 * arithmetic in the reference's order: bit-identical outputs.  Returns when
 * the outputs are written. */
int opiv_average_subtract(opiv_ctx* ctx, const void* const* frames, size_t n, uint32_t img_w, uint32_t img_h,
                          size_t pitch, int32_t dtype, uint16_t* const* out, double* mean);

/* ---- adjoint: displacement sensitivities (no reference counterpart) ----
 *
 * Derivatives of each window's vector (u, v) of main.cpp:227 with respect to
 * the pixels of both frames, with the integer peak held fixed.  The adjoint
 * of the circular correlation at the five stencil shifts is again a
 * correlation; the fit's derivatives use FP64 logs of stencil values
 * recomputed as direct dot products.  Only OPIV_CIRCULAR_FFT_FP32 plans of
 * windows 8 to 64.
 * fwd: the plan's results of opiv_run_pair for the same frames (NULL = run
 * the forward pass here).  Windows flagged FEWER_THAN_2_PEAKS or
 * NONFINITE_FIT contribute nothing.  All pointers host or device; returns
 * when the outputs are written. */

/* grad_a, grad_b (frame-sized, pitch * img_h floats) =
 * sum_i cot[2i] * du_i/d(frame) + cot[2i+1] * dv_i/d(frame) */
int opiv_vjp_pair(opiv_plan* plan, const void* a, const void* b, const opiv_window_result* fwd,
                  const double* cot, float* grad_a, float* grad_b);

/* Same, device pointers only and fwd required, enqueued on the context
 * stream without synchronising (the caller synchronises), as
 * opiv_run_pair_async is to opiv_run_pair. */
int opiv_vjp_pair_async(opiv_plan* plan, const void* d_a, const void* d_b, const opiv_window_result* d_fwd,
                        const double* d_cot, float* d_grad_a, float* d_grad_b);

/* Allocation-free form of opiv_vjp_pair_async.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, kernels and results; the per-window tap weights go to a
 * fixed scratch array of OPIV_NOALLOC_MAX_WINDOWS windows (32 B each, 8 MiB),
 * one per GPU, which each GPU the library runs on holds until the process
 * ends, instead of a buffer the context grows.  Calls no hipMalloc or
 * hipFree.  Plans with more windows get OPIV_E_UNSUPPORTED.  Every call on a
 * GPU shares that GPU's array, so calls that overlap on one GPU (from
 * different contexts or streams) race; calls on one context's stream are
 * ordered and safe.  The array keeps the last call's tap weights; every call
 * writes all of its windows' weights before reading them. */
#define OPIV_NOALLOC_MAX_WINDOWS 262144
int opiv_vjp_pair_async_noalloc(opiv_plan* plan, const void* d_a, const void* d_b, const opiv_window_result* d_fwd,
                                const double* d_cot, float* d_grad_a, float* d_grad_b);

/* var[2i], var[2i+1] = sigma^2 * |d(u_i, v_i)/d(both frames)|^2: the
 * variance of (u_i, v_i) under i.i.d. pixel noise sigma (linear propagation) */
int opiv_variance_pair(opiv_plan* plan, const void* a, const void* b, const opiv_window_result* fwd,
                       double sigma, double* var);

/* ---- batches: many frame pairs and several window passes per launch ----
 *
 * The reference processes one pair and one window size per run of
 * examples/process (main.cpp:137-302) and has no multipass scheme
 * (README.md:280: independent passes).  A batch holds npairs pairs of
 * uneven sizes and up to 4 independent passes (e.g. 64, 32, 16); each pass
 * is one persistent kernel launch over the windows of every pair, and the
 * passes are captured once into a HIP graph that opiv_batch_run replays. */

typedef struct opiv_pair_desc {
    const void* a;           /* frame A, device memory */
    const void* b;           /* frame B, device memory */
    uint32_t img_w, img_h;
    size_t pitch;            /* elements per row (0 = img_w) */
} opiv_pair_desc;

typedef struct opiv_batch_desc {
    int32_t dtype;           /* opiv_dtype shared by all pairs */
    uint32_t npasses;        /* 1..4 */
    uint32_t win[4];         /* square window per pass, power of 2 in [8, 128] */
    double overlap[4];       /* process -o per pass (grid.impl.h:10-22) */
    int32_t variant;         /* opiv_variant */
    int32_t limit_search;
} opiv_batch_desc;

typedef struct opiv_batch opiv_batch;

/* Builds every (pass, pair) grid, uploads the pair table, and picks each
 * pass's launch shape (threads per block, persistent blocks) from the tuning
 * cache, or by timing the candidates on these frames when the cache has no
 * entry.  The shape never changes a window's arithmetic, so results do not
 * depend on it.  With two passes or more and every window at most 64 x 64
 * it also times all passes in one persistent launch (the multi-pass kernel)
 * against one launch per pass and keeps the faster form;
 * OPIV_BATCH_FORM=single|perpass forces one.  The cache is the text file $OPIV_TUNE_CACHE, else
 * $HOME/.cache/openpiv_amd/tune.txt (one line per key: arch, window, pixel
 * type and log2 window count -> threads and blocks per CU x 4; arch, "form",
 * pixel type, passes, log2 windows -> single launch and its blocks per CU x
 * 4); OPIV_TUNE=0 disables timing. */
int opiv_batch_create(opiv_ctx* ctx, const opiv_batch_desc* desc, const opiv_pair_desc* pairs,
                      size_t npairs, opiv_batch** out);
int opiv_batch_destroy(opiv_batch* batch);
/* total windows over all passes and pairs */
int opiv_batch_window_count(const opiv_batch* batch, size_t* count);
/* npasses * npairs + 1 prefix offsets: results of (pass p, pair i) start at
 * offsets[p * npairs + i], in that pair's grid order */
int opiv_batch_offsets(const opiv_batch* batch, size_t* offsets);
/* new frame pointers for the same sizes (takes effect at the next run) */
int opiv_batch_set_pairs(opiv_batch* batch, const opiv_pair_desc* pairs, size_t npairs);
/* the chosen launch shape of pass p: threads per block and block count */
int opiv_batch_shape(const opiv_batch* batch, uint32_t pass, int32_t* threads, int32_t* blocks);
/* the launch form: *single = 1 when every pass runs in one persistent launch
 * of *blocks blocks, 0 for one launch per pass */
int opiv_batch_form(const opiv_batch* batch, int32_t* single, int32_t* blocks);
/* Enqueue every pass as one graph launch on the context stream; d_out is
 * device memory for opiv_batch_window_count results.  Does not synchronise. */
int opiv_batch_run(opiv_batch* batch, opiv_window_result* d_out, size_t n);

#ifdef __cplusplus
}
#endif
#endif
