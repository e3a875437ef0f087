// corr_fft.hip -- fused FP32 FFT window correlation, peak search and
// Gaussian fit for gfx950.
//
// Replaces, per window, the chain of examples/process/main.cpp:189-237:
//   core::extract (image_utils.impl.h:394-419)
//   algos::FFT::cross_correlate (fft.h:188-206) or cross_correlate_real
//   (fft.h:208-227), including swap_quadrants (image_utils.impl.h:364-386)
//   core::find_peaks(plane, 2, 1) (image_utils.impl.h:37-74)
//   core::fit_simple_gaussian (image_utils.impl.h:76-101)
//   the vector / s/n assembly of main.cpp:220-236
// with the correlation plane kept in registers: it never reaches HBM.
//
// Work mapping.  A window of W x W pixels is owned by W consecutive lanes of
// one wavefront (64 / W lane groups per wave).  Lane t loads pixel column t of
// both frames (coalesced: the W lanes read W consecutive pixels of a row),
// packs z = a + i b, and runs a W-point FFT down its column in registers.
// One LDS transpose gives each lane a spectral row for the second FFT; an
// LDS mirror read fetches Z(-k) so the two real spectra separate as
// A = (Z + conj Z(-k)) / 2, B = (Z - conj Z(-k)) / 2i, and the product
// P = B conj(A) is formed in registers.  The inverse runs rows then (after a
// second transpose) columns; the transpose hands lane t the plane column
// that lands in quadrant-swapped column t, so the swapped plane is again a
// register renaming.  find_peaks then runs on registers with DPP neighbour
// moves, a cross-lane top-2 reduction follows, and the fit uses FP64 logs.
// For W <= 32 each lane group runs two windows: the second window's spectrum
// is added as i * P1 to the first's, so one inverse FFT yields both real
// planes (real and imaginary parts), and the FP64 logs of both stencils are
// spread over the group's lanes, one log per lane.
// LDS traffic stays inside a wave, so only wave-level syncs are needed.
#include <hip/hip_runtime.h>

#include <type_traits>

#include <cmath>
#include <cstdint>
#include <limits>

#include "corr_common.hpp"
#include "corr_fft.hpp"
#include "fft_reg.hpp"

namespace opiv {

using fftreg::brev_n;
using fftreg::sfor;

// The frames of one wave task as buffers.  All windows of a task lie in one
// frame pair, and the element offset y0 * pitch + x0 grows with the window
// index, so lane 0's first window has the task's smallest: the buffers start
// at the frames' pixel there (wave-uniform), a lane adds its own window and
// column as a 32-bit byte offset, and row y adds y * pitch bytes as a
// uniform offset -- no per-lane 64-bit address arithmetic per load.
struct TaskFrames {
    __amdgpu_buffer_rsrc_t a, b;
    uint64_t base;                // element offset of the buffer starts
    uint32_t pitch_bytes;
};

template <int DT>
constexpr uint32_t px_bytes() { return DT == 0 ? 1 : DT == 1 ? 2 : DT == 2 ? 4 : 8; }

// Word 3 of a buffer resource.  f32/f64 frames use untyped loads (the data
// format field only has to be valid: 32 bits).  8- and 16-bit frames use
// typed loads of one component (destination X = component X) in the
// unsigned-scaled number format, which return the pixel as a float: the
// conversion happens in the load path instead of a VALU conversion (a
// slow-class instruction per pixel).
template <int DT>
constexpr int rsrc_word3()
{
    constexpr int kDstSelX = 4, kNumFormatUscaled = 2 << 12;
    constexpr int kDataFormat8 = 1 << 15, kDataFormat16 = 2 << 15, kDataFormat32 = 4 << 15;
    return DT == 0 ? kDstSelX | kNumFormatUscaled | kDataFormat8
         : DT == 1 ? kDstSelX | kNumFormatUscaled | kDataFormat16
                   : kDataFormat32;
}

// typed buffer load of one component (the LLVM intrinsic; clang has no
// builtin for it)
extern "C" __device__ float opiv_buffer_load_format_f32(__amdgpu_buffer_rsrc_t rsrc, int voffset, int soffset,
                                                         int aux) __asm("llvm.amdgcn.raw.ptr.buffer.load.format.f32");

__device__ __forceinline__ uint64_t first_lane(uint64_t v)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// g0: this lane's first window; called with every lane of the wave active
template <int DT>
__device__ __forceinline__ TaskFrames task_frames(const WinGeom& g0)
{
    constexpr uint32_t ES = px_bytes<DT>();
    TaskFrames f;
    f.base = first_lane((uint64_t)g0.y0 * g0.pitch + (uint64_t)g0.x0);
    const uint64_t pitch = first_lane(g0.pitch);
    const uint64_t a = first_lane((uint64_t)(uintptr_t)g0.a) + f.base * ES;
    const uint64_t b = first_lane((uint64_t)(uintptr_t)g0.b) + f.base * ES;
    f.a = __builtin_amdgcn_make_buffer_rsrc((void*)(uintptr_t)a, 0, -1, rsrc_word3<DT>());
    f.b = __builtin_amdgcn_make_buffer_rsrc((void*)(uintptr_t)b, 0, -1, rsrc_word3<DT>());
    f.pitch_bytes = (uint32_t)(pitch * ES);
    return f;
}

template <int DT>
__device__ __forceinline__ float load_buf(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t soff)
{
    if constexpr (DT <= 1) return opiv_buffer_load_format_f32(r, (int)off, (int)soff, 0);
    else if constexpr (DT == 2) return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, soff, 0));
    else return (float)__builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, 0));
}

// column t of both frames of window g: z[y] = a + i b
template <int W, int DT>
__device__ __forceinline__ void load_column(const TaskFrames& f, const WinGeom& g, int t, float2 (&z)[W])
{
    const uint32_t off = (uint32_t)(((uint64_t)g.y0 * g.pitch + (uint64_t)(g.x0 + t) - f.base) * px_bytes<DT>());
    sfor<W>([&](auto Y) {
        constexpr uint32_t y = decltype(Y)::value;
        const uint32_t so = y * f.pitch_bytes;
        z[y] = make_float2(load_buf<DT>(f.a, off, so), load_buf<DT>(f.b, off, so));
    });
}

// From 16 x 16 up the lane group's LDS slice holds floats, and every
// complex exchange goes through it in two passes (real parts, then imaginary
// parts): half the LDS, so that six waves (16 x 16), four (32 x 32) or two
// (64 x 64) fit per SIMD.  The kernel is bound by VALU issue; the resident
// waves fill each other's waits on LDS round trips and loads.
template <int W>
constexpr bool split_lds() { return W >= 16; }
template <int W>
using LdsT = std::conditional_t<split_lds<W>(), float, float2>;

// Complex exchange through the lane group's LDS slice: this lane's values
// src(i) go to lds[wi(i)] (i = 0..NW-1), then dst(j, value) receives
// lds[ri(j)] (j = 0..NR-1); i and j are compile-time.  The caller has
// synchronised earlier readers of the slice; src is fully read before dst is
// called.
template <int W, int NW, int NR, class Src, class Wi, class Ri, class Dst>
__device__ __forceinline__ void lds_exchange(LdsT<W> (*lds)[W + 1], Src&& src, Wi&& wi, Ri&& ri, Dst&& dst)
{
    if constexpr (!split_lds<W>()) {
        sfor<NW>([&](auto I) {
            const int2 a = wi(I);
            lds[a.x][a.y] = src(I);
        });
        wave_sync();
        sfor<NR>([&](auto J) {
            const int2 a = ri(J);
            dst(J, lds[a.x][a.y]);
        });
    } else {
        float re[NR];
        sfor<NW>([&](auto I) {
            const int2 a = wi(I);
            lds[a.x][a.y] = src(I).x;
        });
        wave_sync();
        sfor<NR>([&](auto J) {
            const int2 a = ri(J);
            re[decltype(J)::value] = lds[a.x][a.y];
        });
        wave_sync();
        sfor<NW>([&](auto I) {
            const int2 a = wi(I);
            lds[a.x][a.y] = src(I).y;
        });
        wave_sync();
        sfor<NR>([&](auto J) {
            const int2 a = ri(J);
            dst(J, make_float2(re[decltype(J)::value], lds[a.x][a.y]));
        });
    }
}

// Transform the window's column t (z, from load_column) in place: lane t
// then holds spectral row ky = t of Z = FFT(a + i b), Z(t, kx) at
// z[brev_n<W>(kx)].
template <int W>
__device__ __forceinline__ void forward_spectrum(float2 (&z)[W], int t, LdsT<W> (*lds)[W + 1])
{
    // ---- FFT along y (column t), transpose through LDS ----
    fftreg::fft_dit<W, -1>(z);
    wave_sync();                                  // earlier readers of this slice are done
    lds_exchange<W, W, W>(
        lds, [&](auto K) { return z[brev_n<W>(decltype(K)::value)]; },
        [&](auto K) { return make_int2(decltype(K)::value, t); },
        [&](auto X) { return make_int2(t, decltype(X)::value); },
        [&](auto X, float2 v) { z[decltype(X)::value] = v; });
    // ---- FFT along x: lane t holds spectral row ky = t ----
    fftreg::fft_dit<W, -1>(z);
}

// Half of this lane's row of 4 P, P = B conj(A) the cross-power spectrum:
// h[kx] = 4 P(t, kx) for kx = 0..W/2.  With s = Z + conj Z(-k) = 2A and
// d = Z - conj Z(-k) = 2iB, P = -i d conj(s) / 4; the mirror row's Z(-t, -kx)
// comes through LDS, and only its columns 0 and W/2..W-1 are needed.  A and
// B are spectra of real windows, so P(-k) = conj P(k): the row's other half
// is the conjugate of the mirror lane's first half.  The factor 4 is a
// power of two and stays in the planes.
template <int W>
__device__ __forceinline__ void half_product(const float2 (&z)[W], int t, LdsT<W> (*lds)[W + 1],
                                             float2 (&h)[W / 2 + 1])
{
    constexpr int H2 = W / 2;
    const int m = (W - t) & (W - 1);              // row of -ky
    wave_sync();                                  // earlier readers of this slice are done
    lds_exchange<W, H2 + 1, H2 + 1>(
        lds,
        [&](auto I) {
            constexpr int i = decltype(I)::value;
            return z[brev_n<W>(i == 0 ? 0 : H2 - 1 + i)];
        },
        [&](auto I) {
            constexpr int i = decltype(I)::value;
            return make_int2(t, i == 0 ? 0 : H2 - 1 + i);
        },
        [&](auto K) { return make_int2(m, (W - decltype(K)::value) & (W - 1)); },
        [&](auto K, float2 zr) {                  // zr = Z(-k)
            constexpr int kx = decltype(K)::value;
            const float2 zk = z[brev_n<W>(kx)];
            const float sr = zk.x + zr.x, si = zk.y - zr.y;
            const float dr = zk.x - zr.x, di = zk.y + zr.y;
            h[kx] = make_float2(__builtin_fmaf(di, sr, -(dr * si)), -__builtin_fmaf(dr, sr, di * si));
        });
}

// The window's largest |component| of 4 P: the lane's half row covers its
// row's other half through the mirror lane, so the group covers the plane.
template <int W>
__device__ __forceinline__ float spectrum_max(const float2 (&h)[W / 2 + 1])
{
    float m = 0.f;
    sfor<W / 2 + 1>([&](auto K) {
        constexpr int k = decltype(K)::value;
        m = fmaxf(m, fmaxf(fabsf(h[k].x), fabsf(h[k].y)));
    });
    return group_max<W>(m);
}

// The power of two 2^k that equalises the largest components of two packed
// cross-power spectra, so that each plane's FP32 rounding stays relative to
// its own magnitude; multiplying by it is exact.
__device__ __forceinline__ float balance_scale(float pmax0, float pmax1)
{
    int e0, e1;
    frexpf(pmax0, &e0);
    frexpf(pmax1, &e1);
    const int k = (pmax0 == 0.f || pmax1 == 0.f) ? 0 : max(-60, min(60, e0 - e1));
    return ldexpf(1.f, k);
}

// One wave task: ROUNDS x (64 / W) windows, windows
// [(task * ROUNDS + r) * 64/W + slot] for round r and lane group slot.
template <int W, int DT, bool PLANES, int ROUNDS, bool BATCH>
__device__ __forceinline__ void corr_task(const CorrParams& p, int task, LdsT<W> (*lds)[W + 1], int lane)
{
    constexpr int WPW = 64 / W;                 // lane groups per wave
    constexpr int H2 = W / 2;
    const int slot = lane / W;
    const int t = lane % W;

    // round r of this lane group handles window (task * ROUNDS + r) * WPW + slot
    bool active[ROUNDS];
    WinGeom geo[ROUNDS];
    sfor<ROUNDS>([&](auto R) {
        constexpr int r = decltype(R)::value;
        geo[r] = window_geom<BATCH>(p, (task * ROUNDS + r) * WPW + slot);
        active[r] = geo[r].valid;
    });

    // ---- forward transforms; with two rounds the spectra pack as
    // Q = P0 + i 2^k P1, whose inverse is c0 + i 2^k c1 (both planes are
    // real).  The second window's columns are loaded after the first
    // window's half-row product: loading them earlier would cost the
    // registers of the fourth wave per SIMD at 32 x 32. ----
    float2 q[W];
    float inv_scale = 1.f;
    const bool compat = p.real_compat != 0;
    const int m = (W - t) & (W - 1);             // row of -ky
    float2 z0[W], z1[W];
    const TaskFrames frames = task_frames<DT>(geo[0]);
    load_column<W, DT>(frames, geo[0], t, z0);
    forward_spectrum<W>(z0, t, lds);
    float2 h0[H2 + 1];
    half_product<W>(z0, t, lds, h0);
    // q[kx] for kx = 0..W/2, and r[kx] = the mirror lane's q[W - kx] for
    // kx = 1..W/2-1: Q(-t, -kx) = conj P0(t, kx) + i 2^k conj P1(t, kx)
    float2 r[H2 + 1];
    if constexpr (ROUNDS == 2) {
        load_column<W, DT>(frames, geo[ROUNDS - 1], t, z1);
        forward_spectrum<W>(z1, t, lds);
        float2 h1[H2 + 1];
        half_product<W>(z1, t, lds, h1);
        const float scale = balance_scale(spectrum_max<W>(h0), spectrum_max<W>(h1));
        inv_scale = 1.f / scale;
        sfor<H2 + 1>([&](auto K) {
            constexpr int k = decltype(K)::value;
            const float2 a = h0[k], c = h1[k];
            q[k] = make_float2(__builtin_fmaf(-scale, c.y, a.x), __builtin_fmaf(scale, c.x, a.y));
            if constexpr (k >= 1 && k < H2)
                r[k] = make_float2(__builtin_fmaf(scale, c.y, a.x), __builtin_fmaf(scale, c.x, -a.y));
        });
    } else {
        sfor<H2 + 1>([&](auto K) {
            constexpr int k = decltype(K)::value;
            q[k] = h0[k];
            if constexpr (k >= 1 && k < H2) r[k] = make_float2(h0[k].x, -h0[k].y);
        });
    }
    wave_sync();                                 // half_product's readers of the slice are done
    lds_exchange<W, H2 - 1, H2 - 1>(
        lds, [&](auto I) { return r[decltype(I)::value + 1]; },
        [&](auto I) { return make_int2(t, decltype(I)::value + 1); },
        [&](auto J) { return make_int2(m, decltype(J)::value + 1); },
        [&](auto J, float2 v) { q[W - 1 - decltype(J)::value] = v; });
    if (compat) {
        // fft.h:166-183 leaves ky in {0, H/2} and kx = 0 empty (in both
        // packed spectra, so in their sum)
        const bool empty_row = t == 0 || t == H2;
        sfor<W>([&](auto K) {
            constexpr int kx = decltype(K)::value;
            if (empty_row || kx == 0) q[kx] = make_float2(0.f, 0.f);
        });
    }

    // ---- inverse along x (row ky = t), transpose, inverse along y ----
    fftreg::fft_dit<W, +1>(q);
    wave_sync();
    const int xs = (t + H2) & (W - 1);           // plane column that swaps into column t
    lds_exchange<W, W, W>(
        lds, [&](auto X) { return q[brev_n<W>(decltype(X)::value)]; },
        [&](auto X) { return make_int2(t, decltype(X)::value); },
        [&](auto K) { return make_int2(decltype(K)::value, xs); },
        [&](auto K, float2 v) { q[decltype(K)::value] = v; });
    fftreg::fft_dit<W, +1>(q);

    // quadrant-swapped planes, column t: S[h] = c[(h + W/2) mod W]
    float s[ROUNDS][W];
    sfor<W>([&](auto Hh) {
        constexpr int h = decltype(Hh)::value;
        const float2 v = q[brev_n<W>((h + H2) & (W - 1))];
        s[0][h] = v.x;
        if constexpr (ROUNDS == 2) s[ROUNDS - 1][h] = v.y;
    });
    // exact factors from these planes back to the correlation's
    float unscale[ROUNDS];
    unscale[0] = 0.25f;
    if constexpr (ROUNDS == 2) unscale[ROUNDS - 1] = 0.25f * inv_scale;

    // the planes to LDS (round r, column t, row h; the wave's transposes are
    // done) so that the stencil around each peak is read by address.  The
    // split slice holds one plane: then the tail stages them one at a time.
    constexpr bool SEQ = split_lds<W>() && ROUNDS > 1;
    float* pl = reinterpret_cast<float*>(&lds[0][0]);
    auto stage = [&](auto R) {
        constexpr int r = decltype(R)::value;
        wave_sync();                             // earlier readers of the slice are done
        sfor<W>([&](auto Hh) {
            constexpr int h = decltype(Hh)::value;
            pl[((SEQ ? 0 : r) * W + t) * (W + 1) + h] = s[r][h];
        });
        wave_sync();
    };
    if constexpr (!PLANES && !SEQ) {
        wave_sync();
        sfor<ROUNDS>([&](auto R) {
            constexpr int r = decltype(R)::value;
            sfor<W>([&](auto Hh) {
                constexpr int h = decltype(Hh)::value;
                pl[(r * W + t) * (W + 1) + h] = s[r][h];
            });
        });
        wave_sync();
    }
    finish_windows<W, ROUNDS, PLANES, SEQ>(p, s, unscale, geo, active, lane, t, stage, [&](int r, int row, int col) {
        return pl[((SEQ ? 0 : r) * W + col) * (W + 1) + row];
    });
}

// occupancy floor: the most waves whose VGPR budget keeps the
// register-resident lines of the instance without spilling.  The hot
// instances (result records, frames of up to 32 bits, one task per wave,
// 256-thread blocks) run 6 waves per SIMD at 16 x 16 (80 VGPRs) and 4 at
// 32 x 32 (128); plane output, f64 frames, one-wave batch blocks and the
// task loop need more registers and run fewer.
template <int W, bool LOOP, bool HOT>
constexpr int waves_per_simd()
{
    return W <= 8 ? 4 : W == 16 ? (HOT ? 6 : LOOP ? 4 : 5) : W == 32 ? (HOT ? 4 : 3) : 2;
}

// Launch modes: one pair, one task per wave; a batch (each window's pair
// from the device table) with one task per wave; a batch whose waves loop
// over tasks, so that the launch may be persistent.  The loop keeps its
// invariants live across the whole task and costs registers: 16 x 16 and
// 32 x 32 run one wave fewer per SIMD with it.
enum LaunchMode { kOnePair = 0, kBatch = 1, kBatchLoop = 2 };

template <int W, int DT, bool PLANES, int ROUNDS, int NT, int MODE>
__global__ __launch_bounds__(NT) __attribute__((
    amdgpu_waves_per_eu(waves_per_simd<W, MODE == kBatchLoop,
                                       !PLANES && DT != 3 && (MODE == kOnePair || (MODE == kBatch && NT == 256))>())))
void corr_fft_kernel(CorrParams p)
{
    constexpr int WPW = 64 / W;
    constexpr int WAVES = NT / 64;
    constexpr bool BATCH = MODE != kOnePair;
    __shared__ LdsT<W> lds_all[WAVES][WPW][W][W + 1];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int ntasks = (p.n_windows + ROUNDS * WPW - 1) / (ROUNDS * WPW);
    if constexpr (MODE == kBatchLoop) {
        // each XCD strides through its own contiguous share of the tasks
        const XcdShare sh = xcd_share((int)blockIdx.x, (int)gridDim.x, ntasks);
        for (int task = sh.lo + sh.rank * WAVES + wave; task < sh.hi; task += sh.blocks * WAVES)
            corr_task<W, DT, PLANES, ROUNDS, BATCH>(p, task, lds_all[wave][lane / W], lane);
    } else {
        const int task = xcd_block((int)blockIdx.x, (int)gridDim.x) * WAVES + wave;
        if (task < ntasks) corr_task<W, DT, PLANES, ROUNDS, BATCH>(p, task, lds_all[wave][lane / W], lane);
    }
}

// Two windows per lane group share one packed inverse transform: the second
// spectrum fits in VGPRs (at 64 x 64 the kernel is LDS-bound at two waves
// per SIMD, which leaves up to 256 VGPRs per wave), except for f64 frames at
// 64 x 64, whose wider loads would spill.  batch_align() mirrors this.
template <int W, int DT>
constexpr int rounds_for() { return W == 64 && DT == 3 ? 1 : 2; }

template <int W, int DT, int NT, bool BATCH>
static hipError_t launch_wn(const CorrParams& p, bool planes, int blocks, hipStream_t stream)
{
    constexpr int ROUNDS = rounds_for<W, DT>();
    static_assert(batch_align(W, DT) == ROUNDS * (64 / W), "batch pair ranges align to the wave task");
    constexpr int tasks_per_block = NT / 64;
    const int ntasks = (p.n_windows + ROUNDS * (64 / W) - 1) / (ROUNDS * (64 / W));
    const int full = (ntasks + tasks_per_block - 1) / tasks_per_block;
    const int grid = blocks > 0 && blocks < full ? blocks : full;
    if constexpr (BATCH) {
        if (grid == full)    // one task per wave: no loop
            hipLaunchKernelGGL((corr_fft_kernel<W, DT, false, ROUNDS, NT, kBatch>), dim3(grid), dim3(NT), 0, stream, p);
        else
            hipLaunchKernelGGL((corr_fft_kernel<W, DT, false, ROUNDS, NT, kBatchLoop>), dim3(grid), dim3(NT), 0,
                               stream, p);
    } else {
        if (planes)
            hipLaunchKernelGGL((corr_fft_kernel<W, DT, true, ROUNDS, NT, kOnePair>), dim3(grid), dim3(NT), 0, stream, p);
        else
            hipLaunchKernelGGL((corr_fft_kernel<W, DT, false, ROUNDS, NT, kOnePair>), dim3(grid), dim3(NT), 0, stream, p);
    }
    return hipGetLastError();
}

template <int W, int DT>
static hipError_t launch_w(const CorrParams& p, bool planes, int nt, int blocks, hipStream_t stream)
{
    if (p.pairs == nullptr) {
        if (nt != 256) return hipErrorInvalidValue;
        // one task per wave; at 64 x 64 (LDS-bound at two waves per SIMD)
        // one wave per block, so that a finished wave frees its LDS without
        // waiting for the block's others (6.5% faster than 256 threads for
        // u16 frames, as fast for f32, faster in batches).
        if constexpr (W == 64) return launch_wn<W, DT, 64, false>(p, planes, 0, stream);
        else return launch_wn<W, DT, 256, false>(p, planes, 0, stream);
    }
    if (planes) return hipErrorInvalidValue;
    if (nt == 256) return launch_wn<W, DT, 256, true>(p, false, blocks, stream);
    if (nt == 64) return launch_wn<W, DT, 64, true>(p, false, blocks, stream);
    return hipErrorInvalidValue;
}

#ifndef OPIV_CORR_FFT_W64_ONLY
// the 64 x 64 launches for f32 frames, compiled in corr_fft64.hip
template <int DT>
hipError_t launch_w64(const CorrParams& p, bool planes, int nt, int blocks, hipStream_t stream);

template <int DT>
static hipError_t launch_dt(const CorrParams& p, int win, bool planes, int nt, int blocks, hipStream_t stream)
{
    switch (win) {
    case 8: return launch_w<8, DT>(p, planes, nt, blocks, stream);
    case 16: return launch_w<16, DT>(p, planes, nt, blocks, stream);
    case 32: return launch_w<32, DT>(p, planes, nt, blocks, stream);
    case 64:
        if constexpr (DT == 2) return launch_w64<DT>(p, planes, nt, blocks, stream);
        else return launch_w<64, DT>(p, planes, nt, blocks, stream);
    case 128:   // one workgroup of 256 threads per window, never persistent
        return nt == 256 ? launch_corr_fft128(p, DT, planes, stream) : hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_corr_fft_shape(const CorrParams& p, int win, int dtype, bool planes, int nt, int blocks,
                                 hipStream_t stream)
{
    if (p.n_windows <= 0) return hipSuccess;
    switch (dtype) {
    case 0: return launch_dt<0>(p, win, planes, nt, blocks, stream);
    case 1: return launch_dt<1>(p, win, planes, nt, blocks, stream);
    case 2: return launch_dt<2>(p, win, planes, nt, blocks, stream);
    case 3: return launch_dt<3>(p, win, planes, nt, blocks, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_corr_fft(const CorrParams& p, int win, int dtype, bool planes, hipStream_t stream)
{
    return launch_corr_fft_shape(p, win, dtype, planes, 256, 0, stream);
}
#endif // OPIV_CORR_FFT_W64_ONLY

} // namespace opiv
