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
// one wavefront (64 / W windows per wave).  Lane t loads pixel column t of
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
// LDS traffic stays inside a wave, so only wave-level syncs are needed.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <limits>

#include "fft_reg.hpp"
#include "corr_fft.hpp"

namespace opiv {

using fftreg::brev_n;
using fftreg::sfor;

namespace {

// LDS writes of this wave become visible to its later LDS reads; no
// instruction is moved across.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// value of lane+1 / lane-1 (whole-wave DPP shifts; the group-edge lanes
// receive a neighbour group's value, which the scan never uses)
__device__ __forceinline__ float from_right(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ float from_left(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, false));
}

template <int DT>
__device__ __forceinline__ float load_px(const void* base, size_t idx)
{
    if constexpr (DT == 0) return (float)((const uint8_t*)base)[idx];
    else if constexpr (DT == 1) return (float)((const uint16_t*)base)[idx];
    else if constexpr (DT == 2) return ((const float*)base)[idx];
    else return (float)((const double*)base)[idx];
}

struct Cand {
    float v;
    int i;
};

// ordering of find_peaks' sort (descending value); equal values keep scan
// order, i.e. the smaller row-major index comes first
__device__ __forceinline__ bool better(float va, int ia, float vb, int ib)
{
    return va > vb || (va == vb && ia < ib);
}

__device__ __forceinline__ double gauss3(double l, double c, double r)
{
    // image_utils.impl.h:86-94, same operation order
    const double num = log(l) - log(r);
    const double den = 2.0 * (log(l) + log(r) - 2.0 * log(c));
    if (den == 0.0) return 0.0;
    return num / den;
}

} // namespace

template <int W, int DT, bool PLANES>
__global__ __launch_bounds__(kCorrThreads) void corr_fft_kernel(CorrParams p)
{
    constexpr int WPW = 64 / W;                 // windows per wave
    constexpr int WAVES = kCorrThreads / 64;
    constexpr int LS = W + 1;                   // padded LDS row (float2)
    constexpr int H2 = W / 2;
    __shared__ float2 lds_all[WAVES][WPW][W][LS];

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int slot = lane / W;
    const int t = lane % W;
    float2 (*lds)[LS] = lds_all[wave][slot];

    const int local = (blockIdx.x * WAVES + wave) * WPW + slot;
    const bool active = local < p.n_windows;
    const int gid = p.first + (active ? local : p.n_windows - 1);
    const int ix = gid % p.nx, iy = gid / p.nx;
    const int x0 = p.x_start + ix * p.step_x;
    const int y0 = p.y_start + iy * p.step_y;

    // ---- load column t of both windows: z[y] = a + i b ----
    float2 z[W];
    {
        size_t idx = (size_t)y0 * p.pitch + (size_t)(x0 + t);
        sfor<W>([&](auto Y) {
            constexpr int y = decltype(Y)::value;
            z[y] = make_float2(load_px<DT>(p.a, idx + (size_t)y * p.pitch), load_px<DT>(p.b, idx + (size_t)y * p.pitch));
        });
    }

    // ---- forward FFT along y (column t), then transpose ----
    fftreg::fft_dif<W, -1>(z);
    sfor<W>([&](auto K) {
        constexpr int ky = decltype(K)::value;
        lds[ky][t] = z[brev_n<W>(ky)];
    });
    wave_sync();
    sfor<W>([&](auto X) {
        constexpr int x = decltype(X)::value;
        z[x] = lds[t][x];
    });

    // ---- forward FFT along x: lane t holds spectral row ky = t ----
    fftreg::fft_dif<W, -1>(z);
    wave_sync();
    sfor<W>([&](auto K) {
        constexpr int kx = decltype(K)::value;
        lds[t][kx] = z[brev_n<W>(kx)];
    });
    wave_sync();

    // ---- separate the two real spectra and form P = B conj(A) ----
    const int m = (W - t) & (W - 1);            // row of -ky
    const bool compat_row = p.real_compat && (t == 0 || t == H2);
    float2 q[W];
    sfor<W>([&](auto K) {
        constexpr int kx = decltype(K)::value;
        constexpr int kxm = (W - kx) & (W - 1);
        const float2 zk = z[brev_n<W>(kx)];
        const float2 zr = lds[m][kxm];           // Z(-k)
        // s = Z + conj Z(-k) = 2A ; d = Z - conj Z(-k) = 2iB
        const float sr = zk.x + zr.x, si = zk.y - zr.y;
        const float dr = zk.x - zr.x, di = zk.y + zr.y;
        // P = -i d conj(s) / 4
        float pr = 0.25f * (di * sr - dr * si);
        float pi = -0.25f * (dr * sr + di * si);
        // fft.h:166-183 leaves ky in {0, H/2} and kx = 0 empty
        if (compat_row || (kx == 0 && p.real_compat)) { pr = 0.f; pi = 0.f; }
        q[kx] = make_float2(pr, pi);
    });

    // ---- inverse along x (row ky = t), transpose, inverse along y ----
    fftreg::fft_dif<W, +1>(q);
    wave_sync();
    sfor<W>([&](auto X) {
        constexpr int x = decltype(X)::value;
        lds[t][x] = q[brev_n<W>(x)];
    });
    wave_sync();
    const int xs = (t + H2) & (W - 1);           // plane column that swaps into column t
    sfor<W>([&](auto K) {
        constexpr int ky = decltype(K)::value;
        q[ky] = lds[ky][xs];
    });
    fftreg::fft_dif<W, +1>(q);

    // swapped plane, column t: S[h] = c[(h + W/2) mod W]
    float s[W];
    sfor<W>([&](auto Hh) {
        constexpr int h = decltype(Hh)::value;
        s[h] = q[brev_n<W>((h + H2) & (W - 1))].x;
    });

    if constexpr (PLANES) {
        if (active) {
            double* dst = p.planes + (size_t)local * W * W + t;
            sfor<W>([&](auto Hh) {
                constexpr int h = decltype(Hh)::value;
                dst[h * W] = (double)s[h];
            });
        }
        return;
    } else {
        // ---- find_peaks(.., 2, 1): strict 4-neighbour maxima over rows
        // [ry+1, ry+rh-2) and columns [rx+1, rx+rw-1) of the swapped plane
        // (image_utils.impl.h:48-64; the limit-search view is rect::dilate(0.5)) ----
        const int r0 = p.limit_search ? W / 4 : 0;
        const int rlen = p.limit_search ? W / 2 : W;
        const int hlo = r0 + 1, hhi = r0 + rlen - 3;   // inclusive
        const int wlo = r0 + 1, whi = r0 + rlen - 2;
        const bool col_ok = t >= wlo && t <= whi;
        float v1 = -INFINITY, v2 = -INFINITY;
        int i1 = 0x7fffffff, i2 = 0x7fffffff;
        sfor<W - 3>([&](auto Hm) {
            constexpr int h = decltype(Hm)::value + 1;
            const float c = s[h];
            const float l = from_left(c);
            const float r = from_right(c);
            const bool ismax = col_ok && h >= hlo && h <= hhi && l < c && r < c && s[h - 1] < c && s[h + 1] < c;
            // branchless top-2 update; rows arrive in scan order, so a later
            // equal value never displaces an earlier one
            const int idx = h * W + t;
            const bool gt1 = ismax && c > v1;
            const bool gt2 = ismax && !gt1 && c > v2;
            v2 = gt1 ? v1 : (gt2 ? c : v2);
            i2 = gt1 ? i1 : (gt2 ? idx : i2);
            v1 = gt1 ? c : v1;
            i1 = gt1 ? idx : i1;
        });
        // top-2 over the W lanes of the window
#pragma unroll
        for (int off = W / 2; off >= 1; off >>= 1) {
            const float ov1 = __shfl_xor(v1, off, 64), ov2 = __shfl_xor(v2, off, 64);
            const int oi1 = __shfl_xor(i1, off, 64), oi2 = __shfl_xor(i2, off, 64);
            if (better(v1, i1, ov1, oi1)) {
                if (!better(v2, i2, ov1, oi1)) { v2 = ov1; i2 = oi1; }
            } else {
                const bool keep_mine = better(v1, i1, ov2, oi2);
                v2 = keep_mine ? v1 : ov2;
                i2 = keep_mine ? i1 : oi2;
                v1 = ov1;
                i1 = oi1;
            }
        }
        const bool two = i2 != 0x7fffffff;
        const int px = two ? (i1 & (W - 1)) : 1;
        const int py = two ? (i1 / W) : 1;
        // stencil values around the peak: centre column from lane px,
        // horizontal neighbours from lanes px -+ 1
        float mid = 0.f, up = 0.f, dn = 0.f;
        sfor<W - 2>([&](auto Hm) {
            constexpr int h = decltype(Hm)::value + 1;
            if (h == py) { mid = s[h]; up = s[h - 1]; dn = s[h + 1]; }
        });
        const int gbase = lane - t;
        const float fl = __shfl(mid, gbase + px - 1, 64);
        const float fc = __shfl(mid, gbase + px, 64);
        const float fr = __shfl(mid, gbase + px + 1, 64);
        const float fu = __shfl(up, gbase + px, 64);
        const float fd = __shfl(dn, gbase + px, 64);

        if (active && t == 0) {
            opiv_window_result res;
            res.reserved = 0;
            if (!two) {
                // main.cpp:214-218: the default (all-zero) record stays
                res.x = res.y = res.u = res.v = res.sn = 0.0;
                res.peak0 = res.peak1 = 0.0;
                res.px = res.py = 0;
                res.flags = OPIV_FLAG_FEWER_THAN_2_PEAKS;
            } else {
                uint32_t flags = 0;
                // image_utils.impl.h:96-98: x from (l, c, r), y from rows py-1, py, py+1
                const double plx = (double)px + gauss3((double)fl, (double)fc, (double)fr);
                const double ply = (double)py + gauss3((double)fu, (double)fc, (double)fd);
                if (!isfinite(plx) || !isfinite(ply)) flags |= OPIV_FLAG_NONFINITE_FIT;
                const float nmax = fmaxf(fmaxf(fl, fr), fmaxf(fu, fd));
                const float tol = p.tie_tol * fabsf(v1);
                if (v1 - v2 <= tol || v1 - nmax <= tol) flags |= OPIV_FLAG_NEAR_TIE;
                if (v1 == v2) flags |= OPIV_FLAG_EXACT_TIE;
                // main.cpp:220-234
                const int midx = x0 + W / 2, midy = y0 + W / 2;
                res.x = (double)midx;
                res.y = (double)p.img_h - (double)midy;
                res.u = (double)midx - ((double)x0 + plx);
                res.v = (double)midy - ((double)y0 + ply);
                res.sn = (v2 > 0.f) ? (double)v1 / (double)v2 : 0.0;
                res.peak0 = (double)v1;
                res.peak1 = (double)v2;
                res.px = px;
                res.py = py;
                res.flags = flags;
            }
            p.out[local] = res;
        }
    }
}

template <int W, int DT>
static hipError_t launch_w(const CorrParams& p, bool planes, hipStream_t stream)
{
    constexpr int per_block = (kCorrThreads / 64) * (64 / W);
    const int blocks = (p.n_windows + per_block - 1) / per_block;
    if (planes)
        hipLaunchKernelGGL((corr_fft_kernel<W, DT, true>), dim3(blocks), dim3(kCorrThreads), 0, stream, p);
    else
        hipLaunchKernelGGL((corr_fft_kernel<W, DT, false>), dim3(blocks), dim3(kCorrThreads), 0, stream, p);
    return hipGetLastError();
}

template <int DT>
static hipError_t launch_dt(const CorrParams& p, int win, bool planes, hipStream_t stream)
{
    switch (win) {
    case 8: return launch_w<8, DT>(p, planes, stream);
    case 16: return launch_w<16, DT>(p, planes, stream);
    case 32: return launch_w<32, DT>(p, planes, stream);
    case 64: return launch_w<64, DT>(p, planes, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_corr_fft(const CorrParams& p, int win, int dtype, bool planes, hipStream_t stream)
{
    if (p.n_windows <= 0) return hipSuccess;
    switch (dtype) {
    case 0: return launch_dt<0>(p, win, planes, stream);
    case 1: return launch_dt<1>(p, win, planes, stream);
    case 2: return launch_dt<2>(p, win, planes, stream);
    case 3: return launch_dt<3>(p, win, planes, stream);
    default: return hipErrorInvalidValue;
    }
}

} // namespace opiv
