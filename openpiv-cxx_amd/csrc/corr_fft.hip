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

// frames are global memory: address-space-1 pointers give global (not flat)
// loads even when the pointer comes from the batch table
template <class T>
using gptr = const __attribute__((address_space(1))) T*;

template <int DT>
__device__ __forceinline__ float load_px(const void* base, size_t idx)
{
    if constexpr (DT == 0) return (float)((gptr<uint8_t>)base)[idx];
    else if constexpr (DT == 1) return (float)((gptr<uint16_t>)base)[idx];
    else if constexpr (DT == 2) return ((gptr<float>)base)[idx];
    else return (float)((gptr<double>)base)[idx];
}

// ordering of find_peaks' sort (descending value); equal values keep scan
// order, i.e. the smaller row-major index comes first
__device__ __forceinline__ bool better(float va, int ia, float vb, int ib)
{
    return va > vb || (va == vb && ia < ib);
}

} // namespace

// Where one window lives: frames, pitch, bottom-left corner, frame height,
// the index of its result and whether it is a real window of the launch
// (task slots past the end of a pair's range recompute that pair's last
// window and store nothing).
struct WinGeom {
    const void* a;
    const void* b;
    size_t pitch;
    int x0, y0;
    uint32_t img_h;
    int out;
    bool valid;
};

template <class T>
using gptr_t = const __attribute__((address_space(1))) T*;

// Geometry of window `local` of a launch: the single pair of the params, or
// (batch) the pair whose window range [first, first + nx * ny) holds it.
template <bool BATCH>
__device__ __forceinline__ WinGeom window_geom(const CorrParams& p, int local)
{
    WinGeom g;
    g.valid = local < p.n_windows;
    if constexpr (!BATCH) {
        const int l = g.valid ? local : p.n_windows - 1;
        const int gid = p.first + l;
        const int ix = gid % p.nx, iy = gid / p.nx;
        g.a = p.a;
        g.b = p.b;
        g.pitch = p.pitch;
        g.x0 = p.x_start + ix * p.step_x;
        g.y0 = p.y_start + iy * p.step_y;
        g.img_h = p.img_h;
        g.out = l;
    } else {
        const gptr_t<BatchPair> pairs = (gptr_t<BatchPair>)p.pairs;
        const int l = g.valid ? local : p.n_windows - 1;
        int lo = 0, hi = p.npairs;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (pairs[mid].first <= l) lo = mid;
            else hi = mid;
        }
        const auto& bp = pairs[lo];
        const int n = bp.nx * bp.ny;
        int gid = l - bp.first;
        if (gid >= n) {
            g.valid = false;
            gid = n - 1;
        }
        const int ix = gid % bp.nx, iy = gid / bp.nx;
        g.a = bp.a;
        g.b = bp.b;
        g.pitch = bp.pitch;
        g.x0 = bp.x_start + ix * bp.step_x;
        g.y0 = bp.y_start + iy * bp.step_y;
        g.img_h = bp.img_h;
        g.out = bp.out_first + gid;
    }
    return g;
}

// Load window (x0, y0) column t of both frames, transform, and leave in q
// (natural order, lane t = spectral row ky) the cross-power spectrum
// P = B conj(A) of this lane's row -- or, with ACC, add i 2^k P to q.
// Returns the window's max |Z| component (Z = FFT(a + i b)), which bounds
// |P| <= max|Z|^2.  The power of two 2^k equalises that bound between the
// two packed spectra (zmax0 is the first window's), so each plane's FP32
// rounding stays relative to its own magnitude; scaling by 2^k is exact and
// the caller divides it out of the second plane.
template <int W, int DT, bool ACC>
__device__ __forceinline__ float forward_product(const WinGeom& g, bool real_compat, int t,
                                                 float2 (*lds)[W + 1], float2 (&q)[W],
                                                 float zmax0, float& scale)
{
    constexpr int H2 = W / 2;
    // ---- load column t: z[y] = a + i b ----
    float2 z[W];
    {
        const size_t idx = (size_t)g.y0 * g.pitch + (size_t)(g.x0 + t);
        sfor<W>([&](auto Y) {
            constexpr int y = decltype(Y)::value;
            z[y] = make_float2(load_px<DT>(g.a, idx + (size_t)y * g.pitch),
                               load_px<DT>(g.b, idx + (size_t)y * g.pitch));
        });
    }
    // ---- FFT along y (column t), transpose through LDS ----
    fftreg::fft_dif<W, -1>(z);
    wave_sync();                                  // earlier readers of this buffer are done
    sfor<W>([&](auto K) {
        constexpr int ky = decltype(K)::value;
        lds[ky][t] = z[brev_n<W>(ky)];
    });
    wave_sync();
    sfor<W>([&](auto X) {
        constexpr int x = decltype(X)::value;
        z[x] = lds[t][x];
    });
    // ---- FFT along x: lane t holds spectral row ky = t; publish it ----
    fftreg::fft_dif<W, -1>(z);
    wave_sync();
    sfor<W>([&](auto K) {
        constexpr int kx = decltype(K)::value;
        lds[t][kx] = z[brev_n<W>(kx)];
    });
    wave_sync();
    // max |Z| over the window (lanes of the group)
    float zmax = 0.f;
    sfor<W>([&](auto K) {
        constexpr int k = decltype(K)::value;
        zmax = fmaxf(zmax, fmaxf(fabsf(z[k].x), fabsf(z[k].y)));
    });
#pragma unroll
    for (int off = W / 2; off >= 1; off >>= 1) zmax = fmaxf(zmax, __shfl_xor(zmax, off, 64));
    scale = 1.f;
    if constexpr (ACC) {
        int e0, e1;
        frexpf(zmax0, &e0);
        frexpf(zmax, &e1);
        const int k = (zmax0 == 0.f || zmax == 0.f) ? 0 : max(-60, min(60, 2 * (e0 - e1)));
        scale = ldexpf(1.f, k);
    }
    const float f = 0.25f * scale;
    // ---- separate the two real spectra and form P = B conj(A) ----
    const int m = (W - t) & (W - 1);              // row of -ky
    const bool compat_row = real_compat && (t == 0 || t == H2);
    sfor<W>([&](auto K) {
        constexpr int kx = decltype(K)::value;
        constexpr int kxm = (W - kx) & (W - 1);
        const float2 zk = z[brev_n<W>(kx)];
        const float2 zr = lds[m][kxm];            // Z(-k)
        // s = Z + conj Z(-k) = 2A ; d = Z - conj Z(-k) = 2iB ; P = -i d conj(s) / 4
        const float sr = zk.x + zr.x, si = zk.y - zr.y;
        const float dr = zk.x - zr.x, di = zk.y + zr.y;
        float pr = f * (di * sr - dr * si);
        float pi = -f * (dr * sr + di * si);
        // fft.h:166-183 leaves ky in {0, H/2} and kx = 0 empty
        if (compat_row || (kx == 0 && real_compat)) { pr = 0.f; pi = 0.f; }
        if constexpr (ACC) q[kx] = make_float2(q[kx].x - pi, q[kx].y + pr);
        else q[kx] = make_float2(pr, pi);
    });
    return zmax;
}

// find_peaks(.., 2, 1) over one quadrant-swapped plane column held by lane
// t: strict 4-neighbour maxima on rows [hlo, hhi] and columns [wlo, whi]
// (image_utils.impl.h:48-64), reduced to the largest value (and its scan
// index h*W + t; equal values keep scan order) and the second largest value
// over the W lanes of the window.
template <int W>
__device__ __forceinline__ void scan_top2(const float (&s)[W], int t, int hlo, int hhi, bool col_ok,
                                          float& v1, int& i1, float& v2)
{
    v1 = -INFINITY;
    v2 = -INFINITY;
    int h1 = 0x7fff;
    sfor<W - 3>([&](auto Hm) {
        constexpr int h = decltype(Hm)::value + 1;
        const float c = s[h];
        const float nb = fmaxf(fmaxf(from_left(c), from_right(c)), fmaxf(s[h - 1], s[h + 1]));
        const bool ismax = col_ok && h >= hlo && h <= hhi && c > nb;
        const float mc = ismax ? c : -INFINITY;
        const bool gt = mc > v1;                  // rows arrive in scan order: ties keep the first
        v2 = fmaxf(v2, fminf(mc, v1));
        h1 = gt ? h : h1;
        v1 = gt ? mc : v1;
    });
    i1 = h1 * W + t;
#pragma unroll
    for (int off = W / 2; off >= 1; off >>= 1) {
        const float ov1 = __shfl_xor(v1, off, 64);
        const float ov2 = __shfl_xor(v2, off, 64);
        const int oi1 = __shfl_xor(i1, off, 64);
        v2 = fmaxf(fmaxf(v2, ov2), fminf(v1, ov1));
        const bool mine = better(v1, i1, ov1, oi1);
        v1 = mine ? v1 : ov1;
        i1 = mine ? i1 : oi1;
    }
}

// One wave task: ROUNDS x (64 / W) windows, windows
// [(task * ROUNDS + r) * 64/W + slot] for round r and lane group slot.
template <int W, int DT, bool PLANES, int ROUNDS, bool BATCH>
__device__ __forceinline__ void corr_task(const CorrParams& p, int task, float2 (*lds)[W + 1], int lane)
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
    // Q = P0 + i P1, whose inverse is c0 + i c1 (both planes are real) ----
    const int gbase = lane - t;
    float2 q[W];
    float scale = 1.f, inv_scale = 1.f;
    const bool compat = p.real_compat != 0;
    const float zmax0 = forward_product<W, DT, false>(geo[0], compat, t, lds, q, 0.f, scale);
    if constexpr (ROUNDS == 2) {
        forward_product<W, DT, true>(geo[ROUNDS - 1], compat, t, lds, q, zmax0, scale);
        inv_scale = 1.f / scale;
    }

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

    // quadrant-swapped planes, column t: S[h] = c[(h + W/2) mod W]
    float s[ROUNDS][W];
    sfor<W>([&](auto Hh) {
        constexpr int h = decltype(Hh)::value;
        const float2 v = q[brev_n<W>((h + H2) & (W - 1))];
        s[0][h] = v.x;
        if constexpr (ROUNDS == 2) s[ROUNDS - 1][h] = v.y * inv_scale;
    });

    if constexpr (PLANES) {
        sfor<ROUNDS>([&](auto R) {
            constexpr int r = decltype(R)::value;
            if (active[r]) {
                double* dst = p.planes + (size_t)geo[r].out * W * W + t;
                sfor<W>([&](auto Hh) {
                    constexpr int h = decltype(Hh)::value;
                    dst[h * W] = (double)s[r][h];
                });
            }
        });
        return;
    } else {
        // search region: whole plane, or the limit-search view rect::dilate(0.5)
        const int r0 = p.limit_search ? W / 4 : 0;
        const int rlen = p.limit_search ? W / 2 : W;
        const int hlo = r0 + 1, hhi = r0 + rlen - 3;
        const bool col_ok = t >= r0 + 1 && t <= r0 + rlen - 2;

        float v1[ROUNDS], v2[ROUNDS], st[ROUNDS][5];
        int px[ROUNDS], py[ROUNDS];
        bool two[ROUNDS];
        sfor<ROUNDS>([&](auto R) {
            constexpr int r = decltype(R)::value;
            int i1;
            scan_top2<W>(s[r], t, hlo, hhi, col_ok, v1[r], i1, v2[r]);
            two[r] = v2[r] > -INFINITY;
            px[r] = two[r] ? (i1 & (W - 1)) : 1;
            py[r] = two[r] ? (i1 / W) : 1;
            // stencil (l, c, r, up, down) around the peak: the centre column
            // comes from lane px, the horizontal neighbours from lanes px -+ 1
            float mid = 0.f, up = 0.f, dn = 0.f;
            sfor<W - 2>([&](auto Hm) {
                constexpr int h = decltype(Hm)::value + 1;
                const bool at = h == py[r];
                mid = at ? s[r][h] : mid;
                up = at ? s[r][h - 1] : up;
                dn = at ? s[r][h + 1] : dn;
            });
            st[r][0] = __shfl(mid, gbase + px[r] - 1, 64);
            st[r][1] = __shfl(mid, gbase + px[r], 64);
            st[r][2] = __shfl(mid, gbase + px[r] + 1, 64);
            st[r][3] = __shfl(up, gbase + px[r], 64);
            st[r][4] = __shfl(dn, gbase + px[r], 64);
        });

        // ---- FP64 logs of the 5 * ROUNDS stencil values, one per lane ----
        constexpr int NV = 5 * ROUNDS;
        constexpr int NP = (NV + W - 1) / W;
        double lg[NP];
        sfor<NP>([&](auto K) {
            constexpr int k = decltype(K)::value;
            float v = 1.f;
            sfor<NV>([&](auto J) {
                constexpr int j = decltype(J)::value;
                v = (t + k * W == j) ? st[j / 5][j % 5] : v;
            });
            lg[k] = log((double)v);
        });
        double ln[NV];
        sfor<NV>([&](auto J) {
            constexpr int j = decltype(J)::value;
            ln[j] = __shfl(lg[j / W], gbase + (j % W), 64);
        });

        sfor<ROUNDS>([&](auto R) {
            constexpr int r = decltype(R)::value;
            if (!(active[r] && t == r)) return;
            opiv_window_result res;
            res.reserved = 0;
            if (!two[r]) {
                // main.cpp:214-218: the default (all-zero) record stays
                res.x = res.y = res.u = res.v = res.sn = 0.0;
                res.peak0 = res.peak1 = 0.0;
                res.px = res.py = 0;
                res.flags = OPIV_FLAG_FEWER_THAN_2_PEAKS;
            } else {
                // image_utils.impl.h:86-98 on the logs of (l, c, r) and (up, c, down)
                const double lnl = ln[5 * r + 0], lnc = ln[5 * r + 1], lnr = ln[5 * r + 2];
                const double lnu = ln[5 * r + 3], lnd = ln[5 * r + 4];
                const double numx = lnl - lnr, denx = 2.0 * (lnl + lnr - 2.0 * lnc);
                const double numy = lnu - lnd, deny = 2.0 * (lnu + lnd - 2.0 * lnc);
                const double plx = (double)px[r] + (denx == 0.0 ? 0.0 : numx / denx);
                const double ply = (double)py[r] + (deny == 0.0 ? 0.0 : numy / deny);
                uint32_t flags = 0;
                if (!isfinite(plx) || !isfinite(ply)) flags |= OPIV_FLAG_NONFINITE_FIT;
                const float nmax = fmaxf(fmaxf(st[r][0], st[r][2]), fmaxf(st[r][3], st[r][4]));
                const float tol = p.tie_tol * fabsf(v1[r]);
                if (v1[r] - v2[r] <= tol || v1[r] - nmax <= tol) flags |= OPIV_FLAG_NEAR_TIE;
                if (v1[r] == v2[r]) flags |= OPIV_FLAG_EXACT_TIE;
                // main.cpp:220-234
                const int midx = geo[r].x0 + W / 2, midy = geo[r].y0 + W / 2;
                res.x = (double)midx;
                res.y = (double)geo[r].img_h - (double)midy;
                res.u = (double)midx - ((double)geo[r].x0 + plx);
                res.v = (double)midy - ((double)geo[r].y0 + ply);
                res.sn = (v2[r] > 0.f) ? (double)v1[r] / (double)v2[r] : 0.0;
                res.peak0 = (double)v1[r];
                res.peak1 = (double)v2[r];
                res.px = px[r];
                res.py = py[r];
                res.flags = flags;
            }
            p.out[geo[r].out] = res;
        });
    }
}

// occupancy target per window size: the VGPR budget that keeps the
// register-resident lines (nearly) without spilling; batch kernels carry
// per-lane frame pointers and get a little more room at W = 16
template <int W, bool BATCH>
constexpr int waves_per_simd() { return W <= 8 ? 4 : W == 16 ? (BATCH ? 3 : 4) : (W == 32 ? 2 : 1); }

// BATCH reads each window's pair from the device table (per-lane frame
// pointers) and loops over wave tasks, so the launch may be persistent; a
// single-pair launch has uniform geometry and one task per wave.
template <int W, int DT, bool PLANES, int ROUNDS, int NT, bool BATCH>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(waves_per_simd<W, BATCH>())))
void corr_fft_kernel(CorrParams p)
{
    constexpr int WPW = 64 / W;
    constexpr int WAVES = NT / 64;
    __shared__ float2 lds_all[WAVES][WPW][W][W + 1];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int ntasks = (p.n_windows + ROUNDS * WPW - 1) / (ROUNDS * WPW);
    if constexpr (BATCH) {
        for (int task = blockIdx.x * WAVES + wave; task < ntasks; task += gridDim.x * WAVES)
            corr_task<W, DT, PLANES, ROUNDS, BATCH>(p, task, lds_all[wave][lane / W], lane);
    } else {
        const int task = blockIdx.x * WAVES + wave;
        if (task < ntasks) corr_task<W, DT, PLANES, ROUNDS, BATCH>(p, task, lds_all[wave][lane / W], lane);
    }
}

template <int W>
constexpr int rounds_for() { return W <= 32 ? 2 : 1; }   // second spectrum fits in VGPRs up to 32

template <int W, int DT, int NT, bool BATCH>
static hipError_t launch_wn(const CorrParams& p, bool planes, int blocks, hipStream_t stream)
{
    constexpr int ROUNDS = rounds_for<W>();
    constexpr int tasks_per_block = NT / 64;
    const int ntasks = (p.n_windows + ROUNDS * (64 / W) - 1) / (ROUNDS * (64 / W));
    const int full = (ntasks + tasks_per_block - 1) / tasks_per_block;
    const int grid = blocks > 0 && blocks < full ? blocks : full;
    if constexpr (BATCH) {
        hipLaunchKernelGGL((corr_fft_kernel<W, DT, false, ROUNDS, NT, true>), dim3(grid), dim3(NT), 0, stream, p);
    } else {
        if (planes)
            hipLaunchKernelGGL((corr_fft_kernel<W, DT, true, ROUNDS, NT, false>), dim3(grid), dim3(NT), 0, stream, p);
        else
            hipLaunchKernelGGL((corr_fft_kernel<W, DT, false, ROUNDS, NT, false>), dim3(grid), dim3(NT), 0, stream, p);
    }
    return hipGetLastError();
}

template <int W, int DT>
static hipError_t launch_w(const CorrParams& p, bool planes, int nt, int blocks, hipStream_t stream)
{
    if (p.pairs == nullptr) {
        if (nt != 256) return hipErrorInvalidValue;
        return launch_wn<W, DT, 256, false>(p, planes, 0, stream);   // one task per wave
    }
    if (planes) return hipErrorInvalidValue;
    if (nt == 256) return launch_wn<W, DT, 256, true>(p, false, blocks, stream);
    if (nt == 64) return launch_wn<W, DT, 64, true>(p, false, blocks, stream);
    return hipErrorInvalidValue;
}

template <int DT>
static hipError_t launch_dt(const CorrParams& p, int win, bool planes, int nt, int blocks, hipStream_t stream)
{
    switch (win) {
    case 8: return launch_w<8, DT>(p, planes, nt, blocks, stream);
    case 16: return launch_w<16, DT>(p, planes, nt, blocks, stream);
    case 32: return launch_w<32, DT>(p, planes, nt, blocks, stream);
    case 64: return launch_w<64, DT>(p, planes, nt, blocks, stream);
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

} // namespace opiv
