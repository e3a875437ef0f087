// corr_fft128.hip -- the fused FFT correlation, peak search and Gaussian fit
// for 128 x 128 windows: the same per-window chain as corr_fft.hip
// (core::extract, algos::FFT::cross_correlate / cross_correlate_real
// fft.h:188-227 with swap_quadrants, find_peaks(.., 2, 1)
// image_utils.impl.h:37-74, fit_simple_gaussian image_utils.impl.h:76-101,
// the vector of examples/process/main.cpp:220-236), with the plane kept on
// chip (registers and LDS), never in HBM.
//
// A 128-sample line does not fit one lane's registers next to its
// transform, so two lanes share each line: one workgroup of 256 threads
// owns one window, thread (line l = tid / 2, q = tid % 2).
//   * Forward pass over a line (decimation in time): lane q holds samples
//     n = 2m + q (m = 0..63) and runs a 64-point FFT in registers; then
//     X[k] = E[k] + w^k O[k], X[k + 64] = E[k] - w^k O[k] (w = e^{-2 pi i / 128})
//     is one radix-2 stage across the lane pair (a DPP exchange), after which
//     lane q holds the frequency block k + 64 q.
//   * Inverse pass (decimation in frequency, its mirror image): lane q holds
//     the frequency block k + 64 q, the radix-2 stage across the pair comes
//     first, and the 64-point inverse FFT leaves lane q with outputs 2m + q.
// Forward columns (frames a + i b packed), an LDS transpose, forward rows,
// the cross-power spectrum P = B conj(A) of the two real windows, inverse
// rows, an LDS transpose that reads the quadrant-swapped columns, inverse
// columns.  The product uses P(-k) = conj P(k): the lane pair of row r holds
// kx in [0, 64) and [64, 128), the mirror of kx in [1, 64) lies in the other
// block, so the q = 0 lanes form P from Z(k) and Z(-k) and hand conj P to the
// q = 1 lanes of the mirror row (two LDS rounds).  The swapped plane then
// goes to LDS, where the peak search reads each element's neighbours.
//
// LDS: one plane of 128 x 132 floats (66 KB) per workgroup, reused by every
// exchange (complex transposes in two passes, real parts then imaginary
// parts); two workgroups fit per CU.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "corr_common.hpp"
#include "corr_fft.hpp"
#include "fft_reg.hpp"
#include "transform.hpp"

namespace opiv {

using fftreg::brev_n;
using fftreg::cos2pi;
using fftreg::sfor;
using fftreg::sin2pi;

namespace {

constexpr int W = 128;
constexpr int H2 = 64;
constexpr int NT = 256;
constexpr int LP = 132;           // floats per LDS plane row (128 + padding)

// x * e^{SIGN 2 pi i k / 128}, k a compile-time constant
template <int SIGN, int K>
__device__ __forceinline__ float2 twiddle128(float2 v)
{
    if constexpr (K == 0) {
        return v;
    } else {
        constexpr float c = (float)cos2pi(K, W);
        constexpr float s = (float)(SIGN * sin2pi(K, W));
        return make_float2(__builtin_fmaf(-s, v.y, v.x * c), __builtin_fmaf(s, v.x, v.y * c));
    }
}

// value of the other lane of the pair (lane ^ 1)
__device__ __forceinline__ float2 pair_swap(float2 v)
{
    return make_float2(group_partner<1>(v.x), group_partner<1>(v.y));
}

// Forward transform of one line shared by a lane pair: on entry x[m] =
// sample 2m + q, on exit x[brev(k)] = X[k + 64 q].  sg = +1 (q = 0) or -1.
__device__ __forceinline__ void fwd_line(float2 (&x)[H2], int q, float sg)
{
    fftreg::fft_dit<H2, -1>(x);
    sfor<H2>([&](auto Kc) {
        constexpr int k = decltype(Kc)::value;
        const float2 y = x[brev_n<H2>(k)];
        const float2 tw = twiddle128<-1, k>(y);
        const float2 t = q ? tw : y;                  // w^k O[k] on the odd lane
        const float2 p = pair_swap(t);
        // q = 0: E + w^k O; q = 1: E - w^k O
        x[brev_n<H2>(k)] = make_float2(__builtin_fmaf(sg, t.x, p.x), __builtin_fmaf(sg, t.y, p.y));
    });
}

// Inverse transform of one line shared by a lane pair: on entry x[k] =
// X[k + 64 q] (natural order), on exit x[brev(m)] = sample 2m + q.
__device__ __forceinline__ void inv_line(float2 (&x)[H2], int q, float sg)
{
    sfor<H2>([&](auto Kc) {
        constexpr int k = decltype(Kc)::value;
        const float2 v = x[k];
        const float2 p = pair_swap(v);
        // q = 0: X[k] + X[k+64]; q = 1: (X[k] - X[k+64]) w^{-k}
        const float2 d = make_float2(__builtin_fmaf(sg, v.x, p.x), __builtin_fmaf(sg, v.y, p.y));
        const float2 tw = twiddle128<+1, k>(d);
        x[k] = q ? tw : d;
    });
    fftreg::fft_dit<H2, +1>(x);
}

// 64-bit order-preserving key of (value, scan index): larger values win,
// equal values go to the smaller index (the scan order of find_peaks)
__device__ __forceinline__ uint64_t peak_key(float v, int idx)
{
    const uint32_t b = __float_as_uint(v);
    const uint32_t ord = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
    return ((uint64_t)ord << 32) | (uint32_t)(0x7fffffff - idx);
}
__device__ __forceinline__ float key_value(uint64_t k)
{
    const uint32_t ord = (uint32_t)(k >> 32);
    return __uint_as_float((ord & 0x80000000u) ? (ord & 0x7fffffffu) : ~ord);
}
__device__ __forceinline__ int key_index(uint64_t k) { return 0x7fffffff - (int)(uint32_t)k; }

__device__ __forceinline__ uint64_t wave_max_u64(uint64_t k)
{
    sfor<6>([&](auto S) {
        constexpr int off = 1 << decltype(S)::value;
        const uint64_t o = __shfl_xor(k, off, 64);
        k = o > k ? o : k;
    });
    return k;
}

// the largest key over the workgroup (every thread receives it)
__device__ __forceinline__ uint64_t block_max_u64(uint64_t k, uint64_t* slots)
{
    k = wave_max_u64(k);
    const int wave = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) slots[wave] = k;
    __syncthreads();
    uint64_t m = slots[0];
    sfor<NT / 64 - 1>([&](auto Wv) {
        const uint64_t o = slots[decltype(Wv)::value + 1];
        m = o > m ? o : m;
    });
    return m;
}

} // namespace

template <int DT, bool PLANES, bool BATCH>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2))) void corr_fft128_kernel(CorrParams p)
{
    __shared__ __attribute__((aligned(16))) float plane[W * LP];
    __shared__ uint64_t slots[NT / 64];
    float2* cplane = reinterpret_cast<float2*>(plane);     // complex view: rows of LP / 2 values

    const int tid = threadIdx.x;
    const int l = tid >> 1, q = tid & 1;
    const float sg = q ? -1.f : 1.f;
    const int local = xcd_block((int)blockIdx.x, (int)gridDim.x);
    if (local >= p.n_windows) return;                      // the whole workgroup
    const WinGeom g = window_geom<BATCH>(p, local);
    if (!g.valid) return;

    // ---- forward columns: line l = column x0 + l, rows 2m + q of both frames ----
    float2 z[H2];
    {
        const size_t base = (size_t)g.y0 * g.pitch + (size_t)g.x0 + (size_t)l + (size_t)q * g.pitch;
        sfor<H2>([&](auto M) {
            constexpr int m = decltype(M)::value;
            const size_t idx = base + (size_t)(2 * m) * g.pitch;
            z[m] = make_float2(load_px<DT>(g.a, idx), load_px<DT>(g.b, idx));
        });
    }
    fwd_line(z, q, sg);                    // column l: ky = k + 64 q at z[brev(k)]

    // ---- transpose (real parts, then imaginary parts): row pass line l =
    // spectral row ky = l, samples kx = 2m + q ----
    auto transpose_in = [&](auto part) {
        constexpr int pt = decltype(part)::value;
        __syncthreads();
        sfor<H2>([&](auto K) {
            constexpr int k = decltype(K)::value;
            const float2 v = z[brev_n<H2>(k)];
            plane[(k + H2 * q) * LP + l] = pt == 0 ? v.x : v.y;
        });
        __syncthreads();
    };
    float re[H2];
    transpose_in(std::integral_constant<int, 0>{});
    sfor<H2>([&](auto M) {
        constexpr int m = decltype(M)::value;
        re[m] = plane[l * LP + 2 * m + q];
    });
    transpose_in(std::integral_constant<int, 1>{});
    sfor<H2>([&](auto M) {
        constexpr int m = decltype(M)::value;
        z[m] = make_float2(re[m], plane[l * LP + 2 * m + q]);
    });
    fwd_line(z, q, sg);                    // row ky = l: kx = k + 64 q at z[brev(k)]

    // ---- cross-power spectrum 4 P = -i d conj(s), s = Z + conj Z(-k) = 2A,
    // d = Z - conj Z(-k) = 2iB, P = B conj(A), in two rounds through the
    // complex view of the plane (rows of 66 values) ----
    const int r = l, rm = (W - l) & (W - 1);                // this row and the mirror row
    auto prod = [](float2 zk, float2 zr) {
        const float sr = zk.x + zr.x, si = zk.y - zr.y;
        const float dr = zk.x - zr.x, di = zk.y + zr.y;
        return make_float2(__builtin_fmaf(di, sr, -(dr * si)), -__builtin_fmaf(dr, sr, di * si));
    };
    constexpr int CP = LP / 2;                             // complex values per row
    float2 pp[H2];                                         // P(r, k + 64 q), natural order
    __syncthreads();
    if (q == 1)                                            // round A: the upper block
        sfor<H2>([&](auto K) {
            constexpr int k = decltype(K)::value;
            cplane[r * CP + k] = z[brev_n<H2>(k)];
        });
    __syncthreads();
    if (q == 0) {
        sfor<H2 - 1>([&](auto K) {
            constexpr int k = decltype(K)::value + 1;       // kx = k, mirror kx 128 - k
            pp[k] = prod(z[brev_n<H2>(k)], cplane[rm * CP + (H2 - k)]);
        });
    } else {
        pp[0] = prod(z[brev_n<H2>(0)], cplane[rm * CP]);   // kx = 64 is its own mirror column
    }
    __syncthreads();
    if (q == 0) {                                          // round B: conj P of the mirror row
        sfor<H2 - 1>([&](auto K) {
            constexpr int k = decltype(K)::value + 1;
            cplane[rm * CP + (H2 - k)] = make_float2(pp[k].x, -pp[k].y);
        });
        cplane[r * CP + H2] = z[brev_n<H2>(0)];             // Z(r, 0) for the mirror row's lane
    }
    __syncthreads();
    if (q == 1) {
        sfor<H2 - 1>([&](auto K) {
            constexpr int k = decltype(K)::value + 1;       // kx = 64 + k
            pp[k] = cplane[r * CP + k];
        });
    } else {
        pp[0] = prod(z[brev_n<H2>(0)], cplane[rm * CP + H2]);
    }
    if (p.real_compat) {
        // fft.h:166-183 leaves ky in {0, H/2} and kx = 0 empty
        const bool empty_row = r == 0 || r == H2;
        sfor<H2>([&](auto K) {
            constexpr int k = decltype(K)::value;
            if (empty_row || (k == 0 && q == 0)) pp[k] = make_float2(0.f, 0.f);
        });
    }

    // ---- inverse rows, transpose to swapped columns, inverse columns ----
    inv_line(pp, q, sg);                   // row ky = l: x = 2m + q at pp[brev(m)]
    const int cs = (l + H2) & (W - 1);     // plane column that swaps into column l
    auto transpose_out = [&](auto part) {
        constexpr int pt = decltype(part)::value;
        __syncthreads();
        sfor<H2>([&](auto M) {
            constexpr int m = decltype(M)::value;
            const float2 v = pp[brev_n<H2>(m)];
            plane[l * LP + 2 * m + q] = pt == 0 ? v.x : v.y;
        });
        __syncthreads();
    };
    transpose_out(std::integral_constant<int, 0>{});
    sfor<H2>([&](auto K) {
        constexpr int k = decltype(K)::value;
        re[k] = plane[(k + H2 * q) * LP + cs];
    });
    transpose_out(std::integral_constant<int, 1>{});
    sfor<H2>([&](auto K) {
        constexpr int k = decltype(K)::value;
        z[k] = make_float2(re[k], plane[(k + H2 * q) * LP + cs]);
    });
    inv_line(z, q, sg);                    // swapped column l: y = 2m + q at z[brev(m)]

    // ---- the swapped plane S[h][w] = c[(h + 64) % 128][(w + 64) % 128] to
    // LDS: row h = 2m' + q of column l is y = 2((m' + 32) % 64) + q ----
    __syncthreads();
    sfor<H2>([&](auto M) {
        constexpr int mp = decltype(M)::value;
        plane[(2 * mp + q) * LP + l] = z[brev_n<H2>((mp + H2 / 2) & (H2 - 1))].x;
    });
    __syncthreads();

    // the planes carry 4x the reference's (the factor of the product)
    constexpr float kUnscale = 0.25f;
    if constexpr (PLANES) {
        double* dst = p.planes + (size_t)g.out * W * W;
        for (int i = tid; i < W * W; i += NT) dst[i] = (double)(plane[(i / W) * LP + (i % W)] * kUnscale);
        return;
    } else {
        // ---- find_peaks(.., 2, 1): thread (column l, half q) scans rows
        // [64 q, 64 q + 64) of column l; region as rect::dilate(0.5) for
        // limit-search (main.cpp:204-211) ----
        const int r0 = p.limit_search ? W / 4 : 0;
        const int rlen = p.limit_search ? W / 2 : W;
        const int hlo = max(r0 + 1, H2 * q), hhi = min(r0 + rlen - 3, H2 * q + H2 - 1);
        const bool col_ok = l >= r0 + 1 && l <= r0 + rlen - 2;
        uint64_t k1 = 0, k2 = 0;                           // 0 = no peak
        if (col_ok) {
            for (int h = hlo; h <= hhi; ++h) {             // scan order: rows ascending
                const float c = plane[h * LP + l];
                const float nb = fmaxf(fmaxf(plane[(h - 1) * LP + l], plane[(h + 1) * LP + l]),
                                       fmaxf(plane[h * LP + l - 1], plane[h * LP + l + 1]));
                if (c > nb) {
                    const uint64_t k = peak_key(c, h * W + l);
                    if (k > k1) {
                        k2 = k1;
                        k1 = k;
                    } else if (k > k2) {
                        k2 = k;
                    }
                }
            }
        }
        // top two of the window: the best key, then the best of every other
        // thread's first and the winner's second
        const uint64_t g1 = block_max_u64(k1, slots);
        const uint64_t g2 = block_max_u64(k1 == g1 ? k2 : k1, slots);
        if (tid != 0) return;
        opiv_window_result res;
        res.reserved = 0;
        if (g2 == 0) {
            // main.cpp:214-218: the default (all-zero) record stays
            res.x = res.y = res.u = res.v = res.sn = 0.0;
            res.peak0 = res.peak1 = 0.0;
            res.px = res.py = 0;
            res.flags = OPIV_FLAG_FEWER_THAN_2_PEAKS;
        } else {
            const int i1 = key_index(g1), i2 = key_index(g2);
            const int px = i1 % W, py = i1 / W, px2 = i2 % W, py2 = i2 / W;
            const float v1 = key_value(g1), v2 = key_value(g2);
            auto at = [&](int row, int col) { return plane[row * LP + col]; };
            const float l_ = at(py, px - 1), r_ = at(py, px + 1), u_ = at(py - 1, px), d_ = at(py + 1, px);
            const double lnl = ln_of_float(l_), lnc = ln_of_float(v1), lnr = ln_of_float(r_);
            const double lnu = ln_of_float(u_), lnd = ln_of_float(d_);
            // image_utils.impl.h:86-98
            const double numx = lnl - lnr, denx = 2.0 * (lnl + lnr - 2.0 * lnc);
            const double numy = lnu - lnd, deny = 2.0 * (lnu + lnd - 2.0 * lnc);
            const double plx = (double)px + (denx == 0.0 ? 0.0 : numx / denx);
            const double ply = (double)py + (deny == 0.0 ? 0.0 : numy / deny);
            uint32_t flags = 0;
            if (!isfinite(plx) || !isfinite(ply)) flags |= OPIV_FLAG_NONFINITE_FIT;
            const float fv1 = kUnscale * v1, fv2 = kUnscale * v2;
            const float nmax = kUnscale * fmaxf(fmaxf(l_, r_), fmaxf(u_, d_));
            const float nmax2 = kUnscale * fmaxf(fmaxf(at(py2, px2 - 1), at(py2, px2 + 1)),
                                                 fmaxf(at(py2 - 1, px2), at(py2 + 1, px2)));
            const float tol = p.tie_tol * fabsf(fv1);
            if (fv1 - fv2 <= tol || fv1 - nmax <= tol || fv2 - nmax2 <= tol) flags |= OPIV_FLAG_NEAR_TIE;
            if (fv1 == fv2) flags |= OPIV_FLAG_EXACT_TIE;
            // main.cpp:220-234
            const int midx = g.x0 + W / 2, midy = g.y0 + W / 2;
            res.x = (double)midx;
            res.y = (double)g.img_h - (double)midy;
            res.u = (double)midx - ((double)g.x0 + plx);
            res.v = (double)midy - ((double)g.y0 + ply);
            res.sn = (fv2 > 0.f) ? (double)fv1 / (double)fv2 : 0.0;
            res.peak0 = (double)fv1;
            res.peak1 = (double)fv2;
            res.px = px;
            res.py = py;
            res.flags = flags;
        }
        p.out[g.out] = res;
    }
}

// FFT::transform (fft.h:86-119) of 128 x 128 windows of interleaved complex
// FP64 values, un-normalised both ways, FP32 arithmetic: one 256-thread
// workgroup per window, the lane pairs and transposes of the correlation
// kernel (columns, then rows).  The reverse transform runs the forward
// network on conjugates, IDFT(x) = conj(DFT(conj(x))).  Every thread has read
// its window before any thread writes it, so in == out is allowed.
template <bool FORWARD>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2)))
void fft128_transform_kernel(const double2* in, double2* out, int n)
{
    __shared__ __attribute__((aligned(16))) float plane[W * LP];
    const int tid = threadIdx.x;
    const int l = tid >> 1, q = tid & 1;
    const float sg = q ? -1.f : 1.f;
    const float cj = FORWARD ? 1.f : -1.f;                 // conjugation of the reverse transform
    if ((int)blockIdx.x >= n) return;                      // the whole workgroup
    const size_t base = (size_t)blockIdx.x * W * W;

    // ---- column l: rows 2m + q ----
    float2 z[H2];
    sfor<H2>([&](auto M) {
        constexpr int m = decltype(M)::value;
        const double2 v = in[base + (size_t)(2 * m + q) * W + l];
        z[m] = make_float2((float)v.x, cj * (float)v.y);
    });
    fwd_line(z, q, sg);                                    // column l: ky = k + 64 q at z[brev(k)]

    // ---- transpose (real parts, then imaginary parts): line l = row ky = l ----
    auto transpose_in = [&](auto part) {
        constexpr int pt = decltype(part)::value;
        __syncthreads();
        sfor<H2>([&](auto K) {
            constexpr int k = decltype(K)::value;
            const float2 v = z[brev_n<H2>(k)];
            plane[(k + H2 * q) * LP + l] = pt == 0 ? v.x : v.y;
        });
        __syncthreads();
    };
    float re[H2];
    transpose_in(std::integral_constant<int, 0>{});
    sfor<H2>([&](auto M) {
        constexpr int m = decltype(M)::value;
        re[m] = plane[l * LP + 2 * m + q];
    });
    transpose_in(std::integral_constant<int, 1>{});
    sfor<H2>([&](auto M) {
        constexpr int m = decltype(M)::value;
        z[m] = make_float2(re[m], plane[l * LP + 2 * m + q]);
    });
    fwd_line(z, q, sg);                                    // row ky = l: kx = k + 64 q at z[brev(k)]

    sfor<H2>([&](auto K) {
        constexpr int k = decltype(K)::value;
        const float2 v = z[brev_n<H2>(k)];
        out[base + (size_t)l * W + k + H2 * q] = make_double2((double)v.x, (double)(cj * v.y));
    });
}

hipError_t launch_fft_transform128(const double* in, double* out, size_t n, bool forward, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    if (n > 0x7fffffffull) return hipErrorInvalidValue;
    const auto* src = reinterpret_cast<const double2*>(in);
    auto* dst = reinterpret_cast<double2*>(out);
    if (forward)
        hipLaunchKernelGGL((fft128_transform_kernel<true>), dim3((unsigned)n), dim3(NT), 0, stream, src, dst, (int)n);
    else
        hipLaunchKernelGGL((fft128_transform_kernel<false>), dim3((unsigned)n), dim3(NT), 0, stream, src, dst, (int)n);
    return hipGetLastError();
}

template <int DT>
static hipError_t launch128_dt(const CorrParams& p, bool planes, hipStream_t stream)
{
    const dim3 grid((unsigned)p.n_windows), block(NT);
    if (p.pairs != nullptr) {
        if (planes) return hipErrorInvalidValue;
        hipLaunchKernelGGL((corr_fft128_kernel<DT, false, true>), grid, block, 0, stream, p);
    } else if (planes) {
        hipLaunchKernelGGL((corr_fft128_kernel<DT, true, false>), grid, block, 0, stream, p);
    } else {
        hipLaunchKernelGGL((corr_fft128_kernel<DT, false, false>), grid, block, 0, stream, p);
    }
    return hipGetLastError();
}

hipError_t launch_corr_fft128(const CorrParams& p, int dtype, bool planes, hipStream_t stream)
{
    if (p.n_windows <= 0) return hipSuccess;
    switch (dtype) {
    case 0: return launch128_dt<0>(p, planes, stream);
    case 1: return launch128_dt<1>(p, planes, stream);
    case 2: return launch128_dt<2>(p, planes, stream);
    case 3: return launch128_dt<3>(p, planes, stream);
    default: return hipErrorInvalidValue;
    }
}

} // namespace opiv
