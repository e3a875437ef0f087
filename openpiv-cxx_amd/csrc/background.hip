// background.hip -- examples/average_subtract/main.cpp:98-159 on the GPU: the
// mean of inputs 1..n-1 (the reference sizes its accumulator from input 0
// and adds every later input, main.cpp:98-122), subtracted from every input
// (main.cpp:146), and the 16-bit image the PNM writer saves for each result
// (pnm_image_loader.cpp:367-392, algos::find_image_range stats.h:15-34).
//
// Every pixel's FP64 arithmetic is the reference's, in the same order (sums
// in input order, correctly rounded division and truncation), so the outputs
// are bit-identical.  Three bandwidth-bound passes: the mean; per-frame
// min/max of frame - mean (ordered-integer atomics); the scaled images.
// Frames are handled in chunks of 8 per thread so the mean is read once per
// chunk; the scale pass replaces the division by a product with a proven
// fallback (scale16).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "background.hpp"
#include "fft_reg.hpp"

#pragma clang fp contract(off)

namespace opiv {

using fftreg::sfor;

namespace {

constexpr int CH = 8;                  // frames per chunk

template <class T>
using gptr = const __attribute__((address_space(1))) T*;

template <int DT>
__device__ __forceinline__ double load_d(const void* base, size_t idx)
{
    if constexpr (DT == 0) return (double)((gptr<uint8_t>)base)[idx];
    else if constexpr (DT == 1) return (double)((gptr<uint16_t>)base)[idx];
    else if constexpr (DT == 2) return (double)((gptr<float>)base)[idx];
    else return ((gptr<double>)base)[idx];
}

// order-preserving map of doubles (no NaNs) to unsigned 64-bit keys
__device__ __forceinline__ unsigned long long okey(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double ovalue(unsigned long long k)
{
    const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)b);
}

} // namespace

// Thread layout of the three passes: a block covers 256 * PX consecutive
// pixels of one row; thread t handles columns x0 + t + 256 k (k < PX), so
// every load instruction stays coalesced, each thread keeps PX independent
// pixels in flight and a block reads 2 KB runs (uint16) of every frame.
constexpr int PX = 4;
constexpr int SPAN = 256 * PX;

// mean[y][x] = (f1 + ... + f(n-1)) / (n - 1), added in input order
template <int DT>
__global__ __launch_bounds__(256) void bg_mean_kernel(BgParams p)
{
    const int x0 = blockIdx.x * SPAN + (int)threadIdx.x;
    const int y = blockIdx.y;
    const size_t row = (size_t)y * p.pitch;
    double acc[PX];
    sfor<PX>([&](auto K) { acc[decltype(K)::value] = 0.0; });
    constexpr int U = 4;                         // frames loaded ahead of their adds
    for (int i = 1; i < p.n; i += U) {
        double v[U][PX];
        sfor<U>([&](auto Uu) {
            constexpr int u = decltype(Uu)::value;
            const bool fu = i + u < p.n;
            const void* f = p.frames[fu ? i + u : i];
            sfor<PX>([&](auto K) {
                constexpr int k = decltype(K)::value;
                const int x = x0 + 256 * k;
                v[u][k] = (fu && x < p.w) ? load_d<DT>(f, row + x) : 0.0;
            });
        });
        sfor<U>([&](auto Uu) {
            constexpr int u = decltype(Uu)::value;
            if (i + u < p.n) {
                sfor<PX>([&](auto K) {
                    constexpr int k = decltype(K)::value;
                    acc[k] = acc[k] + v[u][k];
                });
            }
        });
    }
    const double div = (double)(p.n - 1);
    sfor<PX>([&](auto K) {
        constexpr int k = decltype(K)::value;
        const int x = x0 + 256 * k;
        if (x < p.w) p.mean[(size_t)y * p.w + x] = acc[k] / div;
    });
}

__global__ void bg_range_init_kernel(unsigned long long* range, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        range[2 * i] = ~0ull;
        range[2 * i + 1] = 0ull;
    }
}

// min and max of frame - mean for the CH frames of chunk blockIdx.z; rows
// strided over gridDim.y so each thread folds many pixels before the block
// reduction and one atomic per frame and block
template <int DT>
__global__ __launch_bounds__(256) void bg_range_kernel(BgParams p)
{
    const int f0 = blockIdx.z * CH;
    const int x0 = blockIdx.x * SPAN + (int)threadIdx.x;
    const void* fr[CH];
    sfor<CH>([&](auto J) {
        constexpr int j = decltype(J)::value;
        fr[j] = p.frames[f0 + j < p.n ? f0 + j : f0];
    });
    double mn[CH], mx[CH];
    sfor<CH>([&](auto J) {
        constexpr int j = decltype(J)::value;
        mn[j] = __longlong_as_double(0x7ff0000000000000ll);     // +inf
        mx[j] = -mn[j];
    });
    for (int y = blockIdx.y; y < p.h; y += gridDim.y) {
        const size_t row = (size_t)y * p.pitch;
        sfor<PX>([&](auto K) {
            constexpr int k = decltype(K)::value;
            const int x = x0 + 256 * k;
            if (x < p.w) {
                const double m = p.mean[(size_t)y * p.w + x];
                sfor<CH>([&](auto J) {
                    constexpr int j = decltype(J)::value;
                    if (f0 + j < p.n) {
                        const double d = load_d<DT>(fr[j], row + x) - m;
                        mn[j] = d < mn[j] ? d : mn[j];
                        mx[j] = d > mx[j] ? d : mx[j];
                    }
                });
            }
        });
    }
    __shared__ double smn[4][CH], smx[4][CH];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    sfor<CH>([&](auto J) {
        constexpr int j = decltype(J)::value;
        double a = mn[j], b = mx[j];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const double oa = __shfl_xor(a, off, 64), ob = __shfl_xor(b, off, 64);
            a = oa < a ? oa : a;
            b = ob > b ? ob : b;
        }
        if (lane == 0) {
            smn[wave][j] = a;
            smx[wave][j] = b;
        }
    });
    __syncthreads();
    if (threadIdx.x < CH && f0 + (int)threadIdx.x < p.n) {
        const int j = threadIdx.x;
        double a = smn[0][j], b = smx[0][j];
        for (int w = 1; w < 4; ++w) {
            a = smn[w][j] < a ? smn[w][j] : a;
            b = smx[w][j] > b ? smx[w][j] : b;
        }
        if (a <= b) {                            // the block saw pixels
            atomicMin(&p.range[2 * (f0 + j)], okey(a));
            atomicMax(&p.range[2 * (f0 + j) + 1], okey(b));
        }
    }
}

// trunc(65535 * fl(x / range)) for x in [0, range].  x * fl(1 / range) is
// within 4 ulp of fl(x / range), so after scaling by 65535 the two differ by
// less than 4e-11: whenever the fast product lies 1e-9 or more from an
// integer both truncate alike; otherwise the correctly rounded quotient is
// used.  Identical to the reference's expression for every input.
__device__ __forceinline__ uint16_t scale16(double x, double range, double rinv)
{
    const double v = 65535.0 * (x * rinv);
    const double fl = floor(v);
    if (v - fl >= 1e-9 && fl + 1.0 - v >= 1e-9) return (uint16_t)fl;
    return (uint16_t)(65535.0 * (x / range));
}

// out = (uint16)(65535 * ((d - min) / range)), range = max - min or 1
template <int DT>
__global__ __launch_bounds__(256) void bg_scale_kernel(BgParams p)
{
    const int x0 = blockIdx.x * SPAN + (int)threadIdx.x;
    const int y = blockIdx.y;
    const int f0 = blockIdx.z * CH;
    const size_t row = (size_t)y * p.pitch;
    double m[PX];
    sfor<PX>([&](auto K) {
        constexpr int k = decltype(K)::value;
        const int x = x0 + 256 * k;
        m[k] = x < p.w ? p.mean[(size_t)y * p.w + x] : 0.0;
    });
    sfor<CH>([&](auto J) {
        constexpr int j = decltype(J)::value;
        const int f = f0 + j;
        if (f < p.n) {
            const double lo = ovalue(p.range[2 * f]), hi = ovalue(p.range[2 * f + 1]);
            double range = hi - lo;
            if (hi == lo) range = 1.0;
            const double rinv = 1.0 / range;
            const void* fr = p.frames[f];
            uint16_t* o = p.out[f] + (size_t)y * p.w;
            sfor<PX>([&](auto K) {
                constexpr int k = decltype(K)::value;
                const int x = x0 + 256 * k;
                if (x < p.w) {
                    const double d = load_d<DT>(fr, row + x) - m[k];
                    o[x] = scale16(d - lo, range, rinv);
                }
            });
        }
    });
}

// ---- the same three passes on runs of V = 8 consecutive pixels per thread,
// each run one (or, for 4- and 8-byte pixels, two or four) 16-byte load, each
// output run one 16-byte store: a wave moves 1-4 KB per instruction where the
// forms above move 128 B.  Every pixel's arithmetic is the same. ----
constexpr int V = 8;
constexpr int VSPAN = 256 * V;

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

template <int DT>
__device__ __forceinline__ void load_run(const void* base, size_t idx, double (&v)[V])
{
    if constexpr (DT == 0) {
        const u32x2 w = *(gptr<u32x2>)((gptr<uint8_t>)base + idx);
        sfor<V>([&](auto E) {
            constexpr int e = decltype(E)::value;
            v[e] = (double)(((e < 4 ? w.x : w.y) >> (8 * (e & 3))) & 0xffu);
        });
    } else if constexpr (DT == 1) {
        const u32x4 w = *(gptr<u32x4>)((gptr<uint16_t>)base + idx);
        const uint32_t q[4] = {w.x, w.y, w.z, w.w};
        sfor<V>([&](auto E) {
            constexpr int e = decltype(E)::value;
            v[e] = (double)((q[e / 2] >> (16 * (e & 1))) & 0xffffu);
        });
    } else if constexpr (DT == 2) {
        const gptr<f32x4> f = (gptr<f32x4>)((gptr<float>)base + idx);
        const f32x4 a = f[0], b = f[1];
        const float q[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        sfor<V>([&](auto E) { v[decltype(E)::value] = (double)q[decltype(E)::value]; });
    } else {
        const gptr<f64x2> f = (gptr<f64x2>)((gptr<double>)base + idx);
        sfor<V / 2>([&](auto E) {
            constexpr int e = decltype(E)::value;
            const f64x2 d = f[e];
            v[2 * e] = d.x;
            v[2 * e + 1] = d.y;
        });
    }
}

__device__ __forceinline__ void load_mean(const double* mean, size_t idx, double (&m)[V])
{
    const gptr<f64x2> f = (gptr<f64x2>)((gptr<double>)mean + idx);
    sfor<V / 2>([&](auto E) {
        constexpr int e = decltype(E)::value;
        const f64x2 d = f[e];
        m[2 * e] = d.x;
        m[2 * e + 1] = d.y;
    });
}

template <int DT>
__global__ __launch_bounds__(256) void bg_mean_vec_kernel(BgParams p)
{
    const int x = (blockIdx.x * 256 + (int)threadIdx.x) * V;
    const int y = blockIdx.y;
    if (x >= p.w) return;
    const size_t row = (size_t)y * p.pitch;
    double acc[V];
    sfor<V>([&](auto E) { acc[decltype(E)::value] = 0.0; });
    constexpr int U = 2;                         // frames loaded ahead of their adds
    for (int i = 1; i < p.n; i += U) {
        double v[U][V];
        sfor<U>([&](auto Uu) {
            constexpr int u = decltype(Uu)::value;
            load_run<DT>(p.frames[i + u < p.n ? i + u : i], row + x, v[u]);
        });
        sfor<U>([&](auto Uu) {
            constexpr int u = decltype(Uu)::value;
            if (i + u < p.n)
                sfor<V>([&](auto E) {
                    constexpr int e = decltype(E)::value;
                    acc[e] = acc[e] + v[u][e];
                });
        });
    }
    const double div = (double)(p.n - 1);
    double2* dst = reinterpret_cast<double2*>(p.mean + (size_t)y * p.w + x);
    sfor<V / 2>([&](auto E) {
        constexpr int e = decltype(E)::value;
        dst[e] = make_double2(acc[2 * e] / div, acc[2 * e + 1] / div);
    });
}

template <int DT>
__global__ __launch_bounds__(256) void bg_range_vec_kernel(BgParams p)
{
    const int f0 = blockIdx.z * CH;
    const int x = (blockIdx.x * 256 + (int)threadIdx.x) * V;
    const bool inside = x < p.w;
    const void* fr[CH];
    sfor<CH>([&](auto J) {
        constexpr int j = decltype(J)::value;
        fr[j] = p.frames[f0 + j < p.n ? f0 + j : f0];
    });
    double mn[CH], mx[CH];
    sfor<CH>([&](auto J) {
        constexpr int j = decltype(J)::value;
        mn[j] = __longlong_as_double(0x7ff0000000000000ll);     // +inf
        mx[j] = -mn[j];
    });
    if (inside) {
        for (int y = blockIdx.y; y < p.h; y += gridDim.y) {
            const size_t row = (size_t)y * p.pitch;
            double m[V];
            load_mean(p.mean, (size_t)y * p.w + x, m);
            sfor<CH>([&](auto J) {
                constexpr int j = decltype(J)::value;
                if (f0 + j < p.n) {
                    double v[V];
                    load_run<DT>(fr[j], row + x, v);
                    sfor<V>([&](auto E) {
                        constexpr int e = decltype(E)::value;
                        const double d = v[e] - m[e];
                        mn[j] = d < mn[j] ? d : mn[j];
                        mx[j] = d > mx[j] ? d : mx[j];
                    });
                }
            });
        }
    }
    __shared__ double smn[4][CH], smx[4][CH];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    sfor<CH>([&](auto J) {
        constexpr int j = decltype(J)::value;
        double a = mn[j], b = mx[j];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const double oa = __shfl_xor(a, off, 64), ob = __shfl_xor(b, off, 64);
            a = oa < a ? oa : a;
            b = ob > b ? ob : b;
        }
        if (lane == 0) {
            smn[wave][j] = a;
            smx[wave][j] = b;
        }
    });
    __syncthreads();
    if (threadIdx.x < CH && f0 + (int)threadIdx.x < p.n) {
        const int j = threadIdx.x;
        double a = smn[0][j], b = smx[0][j];
        for (int w = 1; w < 4; ++w) {
            a = smn[w][j] < a ? smn[w][j] : a;
            b = smx[w][j] > b ? smx[w][j] : b;
        }
        if (a <= b) {                            // the block saw pixels
            atomicMin(&p.range[2 * (f0 + j)], okey(a));
            atomicMax(&p.range[2 * (f0 + j) + 1], okey(b));
        }
    }
}

template <int DT>
__global__ __launch_bounds__(256) void bg_scale_vec_kernel(BgParams p)
{
    const int x = (blockIdx.x * 256 + (int)threadIdx.x) * V;
    const int y = blockIdx.y;
    const int f0 = blockIdx.z * CH;
    if (x >= p.w) return;
    const size_t row = (size_t)y * p.pitch;
    double m[V];
    load_mean(p.mean, (size_t)y * p.w + x, m);
    sfor<CH>([&](auto J) {
        constexpr int j = decltype(J)::value;
        const int f = f0 + j;
        if (f < p.n) {
            const double lo = ovalue(p.range[2 * f]), hi = ovalue(p.range[2 * f + 1]);
            double range = hi - lo;
            if (hi == lo) range = 1.0;
            const double rinv = 1.0 / range;
            double v[V];
            load_run<DT>(p.frames[f], row + x, v);
            uint32_t o[V / 2];
            sfor<V / 2>([&](auto E) {
                constexpr int e = decltype(E)::value;
                const uint32_t lo16 = scale16(v[2 * e] - m[2 * e] - lo, range, rinv);
                const uint32_t hi16 = scale16(v[2 * e + 1] - m[2 * e + 1] - lo, range, rinv);
                o[e] = lo16 | (hi16 << 16);
            });
            *reinterpret_cast<uint4*>(p.out[f] + (size_t)y * p.w + x) = make_uint4(o[0], o[1], o[2], o[3]);
        }
    });
}

template <int DT>
static hipError_t launch_dt(const BgParams& p, hipStream_t stream)
{
    const unsigned gx = (unsigned)((p.w + (p.vec ? VSPAN : SPAN) - 1) / (p.vec ? VSPAN : SPAN)), gy = (unsigned)p.h;
    const unsigned chunks = (unsigned)((p.n + CH - 1) / CH);
    // about 8 rows of pixels per thread in the range pass
    const unsigned ry = gy < 8 ? gy : (gy + 7) / 8;
    if (p.vec) {
        hipLaunchKernelGGL((bg_mean_vec_kernel<DT>), dim3(gx, gy), dim3(256), 0, stream, p);
        hipLaunchKernelGGL(bg_range_init_kernel, dim3((unsigned)((p.n + 255) / 256)), dim3(256), 0, stream, p.range,
                           p.n);
        hipLaunchKernelGGL((bg_range_vec_kernel<DT>), dim3(gx, ry, chunks), dim3(256), 0, stream, p);
        hipLaunchKernelGGL((bg_scale_vec_kernel<DT>), dim3(gx, gy, chunks), dim3(256), 0, stream, p);
        return hipGetLastError();
    }
    hipLaunchKernelGGL((bg_mean_kernel<DT>), dim3(gx, gy), dim3(256), 0, stream, p);
    hipLaunchKernelGGL(bg_range_init_kernel, dim3((unsigned)((p.n + 255) / 256)), dim3(256), 0, stream, p.range, p.n);
    hipLaunchKernelGGL((bg_range_kernel<DT>), dim3(gx, ry, chunks), dim3(256), 0, stream, p);
    hipLaunchKernelGGL((bg_scale_kernel<DT>), dim3(gx, gy, chunks), dim3(256), 0, stream, p);
    return hipGetLastError();
}

hipError_t launch_average_subtract(const BgParams& p, int dtype, hipStream_t stream)
{
    if (p.n < 2 || p.w <= 0 || p.h <= 0) return hipErrorInvalidValue;
    switch (dtype) {
    case 0: return launch_dt<0>(p, stream);
    case 1: return launch_dt<1>(p, stream);
    case 2: return launch_dt<2>(p, stream);
    case 3: return launch_dt<3>(p, stream);
    default: return hipErrorInvalidValue;
    }
}

} // namespace opiv
