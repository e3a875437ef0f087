// fft_reg.hpp -- fully unrolled in-register FFTs for one lane.
//
// A lane holds a whole line of N complex samples in VGPRs (float2 x[N]) and
// transforms it with a radix-2 decimation-in-time network whose loop indices
// are template constants: twiddles fold into literals, trivial twiddles (1,
// +-i) cost adds only, every other butterfly is six fused multiply-adds, and
// the bit-reversed order is a compile-time renaming (read element k as
// x[brev<N>(k)]).  No LDS and no cross-lane traffic is involved.
#pragma once

#include <hip/hip_runtime.h>

#include <utility>

namespace opiv::fftreg {

// ---- compile-time sin/cos (Taylor series in double, rounded to float) ----
constexpr double kPi = 3.141592653589793238462643383279502884;

constexpr double taylor_sin(double x)
{
    double term = x, sum = x;
    for (int n = 1; n < 16; ++n) {
        term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
        sum += term;
    }
    return sum;
}

constexpr double taylor_cos(double x)
{
    double term = 1.0, sum = 1.0;
    for (int n = 1; n < 16; ++n) {
        term *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
        sum += term;
    }
    return sum;
}

// cos and sin of 2*pi*e/n for 0 <= e < n, reduced to [0, pi/2]
constexpr double cos2pi(int e, int n)
{
    e %= n;
    if (e < 0) e += n;
    if (4 * e > 3 * n) return cos2pi(n - e, n);          // 4th quadrant: cos even
    if (2 * e > n) return -cos2pi(e - n / 2, n);         // 3rd: cos(x) = -cos(x - pi)
    if (4 * e > n) return -taylor_sin(2.0 * kPi * e / n - kPi / 2); // 2nd
    return taylor_cos(2.0 * kPi * e / n);
}
constexpr double sin2pi(int e, int n)
{
    e %= n;
    if (e < 0) e += n;
    if (4 * e > 3 * n) return -sin2pi(n - e, n);
    if (2 * e > n) return -sin2pi(e - n / 2, n);
    if (4 * e > n) return taylor_cos(2.0 * kPi * e / n - kPi / 2);
    return taylor_sin(2.0 * kPi * e / n);
}

constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }

constexpr int brev(int v, int bits)
{
    int r = 0;
    for (int i = 0; i < bits; ++i)
        if (v & (1 << i)) r |= 1 << (bits - 1 - i);
    return r;
}

// index of frequency k in the output array of an N-point network
template <int N>
constexpr int brev_n(int k) { return brev(k, ilog2c(N)); }

// ---- static_for ----
template <class F, int... Is>
__device__ __forceinline__ void sfor_impl(F&& f, std::integer_sequence<int, Is...>)
{
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void sfor(F&& f)
{
    sfor_impl(f, std::make_integer_sequence<int, N>{});
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// Radix-2 decimation-in-time butterfly (u, v) <- (u + w v, u - w v), w =
// exp(SIGN * 2*pi*i * E / N), 0 <= E < N/2, in fused multiply-adds only:
// with w = c (1 + i s/c) (or s (c/s + i) where |s| > |c|), t = v (1 + i r)
// takes two FMAs and u +- c t four, against two multiplies, two FMAs and
// four adds for the plain form.  w = 1 and w = SIGN * i are adds only.
template <int SIGN, int N, int E>
__device__ __forceinline__ void bfly_dit(float2& u, float2& v)
{
    static_assert(E >= 0 && 2 * E < N, "a DIT stage uses exponents 0 <= E < N/2");
    if constexpr (E == 0) {
        const float2 a = u;
        u = cadd(a, v);
        v = csub(a, v);
    } else if constexpr (4 * E == N) {           // w = SIGN * i
        const float2 t = SIGN > 0 ? make_float2(-v.y, v.x) : make_float2(v.y, -v.x);
        const float2 a = u;
        u = cadd(a, t);
        v = csub(a, t);
    } else {
        constexpr double c = cos2pi(E, N);
        constexpr double s = SIGN * sin2pi(E, N);
        constexpr bool cos_major = (c > 0 ? c : -c) >= (s > 0 ? s : -s);
        constexpr float m = (float)(cos_major ? c : s);
        constexpr float r = (float)(cos_major ? s / c : c / s);
        float tx, ty;
        if constexpr (cos_major) {               // w v = c ((vx - r vy) + i (vy + r vx))
            tx = __builtin_fmaf(-r, v.y, v.x);
            ty = __builtin_fmaf(r, v.x, v.y);
        } else {                                 // w v = s ((r vx - vy) + i (vx + r vy))
            tx = __builtin_fmaf(r, v.x, -v.y);
            ty = __builtin_fmaf(r, v.y, v.x);
        }
        const float2 a = u;
        u = make_float2(__builtin_fmaf(m, tx, a.x), __builtin_fmaf(m, ty, a.y));
        v = make_float2(__builtin_fmaf(-m, tx, a.x), __builtin_fmaf(-m, ty, a.y));
    }
}

// In-place N-point FFT, SIGN = -1 forward (exp(-2 pi i k n / N)), +1
// inverse, un-normalised: natural-order input, frequency k at
// x[brev_n<N>(k)].  A radix-2 DIT network run on the bit-reversed view of the
// array (logical position p is x[brev(p)]), so the input needs no reordering.
// Butterflies are bfly_dit's FMA form.
template <int N, int SIGN>
__device__ __forceinline__ void fft_dit(float2 (&x)[N])
{
    constexpr int LOG = ilog2c(N);
    sfor<LOG>([&](auto S) {
        constexpr int m = 2 << decltype(S)::value;       // span of this stage's groups
        sfor<N / m>([&](auto G) {
            constexpr int k0 = decltype(G)::value * m;
            sfor<m / 2>([&](auto J) {
                constexpr int j = decltype(J)::value;
                bfly_dit<SIGN, N, j * (N / m)>(x[brev_n<N>(k0 + j)], x[brev_n<N>(k0 + j + m / 2)]);
            });
        });
    });
}

} // namespace opiv::fftreg
