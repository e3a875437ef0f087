// fft_reg.hpp -- fully unrolled in-register FFTs for one lane.
//
// A lane holds a whole line of N complex samples in VGPRs (float2 x[N]) and
// transforms it with a radix-2 decimation-in-frequency network whose loop
// indices are template constants: twiddles fold into literals, trivial
// twiddles (1, +-i, (1+-i)/sqrt2) cost no multiplies, and the bit-reversed
// output order is a compile-time renaming (read element k as x[brev<N>(k)]).
// No LDS and no cross-lane traffic is involved.
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

// index of frequency/sample k in the DIF output array of an N-point network
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

// d * exp(SIGN * 2*pi*i * E / N) with the trivial cases folded
template <int SIGN, int N, int E>
__device__ __forceinline__ float2 twiddle(float2 d)
{
    constexpr int e = ((E % N) + N) % N;
    if constexpr (e == 0) {
        return d;
    } else if constexpr (4 * e == N) {          // SIGN * i
        return SIGN > 0 ? make_float2(-d.y, d.x) : make_float2(d.y, -d.x);
    } else if constexpr (2 * e == N) {          // -1
        return make_float2(-d.x, -d.y);
    } else if constexpr (4 * e == 3 * N) {      // -SIGN * i
        return SIGN > 0 ? make_float2(d.y, -d.x) : make_float2(-d.y, d.x);
    } else if constexpr (8 * e == N || 8 * e == 3 * N || 8 * e == 5 * N || 8 * e == 7 * N) {
        // |cos| == |sin| == sqrt(2)/2: two adds and two multiplies
        constexpr float c = (float)cos2pi(e, N);
        constexpr float s = (float)(SIGN * sin2pi(e, N));
        constexpr float m = c > 0 ? c : -c;
        // (x + iy)(c + is) = (cx - sy) + i(sx + cy)
        constexpr float cs = c > 0 ? 1.f : -1.f;
        constexpr float ss = s > 0 ? 1.f : -1.f;
        return make_float2(m * (cs * d.x - ss * d.y), m * (ss * d.x + cs * d.y));
    } else {
        constexpr float c = (float)cos2pi(e, N);
        constexpr float s = (float)(SIGN * sin2pi(e, N));
        return make_float2(__builtin_fmaf(d.x, c, -(d.y * s)), __builtin_fmaf(d.x, s, d.y * c));
    }
}

// In-place N-point DIF FFT, SIGN = -1 forward (exp(-2 pi i k n / N)), +1
// inverse, un-normalised.  Output frequency k is x[brev_n<N>(k)].
template <int N, int SIGN>
__device__ __forceinline__ void fft_dif(float2 (&x)[N])
{
    constexpr int LOG = ilog2c(N);
    sfor<LOG>([&](auto S) {
        constexpr int len = N >> decltype(S)::value;
        constexpr int half = len / 2;
        sfor<N / len>([&](auto G) {
            constexpr int start = decltype(G)::value * len;
            sfor<half>([&](auto K) {
                constexpr int k = decltype(K)::value;
                const float2 u = x[start + k];
                const float2 v = x[start + k + half];
                x[start + k] = cadd(u, v);
                x[start + k + half] = twiddle<SIGN, N, k * (N / len)>(csub(u, v));
            });
        });
    });
}

} // namespace opiv::fftreg
