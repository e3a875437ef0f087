// fit_log.hpp -- the natural logarithm that the Gaussian sub-pixel fit takes
// of its stencil values (image_utils.impl.h:86-98 applies std::log to the
// three values around the peak in each direction), in FP64, for FP32 plane
// values.  No HIP dependency: the kernels and a host check compile the same
// code (constexpr functions are host and device functions in HIP).
//
// A float is exact in double, so x = m 2^e with m in [sqrt(1/2), sqrt(2)]
// splits without rounding, and
//   ln x = e ln 2 + ln m,   ln m = 2 atanh(u) = 2 (u + u^3/3 + u^5/5 + ...),
//   u = (m - 1) / (m + 1),  |u| <= (sqrt 2 - 1) / (sqrt 2 + 1) < 0.1716,
// where m - 1 and m + 1 are exact (m has at most 24 significant bits) and
// the series stops at u^19/19 (the first term left out is below 3e-17 of the
// sum).  ln 2 is split into a 40-bit head, whose product with any exponent
// of a float is exact, and the rest.  The error is about 1 ulp, against the
// FP32 rounding of the plane values it is applied to (~6e-8 relative); the
// library logarithm spends most of its instructions on compensated sums that
// carry the result to 1 ulp.
#pragma once

#include <cstdint>

namespace opiv {

// ln x in FP64 for a float x: -inf for zero, NaN for negative or NaN x, +inf
// for +inf, like std::log of the widened value
constexpr double ln_of_float(float xf)
{
    constexpr double kSqrt2 = 0x1.6a09e667f3bcdp+0;
    constexpr double kLn2Hi = 0x1.62e42fefa3000p-1;     // 40 significant bits
    constexpr double kLn2Lo = 0x1.3de6af278ece6p-42;    // ln 2 - kLn2Hi
    const double x = (double)xf;
    // the split: x is a positive normal double for every positive finite
    // float (float subnormals included); other inputs are replaced below
    const uint64_t bits = __builtin_bit_cast(uint64_t, x);
    int e = (int)((bits >> 52) & 0x7ff) - 1023;
    double m = __builtin_bit_cast(double, (bits & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > kSqrt2) {
        m *= 0.5;
        e += 1;
    }
    const double u = (m - 1.0) / (m + 1.0);
    const double s = u * u;
    // q = 2 (s/3 + s^2/5 + ... + s^9/19), Horner from the smallest term, so
    // that ln m = 2u + u q with u q below 2% of 2u: the sum's rounding error
    // then rests on the exact 2u
    double q = 2.0 / 19.0;
    q = __builtin_fma(q, s, 2.0 / 17.0);
    q = __builtin_fma(q, s, 2.0 / 15.0);
    q = __builtin_fma(q, s, 2.0 / 13.0);
    q = __builtin_fma(q, s, 2.0 / 11.0);
    q = __builtin_fma(q, s, 2.0 / 9.0);
    q = __builtin_fma(q, s, 2.0 / 7.0);
    q = __builtin_fma(q, s, 2.0 / 5.0);
    q = __builtin_fma(q, s, 2.0 / 3.0);
    q *= s;
    const double ed = (double)e;
    // e hi (exact) + (2u + (u q + e lo)), the small terms summed first
    const double r = __builtin_fma(ed, kLn2Hi, __builtin_fma(2.0, u, __builtin_fma(u, q, ed * kLn2Lo)));
    const bool finite_pos = x > 0.0 && x < __builtin_inf();
    return finite_pos ? r : (x == 0.0 ? -__builtin_inf() : (x > 0.0 ? x : __builtin_nan("")));
}

} // namespace opiv
