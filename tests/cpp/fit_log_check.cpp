// Host check of opiv::ln_of_float (fit_log.hpp), the FP64 logarithm of the
// Gaussian fit: every float of [2^-3, 2^3), every 61st float of the whole
// positive range (subnormals and the largest finite values included) and the
// special values, against the x87 extended-precision logl of the same value.
// Prints the worst error in ulp of the double result and "ok" when it is at
// most 1.5 ulp everywhere and the special values come out as std::log's.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "fit_log.hpp"

static double ulp_error(float x)
{
    const double got = opiv::ln_of_float(x);
    const long double ref = logl((long double)x);
    const double refd = (double)ref;
    if (refd == 0.0) return got == 0.0 ? 0.0 : 1e9;
    int ex;
    std::frexp(refd, &ex);
    const long double ulp = std::ldexp(1.0L, ex - 53);   // ulp of the double result
    return (double)(std::fabs((long double)got - ref) / ulp);
}

int main()
{
    double worst = 0.0;
    float worst_x = 0.f;
    auto visit = [&](uint32_t b) {
        float x;
        std::memcpy(&x, &b, 4);
        const double e = ulp_error(x);
        if (e > worst) {
            worst = e;
            worst_x = x;
        }
    };
    uint32_t lo, hi;
    {
        const float a = 0.125f, b = 8.0f;
        std::memcpy(&lo, &a, 4);
        std::memcpy(&hi, &b, 4);
    }
    for (uint32_t b = lo; b < hi; ++b) visit(b);
    const uint32_t inf_bits = 0x7f800000u;
    for (uint32_t b = 1; b < inf_bits; b += 61) visit(b);
    for (uint32_t b = inf_bits - 4096; b < inf_bits; ++b) visit(b);   // the largest finite floats
    for (uint32_t b = 1; b < 4096; ++b) visit(b);                      // the smallest subnormals
    bool special = true;
    special &= std::isinf(opiv::ln_of_float(0.0f)) && opiv::ln_of_float(0.0f) < 0;
    special &= std::isinf(opiv::ln_of_float(-0.0f)) && opiv::ln_of_float(-0.0f) < 0;
    special &= std::isnan(opiv::ln_of_float(-1.0f));
    special &= std::isnan(opiv::ln_of_float(-INFINITY));
    special &= std::isnan(opiv::ln_of_float(NAN));
    special &= std::isinf(opiv::ln_of_float(INFINITY)) && opiv::ln_of_float(INFINITY) > 0;
    special &= opiv::ln_of_float(1.0f) == 0.0;
    std::printf("worst %.3f ulp at %.9g\n", worst, (double)worst_x);
    std::printf("%s\n", worst <= 1.5 && special ? "ok" : "FAIL");
    return worst <= 1.5 && special ? 0 : 1;
}
