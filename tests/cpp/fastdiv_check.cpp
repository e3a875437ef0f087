// fastdiv_check.cpp -- the grid-width division of geom.hpp (a multiplier and
// shifts, used by every correlation kernel to find a window's grid row)
// against the divide instruction: every divisor up to 70000 with the
// dividends that sit at multiples of it, every power of two and its
// neighbours, and random 32-bit pairs.  Prints "ok" and exits 0, or the
// first mismatch and exits 1.
#include <cstdint>
#include <cstdio>
#include <random>

#include "geom.hpp"

static bool check(uint32_t n, uint32_t d)
{
    const uint32_t q = opiv::fastdiv(n, opiv::ceil_log2_u32(d), opiv::fastdiv_mul(d));
    if (q != n / d) {
        std::printf("mismatch: %u / %u = %u, got %u\n", n, d, n / d, q);
        return false;
    }
    return true;
}

int main()
{
    for (uint32_t d = 1; d <= 70000; ++d) {
        const uint32_t ns[] = { 0u, 1u, d - 1, d, d + 1, 2 * d - 1, 2 * d, 65025u, 1042441u, 0x7fffffffu,
                                0xfffffffeu, 0xffffffffu, (0xffffffffu / d) * d, (0xffffffffu / d) * d - 1 };
        for (uint32_t n : ns)
            if (!check(n, d)) return 1;
        for (uint32_t k = 1; k < 2100 && (uint64_t)k * d < 0xffffffffull; k += 7)
            if (!check(k * d - 1, d) || !check(k * d, d)) return 1;
    }
    for (int e = 0; e < 32; ++e) {
        const uint32_t p = 1u << e;
        for (uint32_t d : { p, p - 1, p + 1 })
            if (d && (!check(0xffffffffu, d) || !check(p, d) || !check(p - 1, d) || !check(12345678u, d)))
                return 1;
    }
    std::mt19937 rng(7);
    for (int i = 0; i < 2000000; ++i) {
        const uint32_t d = (uint32_t)rng() >> (rng() % 32);
        if (d && !check((uint32_t)rng(), d)) return 1;
    }
    std::printf("ok\n");
    return 0;
}
