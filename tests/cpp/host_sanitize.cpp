// host_sanitize.cpp -- drives the library's host-only code (the PNM/TIFF
// decoder of image_decode.cpp and the plan / batch descriptor checks of
// plan_check.cpp) with valid, truncated, bit-flipped and crafted inputs.
// Built with g++ -fsanitize=address,undefined by tests/test_host_sanitizers.py;
// any out-of-bounds access, overflow or other undefined behaviour aborts it.
//
//   host_sanitize FILE...     (image files to decode and mutate)
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <random>
#include <string>
#include <vector>

#include "image_decode.hpp"
#include "plan_check.hpp"

namespace {

int decode_all(const std::vector<uint8_t>& d, const char* name)
{
    opiv_image_info info{};
    int rc = opiv::io::decode_buffer(d.data(), d.size(), name, 0, &info, nullptr);
    if (rc != OPIV_OK) return rc;
    for (uint32_t i = 0; i < info.count && i < 4; ++i) {
        std::vector<uint8_t> px;
        opiv_image_info pi{};
        rc = opiv::io::decode_buffer(d.data(), d.size(), name, i, &pi, &px);
        if (rc == OPIV_OK) {
            const size_t want = (size_t)pi.width * pi.height * (pi.dtype == OPIV_U16 ? 2 : 1);
            if (px.size() != want) {
                std::fprintf(stderr, "%s page %u: %zu bytes decoded, %zu expected\n", name, i, px.size(), want);
                std::abort();
            }
        }
    }
    return rc;
}

void fuzz_file(const std::vector<uint8_t>& d, const char* name, std::mt19937_64& rng, size_t& ok, size_t& bad)
{
    auto count = [&](int rc) { (rc == OPIV_OK ? ok : bad)++; };
    count(decode_all(d, name));
    // every truncation length near the header, then a spread through the data
    for (size_t n = 0; n < std::min<size_t>(d.size(), 600); ++n) count(decode_all({ d.begin(), d.begin() + n }, name));
    for (int k = 0; k < 64; ++k) {
        const size_t n = (size_t)(rng() % (d.size() + 1));
        count(decode_all({ d.begin(), d.begin() + n }, name));
    }
    // flipped bytes, mostly in the first 512 (headers, IFDs) and anywhere
    for (int k = 0; k < 400; ++k) {
        std::vector<uint8_t> m = d;
        const int flips = 1 + (int)(rng() % 4);
        for (int f = 0; f < flips && !m.empty(); ++f) {
            const size_t lim = (k % 2) ? m.size() : std::min<size_t>(m.size(), 512);
            const size_t at = (size_t)(rng() % lim);
            m[at] = (k % 3 == 0) ? 0xff : (uint8_t)rng();
        }
        count(decode_all(m, name));
    }
}

void crafted(size_t& ok, size_t& bad)
{
    auto run = [&](const std::string& s) {
        std::vector<uint8_t> d(s.begin(), s.end());
        d.resize(d.size() + 64, 0);
        (decode_all(d, "crafted") == OPIV_OK ? ok : bad)++;
    };
    const char* dims[] = { "4294967295 4294967295", "2147516416 4294934528", "0 5", "5 0", "99999999999 1", "1 1" };
    for (const char* dm : dims)
        for (const char* mx : { "255", "65535", "65536", "0" })
            for (const char* kind : { "P5", "P6" }) run(std::string(kind) + "\n" + dm + "\n" + mx + "\n");
    run("P5\n# only a comment");
    run("P6");
    run("P5\n1");
}

void plans(std::mt19937_64& rng, size_t& ok, size_t& bad)
{
    const uint32_t u32s[] = { 0u, 1u, 7u, 8u, 16u, 31u, 32u, 64u, 128u, 4096u, 65535u, 0x7fffffffu, 0xffffffffu };
    const double overlaps[] = { 0.0, 0.5, 1.0, -0.1, 1.5, std::numeric_limits<double>::quiet_NaN(),
                                std::numeric_limits<double>::infinity(), 1e-300 };
    const size_t pitches[] = { 0, 1, 4096, (size_t)1 << 24, (size_t)1 << 40, std::numeric_limits<size_t>::max() };
    auto pick = [&](auto& arr) { return arr[rng() % (sizeof(arr) / sizeof(arr[0]))]; };
    for (int k = 0; k < 20000; ++k) {
        opiv_plan_desc d{};
        d.img_w = pick(u32s);
        d.img_h = pick(u32s);
        d.pitch = pick(pitches);
        d.dtype = (int32_t)(rng() % 6) - 1;
        d.win = pick(u32s);
        d.step_x = (k % 3) ? 0 : pick(u32s);
        d.step_y = (k % 3) ? 0 : pick(u32s);
        d.overlap = pick(overlaps);
        d.variant = (int32_t)(rng() % 6) - 1;
        d.limit_search = (int32_t)(rng() % 2);
        opiv::detail::PlanGeom g;
        (opiv::detail::check_plan_desc(&d, &g) == OPIV_OK ? ok : bad)++;

        opiv_pair_desc pr{};
        pr.a = (const void*)(uintptr_t)(rng() % 3 == 0 ? 0 : 4096 * (1 + rng() % 8) + (rng() % 3));
        pr.b = pr.a;
        pr.img_w = pick(u32s);
        pr.img_h = pick(u32s);
        pr.pitch = pick(pitches);
        opiv::BatchPair e{};
        (opiv::detail::check_batch_pair(pr, (size_t)k, 0, pick(u32s), pick(overlaps), (int)(rng() % 5) - 1, &e) ==
                 OPIV_OK ? ok : bad)++;
        uint32_t nx, ny;
        (opiv_grid(pick(u32s), pick(u32s), pick(u32s), pick(u32s), pick(u32s), pick(u32s), &nx, &ny, nullptr) ==
                 OPIV_OK ? ok : bad)++;
    }
}

} // namespace

int main(int argc, char** argv)
{
    std::mt19937_64 rng(20261204);
    size_t ok = 0, bad = 0;
    for (int i = 1; i < argc; ++i) {
        std::ifstream is(argv[i], std::ios::binary);
        std::vector<uint8_t> d((std::istreambuf_iterator<char>(is)), std::istreambuf_iterator<char>());
        if (d.empty()) {
            std::fprintf(stderr, "cannot read %s\n", argv[i]);
            return 2;
        }
        fuzz_file(d, argv[i], rng, ok, bad);
    }
    crafted(ok, bad);
    plans(rng, ok, bad);
    std::printf("host_sanitize: %zu accepted, %zu rejected, no sanitizer report\n", ok, bad);
    return 0;
}
