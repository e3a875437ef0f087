// dropin_core.cpp -- examples/process's correlator registry and per-window
// processor (main.cpp:153-237) compiled against openpiv-cxx's own headers,
// once with algos::FFT and once with the engine's algos::HipFFT
// (include/openpiv_amd_core.hpp) in the same registry, so the two run side
// by side on the same core::gf_image frames: core::extract, the correlator,
// core::find_peaks(.., 2, 1) and core::fit_simple_gaussian are the
// reference's, only the class name of the correlator changes.  Built by
// oracle/Makefile into oracle/_ref/dropin_core (a test driver; it includes
// the reference's headers, so it is built only where they are mounted).
//
//   dropin_core check              size and error behaviour (no GPU needed
//                                  for the errors checked before a context opens)
//   dropin_core run A.bin B.bin W H WIN STEP THREADS
//       frames row-major float64; runs every window through "complex"
//       (algos::FFT) and "hip" (algos::HipFFT) with the reference's
//       ThreadPool "pool" strategy (main.cpp:265-276) and prints
//       "windows N mismatches M max_err E transform_err T auto_err A real_err R"
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <string>
#include <unordered_map>
#include <vector>

#include "algos/fft.h"
#include "core/grid.h"
#include "core/image.h"
#include "core/image_utils.h"
#include "core/vector.h"
#include "openpiv_amd_core.hpp"
#include "threadpool.hpp"

using namespace openpiv;

namespace {

struct record {
    double u = 0, v = 0;
    int px = 0, py = 0;
    bool two = false;
};

using correlator_t = std::function<core::gf_image(const core::gf_image&, const core::gf_image&)>;

// main.cpp:153-186 with the engine registered beside the reference's FFT
std::unordered_map<std::string, correlator_t> registry(core::size ia)
{
    return {
        { "complex",
          [ia](const core::gf_image& a, const core::gf_image& b) -> core::gf_image {
              static algos::FFT fft{ ia };
              return fft.cross_correlate(a, b);
          } },
        { "real",
          [ia](const core::gf_image& a, const core::gf_image& b) -> core::gf_image {
              static algos::FFT fft{ ia };
              return fft.cross_correlate_real(a, b);
          } },
        { "hip",
          [ia](const core::gf_image& a, const core::gf_image& b) -> core::gf_image {
              static algos::HipFFT fft{ ia };
              return fft.cross_correlate(a, b);
          } },
        { "hip_real",
          [ia](const core::gf_image& a, const core::gf_image& b) -> core::gf_image {
              static algos::HipFFT fft{ ia };
              return fft.cross_correlate_real(a, b);
          } },
    };
}

// main.cpp:189-237
void process_one(const std::vector<core::gf_image>& images, const correlator_t& correlator, const core::rect& ia,
                 record& out)
{
    const auto view_a{ core::extract(images[0], ia) };
    const auto view_b{ core::extract(images[1], ia) };
    const core::gf_image output{ correlator(view_a, view_b) };
    auto peaks = core::find_peaks(output, 2, 1);
    if (peaks.size() != 2) return;
    const auto bl = ia.bottomLeft();
    const auto midpoint = ia.midpoint();
    const auto loc = core::fit_simple_gaussian(peaks[0]);
    out.u = midpoint[0] - (bl[0] + loc[0]);
    out.v = midpoint[1] - (bl[1] + loc[1]);
    const auto pm = peaks[0].rect().midpoint();
    out.px = pm[0];
    out.py = pm[1];
    out.two = true;
}

std::vector<double> read_f64(const char* path, size_t n)
{
    std::vector<double> v(n);
    std::ifstream is(path, std::ios::binary);
    is.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(double)));
    if (!is) {
        std::fprintf(stderr, "short read %s\n", path);
        std::exit(2);
    }
    return v;
}

core::gf_image to_image(const std::vector<double>& v, uint32_t w, uint32_t h)
{
    core::gf_image im{ w, h };
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) im[{ x, y }] = v[(size_t)y * w + x];
    return im;
}

int fail(const char* what)
{
    std::printf("FAIL %s\n", what);
    return 1;
}

// errors raised before any context opens: the reference's message for a
// non-power-of-two size (fft.h:82-83)
int cmd_check()
{
    try {
        algos::HipFFT bad{ core::size{ 30, 32 } };
        return fail("non-power-of-two size accepted");
    } catch (const std::runtime_error& e) {
        if (std::string(e.what()) != "dimensions must be power of 2: [30,32]") return fail(e.what());
    }
    std::printf("check ok\n");
    return 0;
}

int cmd_run(int argc, char** argv)
{
    if (argc < 9) return 2;
    const uint32_t W = std::stoul(argv[4]), H = std::stoul(argv[5]), win = std::stoul(argv[6]);
    const uint32_t step = std::stoul(argv[7]);
    const int threads = std::stoi(argv[8]);
    std::vector<core::gf_image> images;
    images.push_back(to_image(read_f64(argv[2], (size_t)W * H), W, H));
    images.push_back(to_image(read_f64(argv[3], (size_t)W * H), W, H));
    const auto ia = core::size{ win, win };
    const auto grid = core::generate_cartesian_grid(images[0].size(), ia, { step, step });
    auto corr = registry(ia);

    auto run_all = [&](const correlator_t& c, std::vector<record>& out) {
        out.assign(grid.size(), record{});
        ThreadPool pool(threads);
        for (size_t i = 0; i < grid.size(); ++i)
            pool.enqueue([i, &grid, &images, &c, &out]() { process_one(images, c, grid[i], out[i]); });
    };
    std::vector<record> ref, hip, ref_real, hip_real;
    run_all(corr["complex"], ref);
    run_all(corr["hip"], hip);
    run_all(corr["real"], ref_real);
    run_all(corr["hip_real"], hip_real);

    size_t mism = 0;
    double err = 0.0, real_err = 0.0;
    for (size_t i = 0; i < grid.size(); ++i) {
        if (ref[i].two != hip[i].two || ref[i].px != hip[i].px || ref[i].py != hip[i].py) {
            ++mism;
            continue;
        }
        if (ref[i].two) err = std::max({ err, std::fabs(ref[i].u - hip[i].u), std::fabs(ref[i].v - hip[i].v) });
        if (ref_real[i].two && hip_real[i].two && ref_real[i].px == hip_real[i].px && ref_real[i].py == hip_real[i].py)
            real_err = std::max({ real_err, std::fabs(ref_real[i].u - hip_real[i].u),
                                  std::fabs(ref_real[i].v - hip_real[i].v) });
        else if (ref_real[i].two != hip_real[i].two || ref_real[i].px != hip_real[i].px)
            ++mism;
    }

    // transform (fft.h:86-119) and auto_correlate (fft.h:229-242) of the
    // first window, against the reference class, relative to the largest
    // magnitude
    const auto wa = core::extract(images[0], grid[0]);
    algos::FFT ref_fft{ ia };
    algos::HipFFT hip_fft{ ia };
    double terr = 0.0, tmax = 0.0, aerr = 0.0, amax = 0.0;
    {
        const core::cf_image rt = ref_fft.transform(wa);   // copies: both return per-thread storage
        const core::cf_image& ht = hip_fft.transform(wa);
        for (size_t i = 0; i < rt.pixel_count(); ++i) {
            tmax = std::max(tmax, std::hypot(rt[i].real, rt[i].imag));
            terr = std::max(terr, std::hypot(rt[i].real - ht[i].real, rt[i].imag - ht[i].imag));
        }
        const core::cf_image ra = ref_fft.auto_correlate(wa);
        const core::cf_image& ha = hip_fft.auto_correlate(wa);
        for (size_t i = 0; i < ra.pixel_count(); ++i) {
            amax = std::max(amax, std::fabs(ra[i].real));
            aerr = std::max(aerr, std::fabs(ra[i].real - ha[i].real) + std::fabs(ha[i].imag));
        }
    }
    // a size mismatch raises the reference's error text (fft.h:94-98)
    bool size_error = false;
    try {
        core::gf_image small{ win / 2, win / 2 };
        (void)hip_fft.transform(small);
    } catch (const std::runtime_error& e) {
        size_error = std::string(e.what()).rfind("image size is different from expected", 0) == 0;
    }
    std::printf("windows %zu mismatches %zu max_err %.3e transform_err %.3e auto_err %.3e real_err %.3e "
                "size_error %d\n",
                grid.size(), mism, err, tmax > 0 ? terr / tmax : terr, amax > 0 ? aerr / amax : aerr, real_err,
                size_error ? 1 : 0);
    return 0;
}

} // namespace

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    const std::string cmd = argv[1];
    if (cmd == "check") return cmd_check();
    if (cmd == "run") return cmd_run(argc, argv);
    return 2;
}
