// ref_harness.cpp -- a small driver around the reference's own headers
// (ErichZimmer/openpiv-cxx, compiled from /root/reference by oracle/Makefile
// into oracle/_ref/).  TEST INFRASTRUCTURE ONLY: it produces golden vectors
// and the "reference" CPU baseline; the product never links or runs it.
//
// The per-window processor below calls the reference's functions in the
// order of examples/process/main.cpp:189-237 (extract -> correlator ->
// find_peaks(.., 2, 1) -> fit_simple_gaussian -> vector), and the timed
// mode drives them with the reference's ThreadPool exactly as the "pool"
// strategy does (main.cpp:265-276).  cxxopts is absent from the image, so
// the argument parsing is this file's own.
//
// usage:
//   ref_piv process A.bin B.bin W H WIN OFF_X OFF_Y VARIANT LIMIT THREADS OUT.bin
//       frames are row-major float64; VARIANT complex|real; writes per
//       window: x y u v sn peak0 peak1 (7 f64) px py flags (3 i32) + pad i32
//   ref_piv planes A.bin B.bin W H WIN OFF_X OFF_Y VARIANT OUT.bin IDX...
//       raw f64 correlation planes of the listed windows
//   ref_piv transform WIN DIR IN.bin OUT.bin   (complex f64 pairs)
//   ref_piv grid IMG_W IMG_H WIN_W WIN_H OFF_X OFF_Y     (prints x,y per line)
//   ref_piv gridf IMG_W IMG_H WIN_W WIN_H FRACTION
//   ref_piv peaks W H NUM RADIUS IN.bin          (prints x y value per peak)
//   ref_piv time A.bin B.bin W H WIN OFF_X OFF_Y THREADS REPS (prints us per rep)
//   ref_piv load FILE OUT.bin (the reference's loaders: "w h", FP64 pixels)
//   ref_piv avgsub OUT_PREFIX F0.pgm F1.pgm ...
//       examples/average_subtract/main.cpp:98-159 on PNM frames loaded by the
//       reference's loader: writes OUT_PREFIX<i>.pnm with the reference's PNM
//       writer and OUT_PREFIXmean.bin (the average, f64)

#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "algos/fft.h"
#include "algos/stats.h"
#include "core/grid.h"
#include "core/image.h"
#include "core/image_utils.h"
#include "core/vector.h"
#include "loaders/image_loader.h"
#include "threadpool.hpp"

using namespace openpiv;

namespace {

std::vector<double> read_f64(const char* path, size_t n)
{
    std::vector<double> v(n);
    std::ifstream is(path, std::ios::binary);
    is.read(reinterpret_cast<char*>(v.data()), n * sizeof(double));
    if (!is) { std::fprintf(stderr, "short read %s\n", path); std::exit(2); }
    return v;
}

core::gf_image to_image(const std::vector<double>& v, uint32_t w, uint32_t h)
{
    core::gf_image im{ w, h };
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x)
            im[{ x, y }] = v[(size_t)y * w + x];
    return im;
}

struct record {
    double x, y, u, v, sn, peak0, peak1;
    int32_t px, py;
    uint32_t flags;
    int32_t pad;
};

using correlator_t = std::function<core::gf_image(const core::gf_image&, const core::gf_image&)>;

correlator_t make_correlator(const std::string& variant, core::size ia)
{
    if (variant == "real")
        return [ia](const core::gf_image& a, const core::gf_image& b) -> core::gf_image {
            static algos::FFT fft{ ia };
            return fft.cross_correlate_real(a, b);
        };
    return [ia](const core::gf_image& a, const core::gf_image& b) -> core::gf_image {
        static algos::FFT fft{ ia };
        return fft.cross_correlate(a, b);
    };
}

// examples/process/main.cpp:189-237, recording the peak values as well
void process_one(const std::vector<core::gf_image>& images, const correlator_t& correlator,
                 bool limit_search, const core::rect& ia, record& out)
{
    out = record{};
    const auto view_a{ core::extract(images[0], ia) };
    const auto view_b{ core::extract(images[1], ia) };
    const core::gf_image output{ correlator(view_a, view_b) };

    constexpr uint16_t num_peaks = 2;
    constexpr uint16_t radius = 1;
    core::peaks_t<core::g_f> peaks;
    if (limit_search) {
        auto centre = core::create_image_view(output, output.rect().dilate(0.5));
        peaks = core::find_peaks(centre, num_peaks, radius);
    } else {
        peaks = core::find_peaks(output, num_peaks, radius);
    }
    if (peaks.size() != num_peaks) {
        out.flags = 1;
        return;
    }
    auto bl = ia.bottomLeft();
    auto midpoint = ia.midpoint();
    auto peak = peaks[0];
    auto peak_location = core::fit_simple_gaussian(peak);
    core::point2<double> xy = midpoint;
    core::vector2<double> vxy = { midpoint[0] - (bl[0] + peak_location[0]), midpoint[1] - (bl[1] + peak_location[1]) };
    xy[1] = images[0].height() - xy[1];
    double sn = 0.0;
    if (peaks[1][{ 1, 1 }] > 0)
        sn = peaks[0][{ 1, 1 }] / peaks[1][{ 1, 1 }];
    out.x = xy[0];
    out.y = xy[1];
    out.u = vxy[0];
    out.v = vxy[1];
    out.sn = sn;
    out.peak0 = peaks[0][{ 1, 1 }];
    out.peak1 = peaks[1][{ 1, 1 }];
    const auto pm = peaks[0].rect().midpoint();
    out.px = pm[0];
    out.py = pm[1];
    if (peaks[0][{ 1, 1 }] == peaks[1][{ 1, 1 }])
        out.flags |= 8;
}

int cmd_process(int argc, char** argv)
{
    if (argc < 13) return 1;
    const uint32_t W = std::stoul(argv[4]), H = std::stoul(argv[5]), win = std::stoul(argv[6]);
    const uint32_t ox = std::stoul(argv[7]), oy = std::stoul(argv[8]);
    const std::string variant = argv[9];
    const bool limit = std::stoi(argv[10]) != 0;
    const int threads = std::stoi(argv[11]);
    std::vector<core::gf_image> images;
    images.push_back(to_image(read_f64(argv[2], (size_t)W * H), W, H));
    images.push_back(to_image(read_f64(argv[3], (size_t)W * H), W, H));
    const auto ia = core::size{ win, win };
    auto grid = core::generate_cartesian_grid(images[0].size(), ia, { ox, oy });
    std::vector<record> found(grid.size());
    auto correlator = make_correlator(variant, ia);
    if (threads <= 1) {
        for (size_t i = 0; i < grid.size(); ++i)
            process_one(images, correlator, limit, grid[i], found[i]);
    } else {
        ThreadPool pool(threads);
        for (size_t i = 0; i < grid.size(); ++i)
            pool.enqueue([i, &grid, &images, &correlator, limit, &found]() {
                process_one(images, correlator, limit, grid[i], found[i]);
            });
    }
    std::ofstream os(argv[12], std::ios::binary);
    os.write(reinterpret_cast<const char*>(found.data()), found.size() * sizeof(record));
    return 0;
}

int cmd_planes(int argc, char** argv)
{
    if (argc < 12) return 1;
    const uint32_t W = std::stoul(argv[4]), H = std::stoul(argv[5]), win = std::stoul(argv[6]);
    const uint32_t ox = std::stoul(argv[7]), oy = std::stoul(argv[8]);
    const std::string variant = argv[9];
    std::vector<core::gf_image> images;
    images.push_back(to_image(read_f64(argv[2], (size_t)W * H), W, H));
    images.push_back(to_image(read_f64(argv[3], (size_t)W * H), W, H));
    const auto ia = core::size{ win, win };
    auto grid = core::generate_cartesian_grid(images[0].size(), ia, { ox, oy });
    auto correlator = make_correlator(variant, ia);
    std::ofstream os(argv[10], std::ios::binary);
    for (int k = 11; k < argc; ++k) {
        const size_t i = std::stoul(argv[k]);
        const core::gf_image out{ correlator(core::extract(images[0], grid[i]), core::extract(images[1], grid[i])) };
        for (uint32_t y = 0; y < win; ++y)
            for (uint32_t x = 0; x < win; ++x) {
                const double v = out[{ x, y }];
                os.write(reinterpret_cast<const char*>(&v), sizeof(v));
            }
    }
    return 0;
}

int cmd_transform(int argc, char** argv)
{
    if (argc < 6) return 1;
    const uint32_t win = std::stoul(argv[2]);
    const int dir = std::stoi(argv[3]);
    auto raw = read_f64(argv[4], (size_t)win * win * 2);
    core::cf_image in{ win, win };
    for (size_t i = 0; i < (size_t)win * win; ++i)
        in[i] = core::c_f{ raw[2 * i], raw[2 * i + 1] };
    algos::FFT fft{ core::size{ win, win } };
    const core::cf_image& out = fft.transform(in, dir == 0 ? algos::direction::FORWARD : algos::direction::REVERSE);
    std::ofstream os(argv[5], std::ios::binary);
    for (size_t i = 0; i < (size_t)win * win; ++i) {
        const double re = out[i].real, im = out[i].imag;
        os.write(reinterpret_cast<const char*>(&re), 8);
        os.write(reinterpret_cast<const char*>(&im), 8);
    }
    return 0;
}

int cmd_grid(int argc, char** argv, bool fraction)
{
    if (argc < (fraction ? 7 : 8)) return 1;
    const core::size img{ (uint32_t)std::stoul(argv[2]), (uint32_t)std::stoul(argv[3]) };
    const core::size ia{ (uint32_t)std::stoul(argv[4]), (uint32_t)std::stoul(argv[5]) };
    std::vector<core::rect> grid;
    try {
        if (fraction)
            grid = core::generate_cartesian_grid(img, ia, std::stod(argv[6]));
        else
            grid = core::generate_cartesian_grid(img, ia, std::array<uint32_t, 2>{ (uint32_t)std::stoul(argv[6]), (uint32_t)std::stoul(argv[7]) });
    } catch (const std::exception& e) {
        std::printf("error %s\n", e.what());
        return 0;
    }
    for (const auto& r : grid)
        std::printf("%d %d\n", r.bottomLeft()[0], r.bottomLeft()[1]);
    return 0;
}

int cmd_peaks(int argc, char** argv)
{
    if (argc < 7) return 1;
    const uint32_t W = std::stoul(argv[2]), H = std::stoul(argv[3]);
    const uint16_t num = (uint16_t)std::stoul(argv[4]);
    const uint32_t radius = std::stoul(argv[5]);
    const auto im = to_image(read_f64(argv[6], (size_t)W * H), W, H);
    auto peaks = core::find_peaks(im, num, radius);
    for (const auto& p : peaks) {
        const auto m = p.rect().midpoint();
        std::printf("%d %d %.17g\n", m[0], m[1], (double)p[{ radius, radius }]);
    }
    return 0;
}

int cmd_time(int argc, char** argv)
{
    if (argc < 11) return 1;
    const uint32_t W = std::stoul(argv[4]), H = std::stoul(argv[5]), win = std::stoul(argv[6]);
    const uint32_t ox = std::stoul(argv[7]), oy = std::stoul(argv[8]);
    const int threads = std::stoi(argv[9]);
    const int reps = std::stoi(argv[10]);
    std::vector<core::gf_image> images;
    images.push_back(to_image(read_f64(argv[2], (size_t)W * H), W, H));
    images.push_back(to_image(read_f64(argv[3], (size_t)W * H), W, H));
    const auto ia = core::size{ win, win };
    auto grid = core::generate_cartesian_grid(images[0].size(), ia, { ox, oy });
    std::vector<record> found(grid.size());
    auto correlator = make_correlator("complex", ia);
    for (int r = 0; r < reps; ++r) {
        // main.cpp:239-309: only the processing loop is timed
        const auto t1 = std::chrono::high_resolution_clock::now();
        if (threads <= 1) {
            for (size_t i = 0; i < grid.size(); ++i)
                process_one(images, correlator, false, grid[i], found[i]);
        } else {
            ThreadPool pool(threads);
            for (size_t i = 0; i < grid.size(); ++i)
                pool.enqueue([i, &grid, &images, &correlator, &found]() {
                    process_one(images, correlator, false, grid[i], found[i]);
                });
        }
        const auto t2 = std::chrono::high_resolution_clock::now();
        const std::chrono::duration<double, std::micro> us = t2 - t1;
        std::printf("%zu %.3f\n", grid.size(), us.count());
        std::fflush(stdout);
    }
    return 0;
}

// load_from_file of examples/average_subtract/main.cpp:21-42 (the error
// path returns an empty image, as there)
core::gf_image load_pnm(const std::string& filename)
{
    std::ifstream is(filename, std::ios::binary);
    core::gf_image image;
    if (!is.is_open()) return image;
    auto loader{ core::image_loader_registry::find(is) };
    if (!loader) return image;
    loader->load(is, image);
    return image;
}

// examples/average_subtract/main.cpp:98-159: the average of every input but
// the first (the first only sets the size), subtracted from every input, each
// result written by the PNM writer
int cmd_avgsub(int argc, char** argv)
{
    if (argc < 5) return 1;
    const std::string prefix = argv[2];
    std::vector<std::string> inputs(argv + 3, argv + argc);
    core::gf_image avg;
    {
        auto image = load_pnm(inputs[0]);
        avg.resize(image.size());
    }
    size_t processed = 0;
    for (size_t i = 1; i < inputs.size(); ++i) {
        auto image = load_pnm(inputs[i]);
        avg = avg + image;
        ++processed;
    }
    if (processed == 0) return 1;
    avg = avg / core::g_f{ processed };
    {
        std::ofstream os(prefix + "mean.bin", std::ios::binary);
        for (uint32_t h = 0; h < avg.height(); ++h)
            os.write(reinterpret_cast<const char*>(avg.line(h)), (std::streamsize)(avg.width() * sizeof(double)));
    }
    std::shared_ptr<core::image_loader> writer{ core::image_loader_registry::find("image/x-portable-anymap") };
    if (!writer) return 1;
    for (size_t i = 0; i < inputs.size(); ++i) {
        auto image = load_pnm(inputs[i]);
        image = image - avg;
        std::fstream os(prefix + std::to_string(i) + ".pnm",
                        std::ios_base::trunc | std::ios_base::out | std::ios_base::binary);
        writer->save(os, image);
    }
    return 0;
}

// the reference's loader registry (image_loader.cpp:23-57) reading a file
// into a gf_image, as examples/process/main.cpp:106-135 loads its frames:
// prints "w h" and writes the pixels as FP64 to OUT
int cmd_load(int argc, char** argv)
{
    if (argc < 4) return 1;
    std::ifstream is(argv[2], std::ios::binary);
    if (!is.is_open()) return 2;
    auto loader{ core::image_loader_registry::find(is) };
    if (!loader) return 3;
    core::gf_image image;
    if (!loader->load(is, image)) return 4;
    std::ofstream os(argv[3], std::ios::binary);
    for (uint32_t h = 0; h < image.height(); ++h)
        os.write(reinterpret_cast<const char*>(image.line(h)), (std::streamsize)(image.width() * sizeof(double)));
    std::printf("%u %u\n", image.width(), image.height());
    return 0;
}

} // namespace

int main(int argc, char** argv)
{
    if (argc < 2) return 1;
    const std::string cmd = argv[1];
    if (cmd == "process") return cmd_process(argc, argv);
    if (cmd == "planes") return cmd_planes(argc, argv);
    if (cmd == "transform") return cmd_transform(argc, argv);
    if (cmd == "grid") return cmd_grid(argc, argv, false);
    if (cmd == "gridf") return cmd_grid(argc, argv, true);
    if (cmd == "peaks") return cmd_peaks(argc, argv);
    if (cmd == "time") return cmd_time(argc, argv);
    if (cmd == "avgsub") return cmd_avgsub(argc, argv);
    if (cmd == "load") return cmd_load(argc, argv);
    return 1;
}
