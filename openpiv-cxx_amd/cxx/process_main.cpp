// process_main.cpp -- drop-in front-end of examples/process (main.cpp:34-317)
// over the HIP engine.
//
// Same options and defaults (main.cpp:53-63): -s/--size 32, -o/--overlap 0.5,
// -i/--input (or two positional files), -t/--thread-count, -e/--exec
// (pool | bulk-pool | async++ | single), -l/--limit-search,
// -f/--ffttype (complex | real | pocket | pocket_real), --loglevel.  Thread
// count and execution strategy select host-thread schedules in the reference;
// here every window of the pair goes through one GPU launch, so they are
// accepted and ignored.  --device picks the GPU.
//
// Frames are read by the engine's loaders (PNM P5 and baseline TIFF, the
// formats of the reference's loader registry, loaders/image_loader.cpp:23-57)
// and are handed to the device as 8/16-bit pixels, without the FP64
// widening of the reference.
//
// Output matches main.cpp:312-313: one "x, y, u, v, sn" line per window in
// grid order, at the stream's default precision.
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <string>
#include <vector>

#include "openpiv_amd.hpp"

namespace {

enum class Level { NONE, ERROR, WARN, INFO, DEBUG };

Level g_level = Level::INFO;

void log(Level l, const std::string& m)
{
    if (l <= g_level && l != Level::NONE) std::cerr << m << "\n";
}

struct Frame {
    uint32_t width = 0, height = 0;
    bool wide = false;              // 16-bit samples
    std::vector<uint8_t> px8;
    std::vector<uint16_t> px16;
    const void* data() const { return wide ? (const void*)px16.data() : (const void*)px8.data(); }
};

// the engine's loaders (opiv_image_read: PNM P5 and baseline TIFF, the
// formats of the reference's loader registry)
bool load_frame(const std::string& path, Frame& f, std::string& err)
{
    opiv_image_info info;
    if (opiv_image_probe(path.c_str(), &info) != OPIV_OK) {
        err = opiv_last_error();
        return false;
    }
    f.width = info.width;
    f.height = info.height;
    f.wide = info.dtype == OPIV_U16;
    const size_t n = (size_t)f.width * f.height;
    void* dst = nullptr;
    if (f.wide) {
        f.px16.resize(n);
        dst = f.px16.data();
    } else {
        f.px8.resize(n);
        dst = f.px8.data();
    }
    if (opiv_image_read(nullptr, path.c_str(), 0, dst, 0) != OPIV_OK) {
        err = opiv_last_error();
        return false;
    }
    return true;
}

void widen(Frame& f)
{
    if (f.wide) return;
    f.px16.assign(f.px8.begin(), f.px8.end());
    f.px8.clear();
    f.wide = true;
}

bool take_value(int& i, int argc, char** argv, const std::string& arg, const std::string& shortf,
                const std::string& longf, std::string& value)
{
    if (arg == shortf || arg == longf) {
        if (i + 1 >= argc) return false;
        value = argv[++i];
        return true;
    }
    if (arg.rfind(longf + "=", 0) == 0) {
        value = arg.substr(longf.size() + 1);
        return true;
    }
    if (shortf.size() == 2 && arg.size() > 2 && arg.rfind(shortf, 0) == 0 && arg[2] != '-') {
        value = arg.substr(2);
        return true;
    }
    return false;
}

} // namespace

int main(int argc, char* argv[])
{
    uint32_t size = 32;
    double overlap = 0.5;
    std::vector<std::string> inputs;
    std::string execution = "pool", fft_type = "complex", level = "INFO";
    bool limit_search = false;
    int device = 0;

    for (int i = 1; i < argc; ++i) {
        const std::string arg = argv[i];
        std::string v;
        try {
            if (arg == "-h" || arg == "--help") {
                std::cout << "usage: " << argv[0]
                          << " [-s size] [-o overlap] [-t threads] [-e exec] [-l] [-f complex|real|pocket|pocket_real]"
                             " [--loglevel LEVEL] [--device N] -i A B | A B\n";
                return 0;
            } else if (take_value(i, argc, argv, arg, "-s", "--size", v)) {
                size = (uint32_t)std::stoul(v);
            } else if (take_value(i, argc, argv, arg, "-o", "--overlap", v)) {
                overlap = std::stod(v);
            } else if (take_value(i, argc, argv, arg, "-t", "--thread-count", v)) {
                (void)std::stoul(v);
            } else if (take_value(i, argc, argv, arg, "-e", "--exec", v)) {
                execution = v;
            } else if (take_value(i, argc, argv, arg, "-f", "--ffttype", v)) {
                fft_type = v;
            } else if (take_value(i, argc, argv, arg, "", "--loglevel", v)) {
                level = v;
            } else if (take_value(i, argc, argv, arg, "", "--device", v)) {
                device = std::stoi(v);
            } else if (arg == "-l" || arg == "--limit-search") {
                limit_search = true;
            } else if (take_value(i, argc, argv, arg, "-i", "--input", v)) {
                inputs.push_back(v);
            } else if (!arg.empty() && arg[0] == '-') {
                std::cerr << "error parsing options: unknown option " << arg << "\n";
                return 1;
            } else {
                inputs.push_back(arg);
            }
        } catch (const std::exception& e) {
            std::cerr << "error parsing options: " << e.what() << "\n";
            return 1;
        }
    }
    if (level == "NONE") g_level = Level::NONE;
    else if (level == "ERROR") g_level = Level::ERROR;
    else if (level == "WARN") g_level = Level::WARN;
    else if (level == "DEBUG") g_level = Level::DEBUG;

    if (inputs.size() != 2) {
        log(Level::ERROR, "require two input images");
        return 1;
    }
    log(Level::INFO, "size: " + std::to_string(size) + " x " + std::to_string(size));
    log(Level::INFO, "overlap: " + std::to_string(overlap));
    log(Level::INFO, "input files: " + inputs[0] + ", " + inputs[1]);
    log(Level::INFO, "execution: " + execution + " (one GPU launch per frame pair)");

    Frame fa, fb;
    std::string err;
    try {
        // (a compressed frame may declare more pixels than memory holds)
        if (!load_frame(inputs[0], fa, err) || !load_frame(inputs[1], fb, err)) {
            log(Level::ERROR, err);
            return 1;
        }
    } catch (const std::exception& e) {
        log(Level::ERROR, std::string("failed to load the input images: ") + e.what());
        return 1;
    }
    if (fa.width != fb.width || fa.height != fb.height) {
        log(Level::ERROR, "image sizes don't match: [" + std::to_string(fa.width) + "," + std::to_string(fa.height) +
                              "], [" + std::to_string(fb.width) + "," + std::to_string(fb.height) + "]");
        return 1;
    }
    if (fa.wide != fb.wide) {
        widen(fa);
        widen(fb);
    }

    int variant;
    if (fft_type == "complex" || fft_type == "pocket" || fft_type == "pocket_real") {
        variant = OPIV_CIRCULAR_FFT_FP32;   // the same exact circular correlation
    } else if (fft_type == "real") {
        variant = OPIV_REAL_COMPAT;
    } else {
        log(Level::ERROR, "unknown fft type: " + fft_type);
        return 1;
    }

    try {
        openpiv_amd::Context ctx(device);
        opiv_plan_desc desc{};
        desc.img_w = fa.width;
        desc.img_h = fa.height;
        desc.pitch = 0;
        desc.dtype = fa.wide ? OPIV_U16 : OPIV_U8;
        desc.win = size;
        desc.overlap = overlap;
        desc.variant = variant;
        desc.limit_search = limit_search ? 1 : 0;
        openpiv_amd::Plan plan(ctx, desc);
        log(Level::INFO, "grid count: " + std::to_string(plan.window_count()));

        const auto t1 = std::chrono::high_resolution_clock::now();
        const auto results = plan.run(fa.data(), fb.data());
        const auto t2 = std::chrono::high_resolution_clock::now();
        const std::chrono::duration<double, std::micro> us = t2 - t1;
        log(Level::INFO, "processing time: " + std::to_string(us.count()) + "us, " +
                             std::to_string(us.count() / (double)results.size()) + "us per interrogation area");
        size_t failed = 0;
        for (const auto& r : results)
            if (r.flags & OPIV_FLAG_FEWER_THAN_2_PEAKS) ++failed;
        if (failed) log(Level::ERROR, "failed to find a peak for " + std::to_string(failed) + " interrogation areas");

        // main.cpp:312-313
        for (const auto& r : results) std::cout << r.x << ", " << r.y << ", " << r.u << ", " << r.v << ", " << r.sn << "\n";
    } catch (const std::exception& e) {   // openpiv_amd::error is a std::runtime_error
        log(Level::ERROR, e.what());
        return 1;
    }
    return 0;
}
