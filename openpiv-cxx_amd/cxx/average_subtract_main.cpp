// average_subtract_main.cpp -- drop-in front-end of examples/average_subtract
// (main.cpp:44-160) over the engine: the mean of every input but the first is
// subtracted from each input on the GPU (opiv_average_subtract), and each
// result is written next to its input as "<input>.avg_sub.pnm" in the format
// of the reference's PNM writer (pnm_image_loader.cpp:367-392: P5, a
// "# created by pnm_image_loader" comment, maxval 65535, big-endian samples).
// The files are byte-identical to the reference's.
//
//   opiv_average_subtract [-i|--input] A.pgm B.tif C.pgm ...   (two or more)
//
// Inputs are PNM P5 or baseline TIFF frames of one size (8- or 16-bit; 8-bit
// frames are widened to 16 bits when the set mixes them -- the same values).
#include <cstdint>
#include <cstdio>
#include <exception>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "openpiv_amd.h"

namespace {

void log_error(const std::string& m) { std::cerr << m << "\n"; }

// main.cpp:123 logs "failed to load image: " and the loader's message; the
// engine's decoder messages may carry that prefix already
std::string load_error()
{
    const std::string prefix = "failed to load image: ", m = opiv_last_error();
    return m.rfind(prefix, 0) == 0 ? m : prefix + m;
}

int run(int argc, char* argv[])
{
    std::vector<std::string> inputs;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-h" || a == "--help") {
            std::cout << "usage: " << argv[0] << " [-i|--input] FILE FILE [FILE ...]\n";
            return 0;
        }
        if (a == "-i" || a == "--input") continue;
        if (a.rfind("--input=", 0) == 0) {
            inputs.push_back(a.substr(8));
            continue;
        }
        inputs.push_back(a);
    }
    if (inputs.size() < 2) {
        log_error("require two or more input images");          // main.cpp:85-89
        return 1;
    }
    std::string joined;
    for (size_t i = 0; i < inputs.size(); ++i) joined += (i ? ", " : "") + inputs[i];
    std::cerr << "input files: " << joined << "\n";

    // load every frame in its own pixel type
    std::vector<opiv_image_info> info(inputs.size());
    for (size_t i = 0; i < inputs.size(); ++i) {
        if (opiv_image_probe(inputs[i].c_str(), &info[i]) != OPIV_OK) {
            log_error(load_error());
            return 1;
        }
        if (info[i].width != info[0].width || info[i].height != info[0].height) {
            log_error("failed to load image: " + inputs[i] + " differs in size from " + inputs[0]);
            return 1;
        }
    }
    bool wide = false;
    for (const auto& f : info) wide = wide || f.dtype == OPIV_U16;
    const uint32_t w = info[0].width, h = info[0].height;
    const size_t n = (size_t)w * h;
    std::vector<std::vector<uint16_t>> px16(inputs.size());
    std::vector<std::vector<uint8_t>> px8(inputs.size());
    std::vector<const void*> frames(inputs.size());
    for (size_t i = 0; i < inputs.size(); ++i) {
        if (info[i].dtype == OPIV_U16 || !wide) {
            void* dst;
            if (info[i].dtype == OPIV_U16) {
                px16[i].resize(n);
                dst = px16[i].data();
            } else {
                px8[i].resize(n);
                dst = px8[i].data();
            }
            if (opiv_image_read(nullptr, inputs[i].c_str(), 0, dst, 0) != OPIV_OK) {
                log_error(load_error());
                return 1;
            }
            frames[i] = dst;
        } else {
            std::vector<uint8_t> tmp(n);
            if (opiv_image_read(nullptr, inputs[i].c_str(), 0, tmp.data(), 0) != OPIV_OK) {
                log_error(load_error());
                return 1;
            }
            px16[i].assign(tmp.begin(), tmp.end());
            frames[i] = px16[i].data();
        }
    }

    opiv_ctx* ctx = nullptr;
    if (opiv_ctx_create(0, &ctx) != OPIV_OK) {
        log_error(std::string("failed to open the GPU: ") + opiv_last_error());
        return 1;
    }
    std::vector<std::vector<uint16_t>> out(inputs.size(), std::vector<uint16_t>(n));
    std::vector<uint16_t*> outp(inputs.size());
    for (size_t i = 0; i < inputs.size(); ++i) outp[i] = out[i].data();
    const int rc = opiv_average_subtract(ctx, frames.data(), inputs.size(), w, h, 0, wide ? OPIV_U16 : OPIV_U8,
                                         outp.data(), nullptr);
    opiv_ctx_destroy(ctx);
    if (rc != OPIV_OK) {
        log_error(opiv_last_error());
        return 1;
    }
    std::cerr << "found average; processed " << inputs.size() - 1 << " image files\n";

    // pnm_image_loader.cpp:376-391
    for (size_t i = 0; i < inputs.size(); ++i) {
        std::ofstream os(inputs[i] + ".avg_sub.pnm", std::ios::binary | std::ios::trunc);
        os << "P5\n"
           << "# created by pnm_image_loader\n"
           << w << " " << h << "\n"
           << "65535\n";
        std::vector<uint8_t> be(2 * n);
        for (size_t k = 0; k < n; ++k) {
            be[2 * k] = (uint8_t)(out[i][k] >> 8);
            be[2 * k + 1] = (uint8_t)(out[i][k] & 0xff);
        }
        os.write(reinterpret_cast<const char*>(be.data()), (std::streamsize)be.size());
    }
    return 0;
}

} // namespace

int main(int argc, char* argv[])
{
    // (a compressed frame may declare more pixels than memory holds)
    try {
        return run(argc, argv);
    } catch (const std::exception& e) {
        log_error(std::string("failed: ") + e.what());
        return 1;
    }
}
