// Drives openpiv_amd.hpp the way examples/process drives algos::FFT: an
// image type with the reference's accessors (width(), height(),
// operator[](size_t), ImageT(w, h)) goes through HipFFT::cross_correlate.
//   dropin_check correlate WIN VARIANT A.bin B.bin OUT.bin   (FP64 windows)
//   dropin_check pow2 W H        -> prints the exception text
//   dropin_check nodevice        -> prints the error code of Context()
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <vector>

#include "openpiv_amd.hpp"

struct gimage {   // stands in for core::gf_image's interface
    uint32_t w, h;
    std::vector<double> px;
    gimage(uint32_t w_, uint32_t h_) : w(w_), h(h_), px((size_t)w_ * h_) {}
    uint32_t width() const { return w; }
    uint32_t height() const { return h; }
    double& operator[](size_t i) { return px[i]; }
    const double& operator[](size_t i) const { return px[i]; }
};

static gimage read_window(const char* path, uint32_t win)
{
    gimage im(win, win);
    std::ifstream is(path, std::ios::binary);
    is.read(reinterpret_cast<char*>(im.px.data()), (std::streamsize)(im.px.size() * sizeof(double)));
    return im;
}

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    const std::string cmd = argv[1];
    try {
        if (cmd == "nodevice") {
            openpiv_amd::Context ctx(0);
            std::cout << "ok\n";
            return 0;
        }
        if (cmd == "pow2" && argc >= 4) {
            try {
                openpiv_amd::Context ctx(0);
                openpiv_amd::HipFFT fft(ctx, (uint32_t)std::stoul(argv[2]), (uint32_t)std::stoul(argv[3]));
            } catch (const std::runtime_error& e) {
                std::cout << e.what() << "\n";
            }
            return 0;
        }
        if (cmd == "correlate" && argc >= 7) {
            const uint32_t win = (uint32_t)std::stoul(argv[2]);
            const std::string variant = argv[3];
            openpiv_amd::Context ctx(0);
            openpiv_amd::HipFFT fft(ctx, win, win);
            const gimage a = read_window(argv[4], win), b = read_window(argv[5], win);
            const gimage c = variant == "real" ? fft.cross_correlate_real(a, b)
                           : variant == "auto" ? fft.auto_correlate(a)
                                               : fft.cross_correlate(a, b);
            std::ofstream os(argv[6], std::ios::binary);
            os.write(reinterpret_cast<const char*>(c.px.data()), (std::streamsize)(c.px.size() * sizeof(double)));
            return 0;
        }
    } catch (const openpiv_amd::error& e) {
        std::cout << "error " << e.code() << " " << e.what() << "\n";
        return 0;
    }
    return 2;
}
