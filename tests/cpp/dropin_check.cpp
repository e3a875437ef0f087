// Drives openpiv_amd.hpp the way examples/process drives algos::FFT: an
// image type with the reference's accessors (width(), height(),
// operator[](size_t), ImageT(w, h)) goes through HipFFT::cross_correlate.
//   dropin_check correlate WIN VARIANT A.bin B.bin OUT.bin   (FP64 windows)
//   dropin_check pow2 W H        -> prints the exception text
//   dropin_check nodevice        -> prints the error code of Context()
//   dropin_check transform WIN forward|reverse IN.bin OUT.bin
//       HipFFT::transform of one window of complex FP64 pixels
//   dropin_check adjoint W H WIN A.bin B.bin COT.bin OUT.bin
//       f32 frames, 50% overlap: Plan::vjp (grad_a, grad_b as f32) then
//       Plan::variance (sigma 2, f64) written one after the other
//   dropin_check batch W H A.bin B.bin OUT.bin
//       f32 frames copied to device memory, one pair x passes 32/16 at 50%
//       through Batch; the results (72 B each) in batch order
#include <hip/hip_runtime_api.h>

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

struct cimage {   // complex pixels with real/imag members, like core::complex<double>
    struct px_t {
        double real, imag;
    };
    uint32_t w, h;
    std::vector<px_t> px;
    cimage(uint32_t w_, uint32_t h_) : w(w_), h(h_), px((size_t)w_ * h_) {}
    uint32_t width() const { return w; }
    uint32_t height() const { return h; }
    const px_t& operator[](size_t i) const { return px[i]; }
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
        if (cmd == "transform" && argc >= 6) {
            const uint32_t win = (uint32_t)std::stoul(argv[2]);
            const bool fwd = std::string(argv[3]) == "forward";
            cimage in(win, win);
            std::ifstream(argv[4], std::ios::binary)
                .read(reinterpret_cast<char*>(in.px.data()), (std::streamsize)(in.px.size() * 16));
            openpiv_amd::Context ctx(0);
            openpiv_amd::HipFFT fft(ctx, win, win);
            const auto out = fft.transform(in, fwd ? OPIV_FORWARD : OPIV_REVERSE);
            std::ofstream os(argv[5], std::ios::binary);
            os.write(reinterpret_cast<const char*>(out.data()), (std::streamsize)(out.size() * 16));
            return 0;
        }
        if (cmd == "adjoint" && argc >= 9) {
            const uint32_t w = (uint32_t)std::stoul(argv[2]), h = (uint32_t)std::stoul(argv[3]);
            const uint32_t win = (uint32_t)std::stoul(argv[4]);
            std::vector<float> a((size_t)w * h), b((size_t)w * h);
            std::ifstream(argv[5], std::ios::binary).read(reinterpret_cast<char*>(a.data()), (std::streamsize)(a.size() * 4));
            std::ifstream(argv[6], std::ios::binary).read(reinterpret_cast<char*>(b.data()), (std::streamsize)(b.size() * 4));
            openpiv_amd::Context ctx(0);
            opiv_plan_desc d{};
            d.img_w = w;
            d.img_h = h;
            d.dtype = OPIV_F32;
            d.win = win;
            d.overlap = 0.5;
            d.variant = OPIV_CIRCULAR_FFT_FP32;
            openpiv_amd::Plan plan(ctx, d);
            std::vector<double> cot(2 * plan.window_count());
            std::ifstream(argv[7], std::ios::binary).read(reinterpret_cast<char*>(cot.data()), (std::streamsize)(cot.size() * 8));
            const auto fwd = plan.run(a.data(), b.data());
            std::vector<float> ga(a.size()), gb(b.size());
            plan.vjp(a.data(), b.data(), fwd.data(), cot.data(), ga.data(), gb.data());
            const auto var = plan.variance(a.data(), b.data(), fwd.data(), 2.0);
            std::ofstream os(argv[8], std::ios::binary);
            os.write(reinterpret_cast<const char*>(ga.data()), (std::streamsize)(ga.size() * 4));
            os.write(reinterpret_cast<const char*>(gb.data()), (std::streamsize)(gb.size() * 4));
            os.write(reinterpret_cast<const char*>(var.data()), (std::streamsize)(var.size() * 8));
            return 0;
        }
        if (cmd == "batch" && argc >= 7) {
            const uint32_t w = (uint32_t)std::stoul(argv[2]), h = (uint32_t)std::stoul(argv[3]);
            const size_t bytes = (size_t)w * h * 4;
            std::vector<float> a((size_t)w * h), b((size_t)w * h);
            std::ifstream(argv[4], std::ios::binary).read(reinterpret_cast<char*>(a.data()), (std::streamsize)bytes);
            std::ifstream(argv[5], std::ios::binary).read(reinterpret_cast<char*>(b.data()), (std::streamsize)bytes);
            openpiv_amd::Context ctx(0);
            void *da = nullptr, *db = nullptr, *dout = nullptr;
            if (hipMalloc(&da, bytes) != hipSuccess || hipMalloc(&db, bytes) != hipSuccess) return 3;
            if (hipMemcpy(da, a.data(), bytes, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(db, b.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
                return 3;
            opiv_batch_desc d{};
            d.dtype = OPIV_F32;
            d.npasses = 2;
            d.win[0] = 32;
            d.win[1] = 16;
            d.overlap[0] = d.overlap[1] = 0.5;
            std::vector<opiv_pair_desc> pairs(1);
            pairs[0].a = da;
            pairs[0].b = db;
            pairs[0].img_w = w;
            pairs[0].img_h = h;
            pairs[0].pitch = 0;
            openpiv_amd::Batch batch(ctx, d, pairs);
            const size_t n = batch.window_count();
            if (hipMalloc(&dout, n * sizeof(opiv_window_result)) != hipSuccess) return 3;
            batch.run(static_cast<opiv_window_result*>(dout), n);
            ctx.synchronize();
            std::vector<opiv_window_result> res(n);
            if (hipMemcpy(res.data(), dout, n * sizeof(opiv_window_result), hipMemcpyDeviceToHost) != hipSuccess) return 3;
            std::ofstream os(argv[6], std::ios::binary);
            os.write(reinterpret_cast<const char*>(res.data()), (std::streamsize)(n * sizeof(opiv_window_result)));
            (void)hipFree(da);
            (void)hipFree(db);
            (void)hipFree(dout);
            return 0;
        }
    } catch (const openpiv_amd::error& e) {
        std::cout << "error " << e.code() << " " << e.what() << "\n";
        return 0;
    }
    return 2;
}
