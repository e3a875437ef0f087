// openpiv_amd.hpp -- header-only C++17 layer over the C ABI of openpiv_amd.h.
//
// Mirrors the interface examples/process uses from openpiv-cxx so the engine
// drops into that driver:
//   * openpiv_amd::HipFFT          ~ algos::FFT (algos/fft.h:28-295): built
//     from a size (any type with width() / height(), as core::size) on the
//     process-wide context, or on an explicit Context; the constructor checks
//     power-of-two sizes with the same message (fft.h:82-83);
//     cross_correlate / cross_correlate_real / auto_correlate take and return
//     images of the caller's type (fft.h:188-242).  openpiv_amd_core.hpp
//     wraps it as openpiv::algos::HipFFT on openpiv-cxx's own core::image /
//     core::cf_image types, for the correlator registry of main.cpp:153-186.
//   * openpiv_amd::Plan            ~ the grid (main.cpp:137-138) plus the whole
//     per-window loop (main.cpp:189-302) as one batched call per frame pair.
// Errors surface as std::runtime_error subclasses carrying the C status, as
// the reference throws std::runtime_error (exception_builder.h:12-31).
//
// ImageT requirements (satisfied by openpiv::core::image<g<double>>):
// width(), height(), operator[](size_t) readable as double and assignable
// from double, and construction as ImageT(width, height).
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <complex>
#include <mutex>
#include <type_traits>
#include <stdexcept>
#include <string>
#include <vector>

#include "openpiv_amd.h"

namespace openpiv_amd {

class error : public std::runtime_error {
public:
    error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const noexcept { return code_; }

private:
    int code_;
};

inline void check(int rc)
{
    if (rc != OPIV_OK) throw error(rc, opiv_last_error());
}

/// One device and one HIP stream (opiv_ctx).
class Context {
public:
    explicit Context(int device = 0) { check(opiv_ctx_create(device, &h_)); }
    ~Context() { opiv_ctx_destroy(h_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    opiv_ctx* get() const { return h_; }
    void synchronize() const { check(opiv_ctx_synchronize(h_)); }
    /// serialises host threads that share this context
    std::mutex& mutex() const { return m_; }

private:
    opiv_ctx* h_ = nullptr;
    mutable std::mutex m_;
};

/// The process-wide context: device $OPIV_DEVICE (else 0), created on first
/// use and kept until the process ends.  Correlators made without an
/// explicit context (HipFFT(size), as `static algos::FFT fft{ia}` is made at
/// examples/process/main.cpp:158) share it.
inline Context& default_context()
{
    static Context ctx([] {
        const char* e = std::getenv("OPIV_DEVICE");
        return e && *e ? std::atoi(e) : 0;
    }());
    return ctx;
}

/// Square-window FFT correlator with the call surface of algos::FFT.
/// Thread-safe like the reference class: calls from several threads are
/// serialised on the context.
class HipFFT {
public:
    /// algos::FFT(const core::size&) (fft.h:76-84) for any size type with
    /// width() and height(), on the process-wide context
    /// (the size is checked before the context is opened, so a bad size
    /// reports the reference's error on any machine)
    template <class SizeT, class = std::void_t<decltype(std::declval<const SizeT&>().width()),
                                               decltype(std::declval<const SizeT&>().height())>>
    explicit HipFFT(const SizeT& size)
        : HipFFT((check_size(static_cast<uint32_t>(size.width()), static_cast<uint32_t>(size.height())),
                  default_context()),
                 static_cast<uint32_t>(size.width()), static_cast<uint32_t>(size.height()))
    {}

    HipFFT(Context& ctx, uint32_t width, uint32_t height) : ctx_(ctx), w_(width), h_(height)
    {
        check_size(w_, h_);
    }

    /// fft.h:82-83, and the engine's square windows
    static void check_size(uint32_t w, uint32_t h)
    {
        const auto pow2 = [](uint32_t v) { return v && (v & (v - 1)) == 0; };
        if (!(pow2(w) && pow2(h)))
            throw error(OPIV_E_NOT_POW2, "dimensions must be power of 2: [" + std::to_string(w) + "," +
                                             std::to_string(h) + "]");
        if (w != h)
            throw error(OPIV_E_UNSUPPORTED, "only square windows are supported: [" + std::to_string(w) + "," +
                                                std::to_string(h) + "]");
    }

    /// fft.h:188-206: real(IFFT(FFT(b) conj(FFT(a)))), quadrant-swapped
    template <class ImageT>
    ImageT cross_correlate(const ImageT& a, const ImageT& b) const
    {
        return correlate(a, &b, OPIV_CIRCULAR_FFT_FP32);
    }

    /// fft.h:208-227 (with the reference's two-real-image unpack)
    template <class ImageT>
    ImageT cross_correlate_real(const ImageT& a, const ImageT& b) const
    {
        return correlate(a, &b, OPIV_REAL_COMPAT);
    }

    /// fft.h:229-242 (real part)
    template <class ImageT>
    ImageT auto_correlate(const ImageT& a) const
    {
        return correlate(a, static_cast<const ImageT*>(nullptr), OPIV_CIRCULAR_FFT_FP32);
    }

    /// fft.h:86-119: un-normalised 2-D DFT (FP32 arithmetic); real or complex
    /// pixels in, complex values out in row-major order
    template <class ImageT>
    std::vector<std::complex<double>> transform(const ImageT& in, opiv_direction d = OPIV_FORWARD) const
    {
        if (in.width() != w_ || in.height() != h_)
            throw error(OPIV_E_SIZE, "image size is different from expected: [" + std::to_string(in.width()) + "," +
                                         std::to_string(in.height()) + "], [" + std::to_string(w_) + "," +
                                         std::to_string(h_) + "]");
        const size_t n = static_cast<size_t>(w_) * h_;
        std::vector<std::complex<double>> z(n), out(n);
        for (size_t i = 0; i < n; ++i) z[i] = to_complex(in[i]);
        std::lock_guard<std::mutex> lock(ctx_.mutex());
        check(opiv_fft_transform_windows(ctx_.get(), w_, 1, d, reinterpret_cast<const double*>(z.data()),
                                         reinterpret_cast<double*>(out.data())));
        return out;
    }

    /// Batched form: n stacked w*h windows (host or device memory).
    void cross_correlate(const double* a, const double* b, size_t n, double* planes,
                         opiv_variant variant = OPIV_CIRCULAR_FFT_FP32) const
    {
        std::lock_guard<std::mutex> lock(ctx_.mutex());
        check(opiv_correlate_windows(ctx_.get(), w_, n, variant, a, b, planes));
    }

    /// transform of n stacked windows of interleaved complex doubles
    /// (host or device memory; in == out allowed)
    void transform_raw(const double* in, double* out, opiv_direction d, size_t n = 1) const
    {
        std::lock_guard<std::mutex> lock(ctx_.mutex());
        check(opiv_fft_transform_windows(ctx_.get(), w_, n, d, in, out));
    }

    /// auto-correlation planes of n stacked real windows
    void auto_correlate_raw(const double* a, double* planes, size_t n = 1) const
    {
        std::lock_guard<std::mutex> lock(ctx_.mutex());
        check(opiv_auto_correlate_windows(ctx_.get(), w_, n, a, planes));
    }

    uint32_t width() const { return w_; }
    uint32_t height() const { return h_; }

private:
    // a real pixel (anything convertible to double, e.g. core::g<T>), a
    // std::complex, or a pixel with real/imag members (core::complex<T>,
    // pixel_types.h:200-201)
    template <class T, class = void>
    struct has_real_member : std::false_type {};
    template <class T>
    struct has_real_member<T, std::void_t<decltype(std::declval<const T&>().real + std::declval<const T&>().imag)>>
        : std::true_type {};

    template <class T>
    static std::complex<double> to_complex(const T& v)
    {
        if constexpr (std::is_convertible_v<T, double>) return { static_cast<double>(v), 0.0 };
        else if constexpr (has_real_member<T>::value) return { static_cast<double>(v.real), static_cast<double>(v.imag) };
        else return { static_cast<double>(v.real()), static_cast<double>(v.imag()) };
    }

    template <class ImageT>
    ImageT correlate(const ImageT& a, const ImageT* b, opiv_variant variant) const
    {
        if (a.width() != w_ || a.height() != h_ || (b && (b->width() != w_ || b->height() != h_)))
            throw error(OPIV_E_SIZE, "image size is different from expected: [" + std::to_string(a.width()) + "," +
                                         std::to_string(a.height()) + "], [" + std::to_string(w_) + "," +
                                         std::to_string(h_) + "]");
        const size_t n = static_cast<size_t>(w_) * h_;
        std::vector<double> da(n), db(b ? n : 0), out(n);
        for (size_t i = 0; i < n; ++i) da[i] = static_cast<double>(a[i]);
        if (b)
            for (size_t i = 0; i < n; ++i) db[i] = static_cast<double>((*b)[i]);
        {
            std::lock_guard<std::mutex> lock(ctx_.mutex());
            if (b)
                check(opiv_correlate_windows(ctx_.get(), w_, 1, variant, da.data(), db.data(), out.data()));
            else
                check(opiv_auto_correlate_windows(ctx_.get(), w_, 1, da.data(), out.data()));
        }
        ImageT result(w_, h_);
        for (size_t i = 0; i < n; ++i) result[i] = out[i];
        return result;
    }

    Context& ctx_;
    uint32_t w_, h_;
};

/// Frame size + window + grid; run() is the batched processing loop.
class Plan {
public:
    Plan(Context& ctx, const opiv_plan_desc& desc) : ctx_(ctx) { check(opiv_plan_create(ctx.get(), &desc, &h_)); }
    ~Plan() { opiv_plan_destroy(h_); }
    Plan(const Plan&) = delete;
    Plan& operator=(const Plan&) = delete;

    size_t window_count() const
    {
        size_t n = 0;
        check(opiv_plan_window_count(h_, &n));
        return n;
    }

    /// bottom-left corners in grid order (generate_cartesian_grid)
    std::vector<int32_t> grid() const
    {
        std::vector<int32_t> xy(2 * window_count());
        check(opiv_plan_grid(h_, xy.data()));
        return xy;
    }

    /// one frame pair (host or device pointers) -> one result per window
    std::vector<opiv_window_result> run(const void* a, const void* b) const
    {
        std::vector<opiv_window_result> out(window_count());
        std::lock_guard<std::mutex> lock(ctx_.mutex());
        check(opiv_run_pair(h_, a, b, out.data(), out.size()));
        return out;
    }

    /// gradients of sum_i cot[2i] u_i + cot[2i+1] v_i with respect to both
    /// frames (pitch * img_h floats each); fwd = run()'s results for the same
    /// frames, or nullptr to run the forward pass inside
    void vjp(const void* a, const void* b, const opiv_window_result* fwd, const double* cot,
             float* grad_a, float* grad_b) const
    {
        std::lock_guard<std::mutex> lock(ctx_.mutex());
        check(opiv_vjp_pair(h_, a, b, fwd, cot, grad_a, grad_b));
    }

    /// (var u_i, var v_i) per window under i.i.d. pixel noise sigma
    std::vector<double> variance(const void* a, const void* b, const opiv_window_result* fwd, double sigma) const
    {
        std::vector<double> var(2 * window_count());
        std::lock_guard<std::mutex> lock(ctx_.mutex());
        check(opiv_variance_pair(h_, a, b, fwd, sigma, var.data()));
        return var;
    }

    opiv_plan* get() const { return h_; }

private:
    Context& ctx_;
    opiv_plan* h_ = nullptr;
};

/// Many frame pairs (device memory) x up to four window passes, replayed as
/// one graph launch per run().
class Batch {
public:
    Batch(Context& ctx, const opiv_batch_desc& desc, const std::vector<opiv_pair_desc>& pairs)
        : ctx_(ctx), npasses_(desc.npasses), npairs_(pairs.size())
    {
        check(opiv_batch_create(ctx.get(), &desc, pairs.data(), pairs.size(), &h_));
    }
    ~Batch() { opiv_batch_destroy(h_); }
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;

    size_t window_count() const
    {
        size_t n = 0;
        check(opiv_batch_window_count(h_, &n));
        return n;
    }

    /// results of (pass p, pair i) start at offsets()[p * npairs + i]
    std::vector<size_t> offsets() const
    {
        std::vector<size_t> off(npasses_ * npairs_ + 1);
        check(opiv_batch_offsets(h_, off.data()));
        return off;
    }

    /// new device frames of the same sizes for the next run()
    void set_pairs(const std::vector<opiv_pair_desc>& pairs) { check(opiv_batch_set_pairs(h_, pairs.data(), pairs.size())); }

    /// enqueue every pass on the context stream; d_out: window_count() results in device memory
    void run(opiv_window_result* d_out, size_t n)
    {
        std::lock_guard<std::mutex> lock(ctx_.mutex());
        check(opiv_batch_run(h_, d_out, n));
    }

    opiv_batch* get() const { return h_; }

private:
    Context& ctx_;
    size_t npasses_, npairs_;
    opiv_batch* h_ = nullptr;
};

} // namespace openpiv_amd
