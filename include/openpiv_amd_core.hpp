// openpiv_amd_core.hpp -- the engine as a drop-in for algos::FFT on
// openpiv-cxx's own image types.
//
// Include this header where openpiv-cxx's headers are on the include path
// (it includes "core/image.h", "core/size.h", "core/pixel_types.h" and
// "algos/fft_common.h" from there).  It defines openpiv::algos::HipFFT with
// algos::FFT's call surface (algos/fft.h:28-295):
//
//   HipFFT(const core::size&)                          fft.h:76-84
//   const core::cf_image& transform(im, direction)     fft.h:86-119
//   image<g<ValueT>> cross_correlate(a, b)             fft.h:188-206
//   image<g<ValueT>> cross_correlate_real(a, b)        fft.h:208-227
//   const core::cf_image& auto_correlate(a)            fft.h:229-242
//
// so the correlator registry of examples/process/main.cpp:153-186 takes it by
// changing the class name only:
//
//   static algos::HipFFT fft{ ia };
//   return fft.cross_correlate(im_a, im_b);
//
// and core::find_peaks / core::fit_simple_gaussian (image_utils.h:37-49) run
// unchanged on the returned plane.  The correlation runs on the GPU of the
// process-wide context (openpiv_amd::default_context(): device
// $OPIV_DEVICE, else 0) in FP32 arithmetic; planes come back as FP64 images.
// As in the reference, the references that transform and auto_correlate
// return point into storage owned per thread and per instance, valid until
// the next call on that thread, and one instance may be called from many
// threads.
// Errors are std::runtime_error (openpiv_amd::error) with the reference's
// texts ("dimensions must be power of 2: [w,h]", "image size is different
// from expected: ...").
#pragma once

#include <atomic>
#include <cstddef>
#include <cstdint>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "algos/fft_common.h"
#include "core/image.h"
#include "core/pixel_types.h"
#include "core/size.h"
#include "openpiv_amd.hpp"

namespace openpiv::algos {

class HipFFT {
public:
    /// fft.h:76-84: power-of-two sizes only ("dimensions must be power of 2");
    /// the engine correlates square windows of 8 to 128 pixels
    explicit HipFFT(const core::size& size) : size_(size), fft_(size) {}
    HipFFT(const core::size& size, openpiv_amd::Context& ctx) : size_(size), fft_(ctx, size.width(), size.height()) {}

    const core::size& size() const { return size_; }

    /// fft.h:86-119: un-normalised 2-D DFT of a real or complex image
    template <template <typename> class ImageT, typename ContainedT,
              typename = typename std::enable_if_t<core::is_imagetype_v<ImageT<ContainedT>>>>
    const core::cf_image& transform(const ImageT<ContainedT>& input, direction d = direction::FORWARD) const
    {
        check_size(input);
        auto& c = cache();
        const uint32_t w = size_.width(), h = size_.height();
        c.z.resize(2 * (size_t)w * h);
        for (uint32_t y = 0; y < h; ++y) {
            const ContainedT* row = input.line(y);
            for (uint32_t x = 0; x < w; ++x) {
                const size_t i = (size_t)y * w + x;
                if constexpr (core::is_real_mono_pixeltype_v<ContainedT>) {
                    c.z[2 * i] = static_cast<double>(row[x]);
                    c.z[2 * i + 1] = 0.0;
                } else {
                    c.z[2 * i] = static_cast<double>(row[x].real);
                    c.z[2 * i + 1] = static_cast<double>(row[x].imag);
                }
            }
        }
        fft_.transform_raw(c.z.data(), c.z.data(), d == direction::FORWARD ? OPIV_FORWARD : OPIV_REVERSE);
        fill_complex(c.output, c.z);
        return c.output;
    }

    /// fft.h:188-206: real(IFFT(FFT(b) conj(FFT(a)))), quadrant-swapped
    template <template <typename> class ImageT, typename ContainedT,
              typename ValueT = typename ContainedT::value_t, typename OutT = core::image<core::g<ValueT>>,
              typename = typename std::enable_if_t<core::is_imagetype_v<ImageT<ContainedT>>>>
    OutT cross_correlate(const ImageT<ContainedT>& a, const ImageT<ContainedT>& b) const
    {
        return correlate<OutT>(a, &b, OPIV_CIRCULAR_FFT_FP32);
    }

    /// fft.h:208-227, including the reference's two-real-image unpack
    /// (spectral row 0, row H/2 and column 0 left empty, fft.h:166-183)
    template <template <typename> class ImageT, typename ContainedT,
              typename ValueT = typename ContainedT::value_t, typename OutT = core::image<core::g<ValueT>>,
              typename = typename std::enable_if_t<core::is_imagetype_v<ImageT<ContainedT>> &&
                                                   core::is_real_mono_pixeltype_v<ContainedT>>>
    OutT cross_correlate_real(const ImageT<ContainedT>& a, const ImageT<ContainedT>& b) const
    {
        return correlate<OutT>(a, &b, OPIV_REAL_COMPAT);
    }

    /// fft.h:229-242: the quadrant-swapped auto-correlation as a complex
    /// image with zero imaginary parts
    template <template <typename> class ImageT, typename ContainedT,
              typename = typename std::enable_if_t<core::is_imagetype_v<ImageT<ContainedT>>>>
    const core::cf_image& auto_correlate(const ImageT<ContainedT>& a) const
    {
        check_size(a);
        auto& c = cache();
        gather(a, c.a);
        c.plane.resize(c.a.size());
        fft_.auto_correlate_raw(c.a.data(), c.plane.data());
        c.output.resize(size_);
        for (size_t i = 0; i < c.plane.size(); ++i) c.output[i] = core::c_f{ c.plane[i], 0.0 };
        return c.output;
    }

private:
    struct data_t {
        core::cf_image output;
        std::vector<double> z, a, b, plane;
    };

    /// per-thread, per-instance storage, as algos::FFT keeps (fft.h:45-73)
    data_t& cache() const
    {
        thread_local std::vector<std::tuple<const HipFFT*, uint64_t, data_t>> storage;
        for (auto& [self, serial, data] : storage)
            if (self == this && serial == serial_) return data;
        auto& e = storage.emplace_back(this, serial_, data_t{});
        return std::get<2>(e);
    }

    template <class ImageT>
    void check_size(const ImageT& im) const
    {
        if (im.width() != size_.width() || im.height() != size_.height())
            throw openpiv_amd::error(OPIV_E_SIZE, "image size is different from expected: [" +
                                                      std::to_string(im.width()) + "," + std::to_string(im.height()) +
                                                      "], [" + std::to_string(size_.width()) + "," +
                                                      std::to_string(size_.height()) + "]");
    }

    template <class ImageT>
    void gather(const ImageT& im, std::vector<double>& out) const
    {
        const uint32_t w = size_.width(), h = size_.height();
        out.resize((size_t)w * h);
        for (uint32_t y = 0; y < h; ++y) {
            const auto* row = im.line(y);
            for (uint32_t x = 0; x < w; ++x) out[(size_t)y * w + x] = static_cast<double>(row[x]);
        }
    }

    void fill_complex(core::cf_image& out, const std::vector<double>& z) const
    {
        out.resize(size_);
        for (size_t i = 0; i < z.size() / 2; ++i) out[i] = core::c_f{ z[2 * i], z[2 * i + 1] };
    }

    template <class OutT, template <typename> class ImageT, typename ContainedT>
    OutT correlate(const ImageT<ContainedT>& a, const ImageT<ContainedT>* b, opiv_variant variant) const
    {
        check_size(a);
        check_size(*b);
        auto& c = cache();
        gather(a, c.a);
        gather(*b, c.b);
        c.plane.resize(c.a.size());
        fft_.cross_correlate(c.a.data(), c.b.data(), 1, c.plane.data(), variant);
        OutT out(size_);
        for (size_t i = 0; i < c.plane.size(); ++i) out[i] = c.plane[i];
        return out;
    }

    static uint64_t next_serial()
    {
        static std::atomic<uint64_t> n{ 0 };
        return ++n;
    }

    core::size size_;
    openpiv_amd::HipFFT fft_;
    uint64_t serial_ = next_serial();   // tells a new instance at a reused address from an old one
};

} // namespace openpiv::algos
