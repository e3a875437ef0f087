// transform.hip -- the complex 2-D DFT of algos::FFT::transform
// (fft.h:86-119): un-normalised, exp(-i 2 pi (ky y / H + kx x / W)) forward
// and exp(+i ...) reverse (fft.h:274-294 twiddle signs), for n stacked square
// windows of interleaved complex FP64 values.  FP32 arithmetic.
//
// Mapping (windows 8 to 64): W lanes own a window (lane t = column t); the
// column FFT runs in VGPRs, an LDS transpose hands lane t row t for the row
// FFT, and a second transpose lets lane t store output column t, so loads and
// stores are coalesced across the lanes of a row.  128 x 128 windows run on
// the lane pairs of the 128 x 128 correlation kernel (corr_fft128.hip).
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "fft_reg.hpp"
#include "transform.hpp"

namespace opiv {

using fftreg::brev_n;
using fftreg::sfor;

namespace {

__device__ __forceinline__ void wave_sync_t()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

} // namespace

template <int W, int SIGN>
__global__ __launch_bounds__(256) void fft_transform_kernel(const double2* in, double2* out, int n)
{
    constexpr int WPW = 64 / W;
    __shared__ float2 lds_all[4][WPW][W][W + 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane / W, t = lane % W;
    const int win = (blockIdx.x * 4 + wave) * WPW + g;
    const bool active = win < n;
    const size_t base = (size_t)(active ? win : n - 1) * W * W;
    float2 (*lds)[W + 1] = lds_all[wave][g];

    float2 z[W];
    sfor<W>([&](auto Y) {
        constexpr int y = decltype(Y)::value;
        const double2 v = in[base + (size_t)y * W + t];
        z[y] = make_float2((float)v.x, (float)v.y);
    });
    fftreg::fft_dit<W, SIGN>(z);                  // column t: frequency ky at z[brev(ky)]
    sfor<W>([&](auto K) {
        constexpr int ky = decltype(K)::value;
        lds[ky][t] = z[brev_n<W>(ky)];
    });
    wave_sync_t();
    sfor<W>([&](auto X) {
        constexpr int x = decltype(X)::value;
        z[x] = lds[t][x];                        // row ky = t
    });
    fftreg::fft_dit<W, SIGN>(z);                  // frequency kx at z[brev(kx)]
    wave_sync_t();                                // the row reads are done
    sfor<W>([&](auto K) {
        constexpr int kx = decltype(K)::value;
        lds[t][kx] = z[brev_n<W>(kx)];
    });
    wave_sync_t();
    if (active) {
        sfor<W>([&](auto K) {
            constexpr int ky = decltype(K)::value;
            const float2 v = lds[ky][t];
            out[base + (size_t)ky * W + t] = make_double2((double)v.x, (double)v.y);
        });
    }
}

template <int W>
static hipError_t launch_w(const double* in, double* out, size_t n, bool forward, hipStream_t stream)
{
    constexpr int per_block = 4 * (64 / W);
    const unsigned blocks = (unsigned)((n + per_block - 1) / per_block);
    const auto* src = reinterpret_cast<const double2*>(in);
    auto* dst = reinterpret_cast<double2*>(out);
    if (forward)
        hipLaunchKernelGGL((fft_transform_kernel<W, -1>), dim3(blocks), dim3(256), 0, stream, src, dst, (int)n);
    else
        hipLaunchKernelGGL((fft_transform_kernel<W, +1>), dim3(blocks), dim3(256), 0, stream, src, dst, (int)n);
    return hipGetLastError();
}

hipError_t launch_fft_transform(const double* in, double* out, int win, size_t n, bool forward, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    switch (win) {
    case 8: return launch_w<8>(in, out, n, forward, stream);
    case 16: return launch_w<16>(in, out, n, forward, stream);
    case 32: return launch_w<32>(in, out, n, forward, stream);
    case 64: return launch_w<64>(in, out, n, forward, stream);
    case 128: return launch_fft_transform128(in, out, n, forward, stream);
    default: return hipErrorInvalidValue;
    }
}

} // namespace opiv
