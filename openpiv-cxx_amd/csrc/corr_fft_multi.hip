// corr_fft_multi.hip -- every window pass of a batch in one persistent
// launch: the waves of one grid take tasks from all passes (window sizes 8 to
// 64), each task running the same per-window code as the per-pass kernels
// (corr_task of corr_fft.hip), so a window's results do not depend on which
// launch form computed it.  The reference runs one process per pair and
// window size (examples/process/main.cpp:137-302); this is the batch form
// of that loop.
//
// Each XCD takes a contiguous share of every pass (xcd_share per pass): the
// overlapping windows of a pass keep their frame rows in one L2, and every
// wave works through pass after pass, so the passes' unequal task costs
// even out over the grid.  The kernel's registers and LDS are those of the
// most demanding pass it can run (64 x 64: one 16.6 KB slice and up to 256
// VGPRs per wave), which caps every pass at two waves per SIMD; the batch's
// tuner times this form against one launch per pass and keeps the faster.
#define OPIV_CORR_FFT_W64_ONLY 1
#include "corr_fft.hip"

namespace opiv {

namespace {

constexpr int kMultiNT = 256;
constexpr int kMultiWaves = kMultiNT / 64;

// LDS of one wave for window size W: the per-pass kernels' slices
template <int W>
constexpr int slice_bytes() { return (64 / W) * W * (W + 1) * (int)sizeof(LdsT<W>); }
constexpr int kMultiSlice = slice_bytes<64>() > slice_bytes<8>() ? slice_bytes<64>() : slice_bytes<8>();

template <int W, int DT>
__device__ __forceinline__ void multi_pass(const CorrParams& p, unsigned char* wave_lds, int lane)
{
    constexpr int ROUNDS = rounds_for<W, DT>();
    constexpr int WPW = 64 / W;
    auto* lds = reinterpret_cast<LdsT<W> (*)[W][W + 1]>(wave_lds);
    const int wave = threadIdx.x >> 6;
    const int ntasks = (p.n_windows + ROUNDS * WPW - 1) / (ROUNDS * WPW);
    const XcdShare sh = xcd_share((int)blockIdx.x, (int)gridDim.x, ntasks);
    for (int task = sh.lo + sh.rank * kMultiWaves + wave; task < sh.hi; task += sh.blocks * kMultiWaves)
        corr_task<W, DT, false, ROUNDS, true>(p, task, lds[lane / W], lane);
}

} // namespace

template <int DT>
__global__ __launch_bounds__(kMultiNT) __attribute__((amdgpu_waves_per_eu(2)))
void corr_fft_multipass_kernel(MultiPassParams mp)
{
    __shared__ __attribute__((aligned(16))) unsigned char lds_all[kMultiWaves][kMultiSlice];
    const int lane = threadIdx.x & 63;
    unsigned char* mine = lds_all[threadIdx.x >> 6];
    for (int ps = 0; ps < mp.npasses; ++ps) {      // uniform: every wave sees the same pass order
        const CorrParams& p = mp.pass[ps];
        switch (mp.win[ps]) {
        case 8: multi_pass<8, DT>(p, mine, lane); break;
        case 16: multi_pass<16, DT>(p, mine, lane); break;
        case 32: multi_pass<32, DT>(p, mine, lane); break;
        case 64: multi_pass<64, DT>(p, mine, lane); break;
        default: break;
        }
    }
}

hipError_t launch_corr_fft_multipass(const MultiPassParams& mp, int dtype, int blocks, hipStream_t stream)
{
    if (mp.npasses < 1 || mp.npasses > 4 || blocks <= 0) return hipErrorInvalidValue;
    for (int ps = 0; ps < mp.npasses; ++ps)
        if (mp.win[ps] < 8 || mp.win[ps] > 64 || mp.pass[ps].pairs == nullptr) return hipErrorInvalidValue;
    switch (dtype) {
    case 0: hipLaunchKernelGGL((corr_fft_multipass_kernel<0>), dim3(blocks), dim3(kMultiNT), 0, stream, mp); break;
    case 1: hipLaunchKernelGGL((corr_fft_multipass_kernel<1>), dim3(blocks), dim3(kMultiNT), 0, stream, mp); break;
    case 2: hipLaunchKernelGGL((corr_fft_multipass_kernel<2>), dim3(blocks), dim3(kMultiNT), 0, stream, mp); break;
    case 3: hipLaunchKernelGGL((corr_fft_multipass_kernel<3>), dim3(blocks), dim3(kMultiNT), 0, stream, mp); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

} // namespace opiv
