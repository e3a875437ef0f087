// The 64 x 64 instances of corr_fft_kernel for f32 frames, compiled on their
// own with the max-memory-clause machine scheduler (Makefile): it keeps their
// two-window tasks within 256 VGPRs without scratch, where the default
// scheduler spills ~20 VGPRs per lane (4096^2 pair: 2% faster, HBM traffic
// 2.37x -> 1.82x the algorithmic bytes; the batch's 64 x 64 pass ~5% faster).
// 8- and 16-bit and f64 frames, and the 8-32 windows, keep the default
// scheduler (corr_fft.hip): for u16 frames at 64 x 64 this one is 6% slower.
#define OPIV_CORR_FFT_W64_ONLY 1
#include "corr_fft.hip"

namespace opiv {

template <int DT>
hipError_t launch_w64(const CorrParams& p, bool planes, int nt, int blocks, hipStream_t stream)
{
    return launch_w<64, DT>(p, planes, nt, blocks, stream);
}

template hipError_t launch_w64<2>(const CorrParams&, bool, int, int, hipStream_t);

} // namespace opiv
