// corr_direct32.hip -- direct circular cross-correlation of 32 x 32 windows
// on int8 MFMA for 8- and 16-bit frames, fused with find_peaks and the
// Gaussian fit.  Same semantics and outputs as the FFT kernel (corr_fft.hip):
// algos::FFT::cross_correlate (fft.h:188-206) + swap_quadrants +
// find_peaks(.., 2, 1) + fit_simple_gaussian + the vector of
// examples/process/main.cpp:220-234.
//
// The correlation of one window is one GEMM.  With k = (y, x),
//   C[sy][sx] = sum_k P[sy][k] Q[k][sx],  P[sy][(y,x)] = a[(y - sy) % 32][x],
//                                          Q[(y,x)][sx] = b[y][(x + sx) % 32]
// (M = N = 32, K = 1024), as four 16 x 16 tiles (T, U) of
// v_mfma_i32_16x16x64_i8 over 16 K-chunks.  K-chunk c gives lane group g
// (lanes 16 g .. 16 g + 15) window row y = 2c + g / 2 and columns
// 16 (g % 2) .. + 15; the same k slots in A and B, so the MFMA sums exactly
// the window's products.
//   * A (tile T, lane i = lane % 16): output row sy = 2 i + T, 16 bytes of
//     row (y - sy) % 32 of a.  From one chunk to the next y grows by 2, which
//     is what row sy - 2, lane i - 1 of the same tile, held (lane 0 takes
//     lane 15): one DPP row_ror:1 per operand register, and the A operands
//     never return to LDS.
//   * B (tile U): 16 bytes of row y of b starting at column
//     (16 (g % 2) + 16 U + i) % 32, read from one of four byte-shifted copies
//     of the doubled row so the run is dword-aligned.
// Pixels enter as signed bytes p - 128 (the top bit flipped); for a circular
// correlation sum_n (a' + 128)(b' + 128)[n + s] = C'[s] + 128 (Sa' + Sb') +
// 1024 * 128^2 with Sa', Sb' the window sums, all exact in int32.  16-bit
// pixels split exactly into a high and a low byte, p = 256 h + l, each byte
// plane offset the same way, and the four byte-plane correlations combine as
// 65536 C_hh + 256 (C_hl + C_lh) + C_ll.  The integer correlation is rounded
// once, to FP32, for the peak search.
//
// A wave correlates four windows in turn, two per 32-lane group, keeps
// their planes in registers (16 values per lane each), then moves them, one
// round of two windows at a time, through the wave's LDS slice into the
// column layout of the FFT kernel's tail (finish_windows: lane t of a group
// holds column t of its windows' planes), whose fits, reductions and logs
// then serve two windows per group.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <type_traits>

#include "corr_common.hpp"
#include "corr_direct.hpp"
#include "fft_reg.hpp"

namespace opiv {

using fftreg::sfor;

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int W = 32;
constexpr int WAVES = 4;                   // waves per block, one task each
constexpr int ROUNDS = 2, WPW = 2, PER = ROUNDS * WPW;

// LDS staging of one window, per byte plane: a as 32 rows of 8 dwords
// (padded to 12: the rows of one b128 read group start in distinct bank
// quads), b as four copies of its rows extended periodically to 48 bytes
// (12 dwords, enough for a 16-byte run from any of the 32 starts; copy s
// shifted by s bytes: byte m of row y holds b[y][(m + s) % 32]).  The copies
// start 8 banks apart, so the 32 lanes of one B-operand read pass (4 copies
// x 2 column halves x 4 dword offsets) hit 32 distinct banks.
constexpr int APITCH = 12;
constexpr int BROW = 12;
constexpr int BCOPY = W * BROW + 8;
struct PlaneStage {
    uint32_t a[W * APITCH];
    uint32_t b[4][BCOPY];
};
// the wave's tail slices: one plane per lane group, column-major, W + 1 rows
constexpr int SLICE = WPW * W * (W + 1);

template <int DT>
constexpr int nplanes() { return DT == 0 ? 1 : 2; }

template <int DT>
constexpr int lds_words()
{
    constexpr int stage = nplanes<DT>() * (int)(sizeof(PlaneStage) / 4);
    return stage > SLICE ? stage : SLICE;
}

__device__ __forceinline__ uint32_t dpp_ror1(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x121, 0xf, 0xf, false);
}

// 16 consecutive pixels of one frame row (as bytes for 8-bit frames, as
// 16-bit values in pairs for 16-bit frames): dword loads when every row of
// the window is dword-aligned, else element loads
template <int DT>
struct Run {
    static constexpr int NW = DT == 0 ? 4 : 8;
    uint32_t w[NW];
    __device__ __forceinline__ void load(const void* base, size_t idx, bool aligned)
    {
        if (aligned) {
            const gptr<uint32_t> q = (gptr<uint32_t>)((gptr<uint8_t>)base + idx * (DT == 0 ? 1 : 2));
            sfor<NW>([&](auto J) { w[decltype(J)::value] = q[decltype(J)::value]; });
        } else if constexpr (DT == 0) {
            const gptr<uint8_t> q = (gptr<uint8_t>)base + idx;
            sfor<NW>([&](auto J) {
                constexpr int j = decltype(J)::value;
                w[j] = (uint32_t)q[4 * j] | ((uint32_t)q[4 * j + 1] << 8) | ((uint32_t)q[4 * j + 2] << 16) |
                       ((uint32_t)q[4 * j + 3] << 24);
            });
        } else {
            const gptr<uint16_t> q = (gptr<uint16_t>)base + idx;
            sfor<NW>([&](auto J) {
                constexpr int j = decltype(J)::value;
                w[j] = (uint32_t)q[2 * j] | ((uint32_t)q[2 * j + 1] << 16);
            });
        }
    }
    // the run as NP byte planes of 4 dwords, offset to signed bytes
    // (plane 0: the pixel (8-bit) or its high byte; plane 1: the low byte)
    __device__ __forceinline__ void planes(uint32_t (&pl)[nplanes<DT>()][4]) const
    {
        if constexpr (DT == 0) {
            sfor<4>([&](auto J) { pl[0][decltype(J)::value] = w[decltype(J)::value] ^ 0x80808080u; });
        } else {
            sfor<4>([&](auto J) {
                constexpr int j = decltype(J)::value;
                // w[2j] = pixels 4j, 4j+1; w[2j+1] = 4j+2, 4j+3 (little-endian halves)
                pl[0][j] = __builtin_amdgcn_perm(w[2 * j + 1], w[2 * j], 0x07050301u) ^ 0x80808080u;
                pl[1][j] = __builtin_amdgcn_perm(w[2 * j + 1], w[2 * j], 0x06040200u) ^ 0x80808080u;
            });
        }
    }
};

// sum over the wave
__device__ __forceinline__ int wave_sum(int v)
{
    sfor<6>([&](auto K) { v += group_partner<(1 << decltype(K)::value)>(v); });
    return v;
}

} // namespace

template <int DT, bool PLANES>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DT == 0 ? 4 : 2)))
void corr_direct32_kernel(CorrParams p)
{
    constexpr int NP = nplanes<DT>();
    constexpr int NPROD = NP * NP;             // byte-plane products
    __shared__ __attribute__((aligned(16))) uint32_t lds_all[WAVES][lds_words<DT>()];

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int task = xcd_block((int)blockIdx.x, (int)gridDim.x) * WAVES + wave;
    const int ntasks = (p.n_windows + PER - 1) / PER;
    if (task >= ntasks) return;                // whole wave idle
    const int slot = lane / W, t = lane % W;

    bool active[ROUNDS];
    WinGeom geo[ROUNDS];
    sfor<ROUNDS>([&](auto R) {
        constexpr int r = decltype(R)::value;
        geo[r] = window_geom<false>(p, (task * ROUNDS + r) * WPW + slot);
        active[r] = geo[r].valid;
    });

    PlaneStage* st = reinterpret_cast<PlaneStage*>(&lds_all[wave][0]);
    float* slice = reinterpret_cast<float*>(&lds_all[wave][0]);

    const int i16 = lane & 15, g = lane >> 4;
    const int srow = lane >> 1, shalf = lane & 1;          // staging: row, 16-pixel half
    // dword alignment of every row of the frames (uniform)
    constexpr size_t ES = DT == 0 ? 1 : 2;
    const bool aligned = (((uintptr_t)p.a | (uintptr_t)p.b) & 3) == 0 && ((p.pitch * ES) & 3) == 0;

    auto window_xy = [&](auto I, int& x0, int& y0) {
        constexpr int i = decltype(I)::value;
        x0 = __builtin_amdgcn_readlane(geo[i / WPW].x0, (i % WPW) * W);
        y0 = __builtin_amdgcn_readlane(geo[i / WPW].y0, (i % WPW) * W);
    };
    auto fetch = [&](auto I, Run<DT>& ra, Run<DT>& rb) {
        int x0, y0;
        window_xy(I, x0, y0);
        const bool al = aligned && (((size_t)x0 * ES) & 3) == 0;
        const size_t idx = (size_t)(y0 + srow) * p.pitch + (size_t)x0 + 16 * (size_t)shalf;
        ra.load(p.a, idx, al);
        rb.load(p.b, idx, al);
    };

    // the four windows' quadrant-swapped planes, MFMA layout: pv[i][T][U][r]
    // = S[(8 g + 2 r + T + 16) % 32][16 (1 - U) + i16] (tile row m = 4 g + r
    // is correlation row 2 m + T)
    float pv[PER][2][2][4];
    Run<DT> ra[2], rb[2];
    fetch(std::integral_constant<int, 0>{}, ra[0], rb[0]);

    sfor<PER>([&](auto I) {
        constexpr int i = decltype(I)::value;
        // ---- stage window i ----
        uint32_t pa[NP][4], pb[NP][4];
        ra[i & 1].planes(pa);
        rb[i & 1].planes(pb);
        if constexpr (i + 1 < PER) fetch(std::integral_constant<int, i + 1>{}, ra[(i + 1) & 1], rb[(i + 1) & 1]);
        int sums[NP][2];                                    // window sums of a' and b' per plane
        sfor<NP>([&](auto K) {
            constexpr int k = decltype(K)::value;
            int sa = 0, sb = 0;
            sfor<4>([&](auto J) {
                constexpr int j = decltype(J)::value;
                sa = __builtin_amdgcn_sdot4((int)pa[k][j], 0x01010101, sa, false);
                sb = __builtin_amdgcn_sdot4((int)pb[k][j], 0x01010101, sb, false);
            });
            sums[k][0] = wave_sum(sa);
            sums[k][1] = wave_sum(sb);
        });
        wave_sync();                                        // earlier readers of the staging are done
        sfor<NP>([&](auto K) {
            constexpr int k = decltype(K)::value;
            *reinterpret_cast<uint4*>(&st[k].a[srow * APITCH + 4 * shalf]) =
                make_uint4(pa[k][0], pa[k][1], pa[k][2], pa[k][3]);
            // the dword after this run: the other half's first (mod 32 bytes)
            const uint32_t next = (uint32_t)group_partner<1>((int)pb[k][0]);
            sfor<4>([&](auto S) {
                constexpr int s = decltype(S)::value;
                uint32_t wv[4];
                sfor<4>([&](auto J) {
                    constexpr int j = decltype(J)::value;
                    const uint32_t hi = j < 3 ? pb[k][j + 1 < 4 ? j + 1 : 3] : next;
                    wv[j] = s == 0 ? pb[k][j] : __builtin_amdgcn_alignbyte(hi, pb[k][j], s);
                });
                const uint4 v = make_uint4(wv[0], wv[1], wv[2], wv[3]);
                // dwords 0..7 of the row, then 8..11 repeat 0..3
                *reinterpret_cast<uint4*>(&st[k].b[s][srow * BROW + 4 * shalf]) = v;
                if (shalf == 0) *reinterpret_cast<uint4*>(&st[k].b[s][srow * BROW + 8]) = v;
            });
        });
        wave_sync();

        // ---- C = P Q: A operands of chunk 0, then 16 chunks ----
        uint32_t A[NP][2][4];
        sfor<NP>([&](auto K) {
            constexpr int k = decltype(K)::value;
            sfor<2>([&](auto Tt) {
                constexpr int T = decltype(Tt)::value;
                const int row = ((g >> 1) - 2 * i16 - T) & (W - 1);
                const uint4 v = *reinterpret_cast<const uint4*>(&st[k].a[row * APITCH + 4 * (g & 1)]);
                A[k][T][0] = v.x;
                A[k][T][1] = v.y;
                A[k][T][2] = v.z;
                A[k][T][3] = v.w;
            });
        });
        v4i acc[NPROD][2][2];
        sfor<NPROD>([&](auto Pp) {
            sfor<2>([&](auto Tt) {
                sfor<2>([&](auto Uu) {
                    acc[decltype(Pp)::value][decltype(Tt)::value][decltype(Uu)::value] = v4i{0, 0, 0, 0};
                });
            });
        });
#pragma unroll 2
        for (int c = 0; c < 16; ++c) {
            const int y = 2 * c + (g >> 1);
            v4i B[NP][2];
            sfor<NP>([&](auto K) {
                constexpr int k = decltype(K)::value;
                sfor<2>([&](auto Uu) {
                    constexpr int U = decltype(Uu)::value;
                    const int q = 4 * ((g & 1) ^ U) + (i16 >> 2);
                    const uint32_t* src = &st[k].b[i16 & 3][y * BROW + q];
                    B[k][U] = v4i{(int)src[0], (int)src[1], (int)src[2], (int)src[3]};
                });
            });
            sfor<NP>([&](auto KA) {
                constexpr int ka = decltype(KA)::value;
                sfor<2>([&](auto Tt) {
                    constexpr int T = decltype(Tt)::value;
                    const v4i a4 = {(int)A[ka][T][0], (int)A[ka][T][1], (int)A[ka][T][2], (int)A[ka][T][3]};
                    sfor<NP>([&](auto KB) {
                        constexpr int kb = decltype(KB)::value;
                        sfor<2>([&](auto Uu) {
                            constexpr int U = decltype(Uu)::value;
                            acc[ka * NP + kb][T][U] =
                                __builtin_amdgcn_mfma_i32_16x16x64_i8(a4, B[kb][U], acc[ka * NP + kb][T][U], 0, 0, 0);
                        });
                    });
                });
            });
            // the next chunk's A operands (the last chunk's rotation is unused)
            sfor<NP>([&](auto K) {
                sfor<2>([&](auto Tt) {
                    sfor<4>([&](auto D) {
                        uint32_t& v = A[decltype(K)::value][decltype(Tt)::value][decltype(D)::value];
                        v = dpp_ror1(v);
                    });
                });
            });

        }

        // ---- the integer correlation, rounded once to FP32 ----
        // byte-plane correlation x.y of offset bytes: + 128 (Sx' + Sy') + 1024 * 128^2
        auto corr = [&](int prod_idx, int v) -> int {
            const int ka = prod_idx / NP, kb = prod_idx % NP;
            return v + 128 * (sums[ka][0] + sums[kb][1]) + W * W * 128 * 128;
        };
        sfor<2>([&](auto Tt) {
            constexpr int T = decltype(Tt)::value;
            sfor<2>([&](auto Uu) {
                constexpr int U = decltype(Uu)::value;
                sfor<4>([&](auto R) {
                    constexpr int r = decltype(R)::value;
                    float v;
                    if constexpr (NP == 1) {
                        v = (float)corr(0, acc[0][T][U][r]);
                    } else {
                        const int hh = corr(0, acc[0][T][U][r]), hl = corr(1, acc[1][T][U][r]);
                        const int lh = corr(2, acc[2][T][U][r]), ll = corr(3, acc[3][T][U][r]);
                        v = __builtin_fmaf(65536.f, (float)hh, __builtin_fmaf(256.f, (float)(hl + lh), (float)ll));
                    }
                    pv[i][T][U][r] = v;
                });
            });
        });
    });

    // ---- planes to the column layout of the FFT tail: lane (slot, t) reads
    // column t of its windows' planes, one round at a time, the last round
    // first, so that the slice still holds the first round's planes for the
    // tail ----
    float s[ROUNDS][W];
    sfor<ROUNDS>([&](auto Rq) {
        constexpr int r = ROUNDS - 1 - decltype(Rq)::value;
        wave_sync();                                        // earlier readers of the slice are done
        sfor<WPW>([&](auto Sl) {
            constexpr int sl = decltype(Sl)::value;
            float* pl = slice + sl * W * (W + 1);
            sfor<2>([&](auto Tt) {
                constexpr int T = decltype(Tt)::value;
                sfor<2>([&](auto Uu) {
                    constexpr int U = decltype(Uu)::value;
                    const int col = 16 * (1 - U) + i16;
                    sfor<4>([&](auto Rr) {
                        constexpr int rr = decltype(Rr)::value;
                        pl[col * (W + 1) + ((8 * g + 2 * rr + T + W / 2) & (W - 1))] = pv[r * WPW + sl][T][U][rr];
                    });
                });
            });
        });
        wave_sync();
        const float* mine = slice + slot * W * (W + 1) + t * (W + 1);
        sfor<W>([&](auto Hh) { s[r][decltype(Hh)::value] = mine[decltype(Hh)::value]; });
    });

    // the sums are the correlation divided by W * W (the reference's
    // unnormalised inverse FFT), an exact power of two that the tail restores
    float unscale[ROUNDS];
    sfor<ROUNDS>([&](auto R) { unscale[decltype(R)::value] = (float)(W * W); });
    // the slice holds one round's planes (both groups): the first round's
    // are there, the second's go back from the column layout when the tail
    // asks for them
    float* pl = slice + slot * W * (W + 1);
    auto stage = [&](auto R) {
        constexpr int r = decltype(R)::value;
        if constexpr (r > 0) {
            wave_sync();                                    // the first round's readers are done
            sfor<W>([&](auto Hh) {
                constexpr int h = decltype(Hh)::value;
                pl[t * (W + 1) + h] = s[r][h];
            });
            wave_sync();
        }
    };
    finish_windows<W, ROUNDS, PLANES, true>(p, s, unscale, geo, active, lane, t, stage,
                                            [&](int, int row, int col) { return pl[col * (W + 1) + row]; });
}

template <int DT>
static hipError_t launch_direct32_dt(const CorrParams& p, bool planes, hipStream_t stream)
{
    const int blocks = (p.n_windows + WAVES * PER - 1) / (WAVES * PER);
    if (planes)
        hipLaunchKernelGGL((corr_direct32_kernel<DT, true>), dim3(blocks), dim3(256), 0, stream, p);
    else
        hipLaunchKernelGGL((corr_direct32_kernel<DT, false>), dim3(blocks), dim3(256), 0, stream, p);
    return hipGetLastError();
}

hipError_t launch_corr_direct32(const CorrParams& p, int dtype, bool planes, hipStream_t stream)
{
    if (p.n_windows <= 0) return hipSuccess;
    if (p.real_compat) return hipErrorInvalidValue;
    switch (dtype) {
    case 0: return launch_direct32_dt<0>(p, planes, stream);
    case 1: return launch_direct32_dt<1>(p, planes, stream);
    default: return hipErrorInvalidValue;
    }
}

} // namespace opiv
