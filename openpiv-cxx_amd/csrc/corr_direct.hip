// corr_direct.hip -- direct circular cross-correlation on MFMA (int8 for
// 8-bit frames, split bf16 operands otherwise; windows of 16 x 16), fused with
// find_peaks and the Gaussian fit.  Same semantics and outputs as the FFT kernel
// (corr_fft.hip): algos::FFT::cross_correlate (fft.h:188-206) +
// swap_quadrants + find_peaks(.., 2, 1) + fit_simple_gaussian + the vector
// of examples/process/main.cpp:220-234.
//
// The correlation of one window is one GEMM.  With k = (y, x),
//   C[sy][sx] = sum_k P[sy][k] Q[k][sx],  P[sy][(y,x)] = a[(y - sy) % W][x],
//                                          Q[(y,x)][sx] = b[y][(x + sx) % W]
// (M = N = 16, K = 256).  Both operands are read from LDS copies of the
// window: P as aligned 16-byte row pieces of a (the row index carries the
// shift), Q from doubled rows of b held in shifted copies (two for bf16
// elements, four for bytes), so that every run starting at any column is
// dword-aligned in one of the copies.
//
// Precision: 8-bit frames run on int8 MFMA (pixels offset by 128, the exact
// integer correlation; see stage8 / gemm8 below).  Other pixels are split
// into bf16 terms, x = x0 + x1 (+ x2), and the cross products are
// accumulated in FP32 by the MFMA:
//   u16 : x0 + x1 exactly (8 + 8 bits), all four products
//   f32 : three terms (f64 frames are narrowed to f32 first), the six
//         products x_i y_j with i + j <= 2 (relative error ~2^-24, Henry,
//         Tang & Heinecke, "Leveraging the bfloat16 AI datatype for
//         higher-precision computations", ARITH 2019)
// Each cross product has its own FP32 accumulator; the partial sums are
// added smallest product first.
//
// Mapping: a wave runs 8 windows through the MFMA one after another, then
// finishes all 8 with the tail shared with the FFT kernel (one FP64 log per
// lane for 8 windows).  MFMA lane layout (lane l): A[i][k], B[k][j] with
// i = j = l % 16, lane group l / 16 holding one run of k (8 bf16 values of
// the 16x16x32 form, 16 bytes of the 16x16x64 int8 form, the same k for the
// same position in A and B); the result D[i][j] has j = l % 16 and
// i = 4 (l / 16) + r in accumulator r.
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

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// number of bf16 terms per pixel type
template <int DT>
constexpr int nparts() { return DT == 0 ? 1 : (DT == 1 ? 2 : 3); }

__device__ __forceinline__ uint32_t bf16_bits(float x)
{
    return (uint32_t)__builtin_bit_cast(uint16_t, (__bf16)x);
}
__device__ __forceinline__ float bf16_val(uint32_t b)
{
    return __builtin_bit_cast(float, b << 16);
}

// split x into NP bf16 terms (round to nearest each time; every residual is
// exact in FP32)
template <int NP>
__device__ __forceinline__ void split(float x, uint32_t (&t)[NP])
{
    float r = x;
    sfor<NP>([&](auto K) {
        constexpr int k = decltype(K)::value;
        t[k] = bf16_bits(r);
        r = r - bf16_val(t[k]);
    });
}

// four consecutive pixels of one frame row in their stored type (converted to
// FP32 only when staged, so that loads stay in flight across earlier windows)
template <int DT>
struct RawQuad {
    using T = std::conditional_t<DT == 0, uint8_t, std::conditional_t<DT == 1, uint16_t,
              std::conditional_t<DT == 2, float, double>>>;
    T v[4];
    __device__ __forceinline__ void load(const void* base, size_t idx)
    {
        const gptr<T> q = (gptr<T>)base + idx;
        sfor<4>([&](auto J) {
            constexpr int j = decltype(J)::value;
            v[j] = q[j];
        });
    }
    __device__ __forceinline__ float get(int j) const { return (float)v[j]; }
};

// the value of the next lane of the quad (DPP quad_perm [1, 2, 3, 0])
__device__ __forceinline__ float from_quad_next(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x39, 0xf, 0xf, false));
}

} // namespace

// 8-bit frames without plane output take the tail one round (four windows)
// at a time: the wave's LDS then holds one window's operands or four planes
// (4.3 KB instead of 8.7 KB), which lets six waves per SIMD fit instead of
// four.  The kernel is latency-bound (the GEMMs, their staging and the tail
// are short dependent chains), so the extra waves are what it needs.
template <int DT, bool PLANES>
constexpr bool by_rounds() { return DT == 0 && !PLANES; }

template <int DT, bool PLANES>
constexpr int direct16_waves() { return by_rounds<DT, PLANES>() ? 6 : DT <= 1 ? 4 : (DT == 2 ? 3 : 2); }

// Wave task: 8 windows, (task * 2 + r) * 4 + slot for round r and slot --
// the numbering and column layout of the 16 x 16 FFT kernel, so the shared
// tail (finish_windows) turns the planes into results.  Each window in turn:
// stage its bf16 terms in LDS, run the MFMA GEMM, and park the scaled,
// quadrant-swapped plane in LDS; then lane (slot, t) reads column t of its
// two windows' planes.  (8-bit frames: each round's four windows, then that
// round's part of the tail.)
template <int DT, bool PLANES>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(direct16_waves<DT, PLANES>()))) void corr_direct16_kernel(CorrParams p)
{
    constexpr int W = 16;
    constexpr int NP = nparts<DT>();
    constexpr int WAVES = 4;
    constexpr int ROUNDS = 2, WPW = 4, PER = ROUNDS * WPW;
    constexpr bool RND = by_rounds<DT, PLANES>();
    // per wave: the operand staging of one window (rows of a; doubled rows of
    // b, two copies), later overlaid by the 8 swapped planes (RND: the 4
    // planes of one round)
    // A rows padded to 12 dwords: the 8 rows of a b128 read group start in
    // distinct bank quads.  The odd-shift copy of B sits 16 dwords off the
    // even one, so that even and odd lanes of a read hit different banks.
    constexpr int APITCH = 24, BCOPY = 2 * W * W + 32;
    typedef uint16_t StageA[NP][W][APITCH];
    typedef uint16_t StageB[NP][2][BCOPY];
    typedef float Planes[RND ? WPW : PER][W][W + 1];
    constexpr int STAGE = RND ? 0 : (int)(sizeof(StageA) + sizeof(StageB));
    constexpr int BYTES = STAGE > (int)sizeof(Planes) ? STAGE : (int)sizeof(Planes);
    __shared__ __attribute__((aligned(16))) unsigned char lds[WAVES][BYTES];

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int task = xcd_block((int)blockIdx.x, (int)gridDim.x) * WAVES + wave;
    const int ntasks = (p.n_windows + PER - 1) / PER;
    if (task >= ntasks) return;                 // whole wave idle
    const int slot = lane / W, t = lane % W;

    bool active[ROUNDS];
    WinGeom geo[ROUNDS];
    sfor<ROUNDS>([&](auto R) {
        constexpr int r = decltype(R)::value;
        geo[r] = window_geom<false>(p, (task * ROUNDS + r) * WPW + slot);
        active[r] = geo[r].valid;
    });

    StageA& lds_a = *reinterpret_cast<StageA*>(&lds[wave][0]);
    StageB& lds_b = *reinterpret_cast<StageB*>(&lds[wave][RND ? 0 : sizeof(StageA)]);
    Planes& lds_p = *reinterpret_cast<Planes*>(&lds[wave][0]);

    const int i16 = lane & 15;                  // MFMA: row of A / column of B / column of D
    const int g = lane >> 4;                    // MFMA: k group / row quad of D
    const int srow = lane >> 2, scol = 4 * (lane & 3);   // staging: row and first column
    floatx4 accs[PER];

    // raw pixels of the next windows, loaded DEPTH windows ahead of their staging
    constexpr int DEPTH = DT <= 1 ? PER : (DT == 2 ? 4 : 2);
    RawQuad<DT> ra_[DEPTH], rb_[DEPTH];
    // window i = (round i / 4, slot i % 4): its corner is that lane group's
    // geometry, read into scalars
    const size_t lane_off = (size_t)srow * p.pitch + (size_t)scol;
    auto fetch = [&](auto I, RawQuad<DT>& qa, RawQuad<DT>& qb) {
        constexpr int i = decltype(I)::value;
        const int x0 = __builtin_amdgcn_readlane(geo[i / WPW].x0, (i % WPW) * W);
        const int y0 = __builtin_amdgcn_readlane(geo[i / WPW].y0, (i % WPW) * W);
        const size_t idx = (size_t)y0 * p.pitch + (size_t)x0 + lane_off;
        qa.load(p.a, idx);
        qb.load(p.b, idx);
    };
    sfor<DEPTH>([&](auto I) {
        constexpr int i = decltype(I)::value;
        fetch(I, ra_[i], rb_[i]);
    });

    // ---- stage window i: lane holds row srow, columns scol .. scol + 3 ----
    auto stage = [&](auto I) {
        constexpr int i = decltype(I)::value;
        float av[4], bv[5];
        sfor<4>([&](auto J) {
            constexpr int j = decltype(J)::value;
            av[j] = ra_[i % DEPTH].get(j);
            bv[j] = rb_[i % DEPTH].get(j);
        });
        if constexpr (i + DEPTH < PER)
            fetch(std::integral_constant<int, i + DEPTH>{}, ra_[i % DEPTH], rb_[i % DEPTH]);
        bv[4] = from_quad_next(bv[0]);          // b[row][(scol + 4) % 16]
        uint32_t ta[4][NP], tb[5][NP];
        sfor<4>([&](auto J) {
            constexpr int j = decltype(J)::value;
            split<NP>(av[j], ta[j]);
        });
        sfor<5>([&](auto J) {
            constexpr int j = decltype(J)::value;
            split<NP>(bv[j], tb[j]);
        });
        sfor<NP>([&](auto K) {
            constexpr int k = decltype(K)::value;
            const uint2 pa = make_uint2(ta[0][k] | (ta[1][k] << 16), ta[2][k] | (ta[3][k] << 16));
            const uint2 pe = make_uint2(tb[0][k] | (tb[1][k] << 16), tb[2][k] | (tb[3][k] << 16));
            const uint2 po = make_uint2(tb[1][k] | (tb[2][k] << 16), tb[3][k] | (tb[4][k] << 16));
            *reinterpret_cast<uint2*>(&lds_a[k][srow][scol]) = pa;
            // copy 0: E[m] = b[m % 16]; copy 1: O[m] = b[(m + 1) % 16]
            *reinterpret_cast<uint2*>(&lds_b[k][0][srow * 2 * W + scol]) = pe;
            *reinterpret_cast<uint2*>(&lds_b[k][0][srow * 2 * W + scol + W]) = pe;
            *reinterpret_cast<uint2*>(&lds_b[k][1][srow * 2 * W + scol]) = po;
            *reinterpret_cast<uint2*>(&lds_b[k][1][srow * 2 * W + scol + W]) = po;
        });
    };

    // ---- C = P Q on bf16 MFMA (u16, f32, f64 frames): 8 k-chunks of 32, one
    // accumulator per cross product x_i y_j (u16: all four; f32: i + j <= 2),
    // so the dependent MFMA chains are short.  Generic, so that it is
    // compiled only for the pixel types that call it. ----
    auto gemm = [&](auto...) {
        static_assert(NP >= 2, "8-bit frames take the int8 path");
        constexpr int NACC = NP == 2 ? 4 : 6;
        floatx4 part[NACC];
        sfor<NACC>([&](auto Q) { part[decltype(Q)::value] = floatx4{0.f, 0.f, 0.f, 0.f}; });
        sfor<8>([&](auto Cc) {
            constexpr int c = decltype(Cc)::value;
            const int ra = (2 * c + (g >> 1) - i16) & (W - 1);
            const int rb = 2 * c + (g >> 1);
            const int o = 8 * (g & 1) + i16;            // first column of the B run
            const int copy = o & 1, base = o & ~1;
            bf16x8 A[NP], B[NP];
            sfor<NP>([&](auto K) {
                constexpr int k = decltype(K)::value;
                A[k] = *reinterpret_cast<const bf16x8*>(&lds_a[k][ra][8 * (g & 1)]);
                const uint32_t* q = reinterpret_cast<const uint32_t*>(&lds_b[k][copy][rb * 2 * W + base]);
                uint32_t w4[4] = {q[0], q[1], q[2], q[3]};
                B[k] = __builtin_bit_cast(bf16x8, w4);
            });
            if constexpr (NP == 2) {
                part[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[1], B[1], part[0], 0, 0, 0);
                part[1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[1], B[0], part[1], 0, 0, 0);
                part[2] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[0], B[1], part[2], 0, 0, 0);
                part[3] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[0], B[0], part[3], 0, 0, 0);
            } else {
                part[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[2], B[0], part[0], 0, 0, 0);
                part[1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[1], B[1], part[1], 0, 0, 0);
                part[2] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[0], B[2], part[2], 0, 0, 0);
                part[3] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[1], B[0], part[3], 0, 0, 0);
                part[4] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[0], B[1], part[4], 0, 0, 0);
                part[5] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[0], B[0], part[5], 0, 0, 0);
            }
        });
        // smallest products first
        floatx4 acc = part[0];
        sfor<NACC - 1>([&](auto Q) {
            constexpr int q = decltype(Q)::value + 1;
            sfor<4>([&](auto R) {
                constexpr int r = decltype(R)::value;
                acc[r] = acc[r] + part[q][r];
            });
        });
        return acc;
    };

    // ---- 8-bit frames: the exact integer correlation on
    // v_mfma_i32_16x16x64_i8.  Pixels enter as v - 128 (the byte with its
    // top bit flipped, a signed byte); for a circular correlation
    //   sum_n (a' + 128)(b' + 128)[n + s] = C'[s] + 128 (Sa' + Sb') + 256 * 128^2,
    // Sa', Sb' the window sums, and every term is an integer below 2^24, so
    // the FP32 plane is exact.  K-chunk c (64 = 4 rows x 16 columns) gives
    // lane group g window row y = 4c + g: the A operand of lane i16 is row
    // (y - i16) mod 16 of a' (chunk c + 1's is chunk c's of lane i16 - 4: DPP
    // row_ror:4), the B operand row y of b' rotated by i16, read from one of
    // four byte-shifted copies of the doubled row so that it is dword-aligned.
    // B rows are padded to 12 dwords and the copies start 8 banks apart
    // (200 dwords), so the 32 lanes of each pass of a B read (4 copies x 2
    // rows x 4 dword offsets) and of a staging store (8 rows x 4 dwords) hit
    // 32 distinct banks. ----
    constexpr int B8ROW = 12, B8COPY = 16 * B8ROW + 8;
    uint32_t (&a8)[W][4] = *reinterpret_cast<uint32_t (*)[W][4]>(&lds[wave][0]);
    uint32_t (&b8)[4][B8COPY] = *reinterpret_cast<uint32_t (*)[4][B8COPY]>(&lds[wave][W * 16]);
    static_assert(W * 16 + 4 * B8COPY * 4 <= BYTES, "8-bit staging fits the wave's LDS");
    typedef int v4i __attribute__((ext_vector_type(4)));
    // stage window i; returns Sa' + Sb' (every lane)
    auto stage8 = [&](auto I) -> int {
        constexpr int i = decltype(I)::value;
        uint32_t pa = 0, pb = 0;
        sfor<4>([&](auto J) {
            constexpr int j = decltype(J)::value;
            pa |= (uint32_t)ra_[i % DEPTH].v[j] << (8 * j);
            pb |= (uint32_t)rb_[i % DEPTH].v[j] << (8 * j);
        });
        if constexpr (i + DEPTH < PER)
            fetch(std::integral_constant<int, i + DEPTH>{}, ra_[i % DEPTH], rb_[i % DEPTH]);
        pa ^= 0x80808080u;
        pb ^= 0x80808080u;
        // b'[srow][(scol + 4) % 16 ..] from the next lane of the quad
        const uint32_t nb = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)pb, 0x39, 0xf, 0xf, false);
        a8[srow][lane & 3] = pa;
        sfor<4>([&](auto Jc) {                  // copy j: byte m holds b'[srow][(m + j) % 16]
            constexpr int j = decltype(Jc)::value;
            const uint32_t w = j == 0 ? pb : __builtin_amdgcn_alignbyte(nb, pb, j);
            b8[j][srow * B8ROW + (lane & 3)] = w;
            b8[j][srow * B8ROW + (lane & 3) + 4] = w;
        });
        int sum = __builtin_amdgcn_sdot4((int)pa, 0x01010101, 0, false) +
                  __builtin_amdgcn_sdot4((int)pb, 0x01010101, 0, false);
        sfor<6>([&](auto K) {
            constexpr int off = 1 << decltype(K)::value;
            sum += group_partner<off>(sum);
        });
        return sum;
    };
    auto gemm8 = [&](int corr) {
        v4i acc[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
        const uint4 a0 = *reinterpret_cast<const uint4*>(&a8[(g - i16) & (W - 1)][0]);
        uint32_t Aw[4] = {a0.x, a0.y, a0.z, a0.w};
        const uint32_t* brow = &b8[i16 & 3][g * B8ROW + (i16 >> 2)];
        sfor<4>([&](auto Cc) {
            constexpr int c = decltype(Cc)::value;
            if constexpr (c > 0) {
                sfor<4>([&](auto Mm) {
                    constexpr int m = decltype(Mm)::value;
                    Aw[m] = (uint32_t)__builtin_amdgcn_mov_dpp((int)Aw[m], 0x124, 0xf, 0xf, false);
                });
            }
            const uint32_t* q = brow + 4 * c * B8ROW;
            const v4i A = {(int)Aw[0], (int)Aw[1], (int)Aw[2], (int)Aw[3]};
            const v4i B = {(int)q[0], (int)q[1], (int)q[2], (int)q[3]};
            acc[c & 1] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A, B, acc[c & 1], 0, 0, 0);
        });
        floatx4 out;
        sfor<4>([&](auto R) {
            constexpr int r = decltype(R)::value;
            out[r] = (float)(acc[0][r] + acc[1][r] + corr);
        });
        return out;
    };

    // D[sy][sx]: sx = i16, sy = 4 g + r.  The reference's plane is the
    // un-normalised inverse FFT, W^2 times C (an exact power of two), and the
    // quadrant swap puts C[sy][sx] at plane (sy + 8) % 16, (sx + 8) % 16.
    const int wcol = (i16 + 8) & (W - 1);
    // the sums are the correlation divided by W * W (the reference's
    // unnormalised inverse FFT), an exact power of two that the tail restores
    float unscale[ROUNDS];
    sfor<ROUNDS>([&](auto R) { unscale[decltype(R)::value] = (float)(W * W); });

    if constexpr (RND) {
        // round r: its four windows' GEMMs, their planes to LDS, column t of
        // this lane group's plane; the tail scans it and reads the stencils
        // it needs before the next round overwrites the LDS
        float s[ROUNDS][W];
        auto run_round = [&](auto R) {
            constexpr int r = decltype(R)::value;
            floatx4 acc[WPW];
            wave_sync();                        // the previous round's readers are done
            sfor<WPW>([&](auto S) {
                constexpr int i = r * WPW + decltype(S)::value;
                const int sums = stage8(std::integral_constant<int, i>{});
                wave_sync();
                acc[decltype(S)::value] = gemm8(128 * sums + W * W * 128 * 128);
                wave_sync();                    // staging buffers free for the next window
            });
            sfor<WPW>([&](auto S) {
                constexpr int j = decltype(S)::value;
                sfor<4>([&](auto Rr) {
                    constexpr int q = decltype(Rr)::value;
                    lds_p[j][(4 * g + q + 8) & (W - 1)][wcol] = acc[j][q];
                });
            });
            wave_sync();
            sfor<W>([&](auto Hh) {
                constexpr int h = decltype(Hh)::value;
                s[r][h] = lds_p[slot][h][t];
            });
        };
        finish_windows<W, ROUNDS, false, true>(p, s, unscale, geo, active, lane, t, run_round,
                                                [&](int, int row, int col) { return lds_p[slot][row][col]; });
        return;
    } else {

    sfor<PER>([&](auto I) {
        constexpr int i = decltype(I)::value;
        if constexpr (DT == 0) {
            const int sums = stage8(I);
            wave_sync();
            accs[i] = gemm8(128 * sums + W * W * 128 * 128);
        } else {
            stage(I);
            wave_sync();
            accs[i] = gemm();
        }
        wave_sync();                            // staging buffers free for the next window
    });

    sfor<PER>([&](auto I) {
        constexpr int i = decltype(I)::value;
        sfor<4>([&](auto R) {
            constexpr int r = decltype(R)::value;
            lds_p[i][(4 * g + r + 8) & (W - 1)][wcol] = accs[i][r];
        });
    });
    wave_sync();

    // ---- column t of this lane group's two planes, then the common tail ----
    float s[ROUNDS][W];
    sfor<ROUNDS>([&](auto R) {
        constexpr int r = decltype(R)::value;
        sfor<W>([&](auto Hh) {
            constexpr int h = decltype(Hh)::value;
            s[r][h] = lds_p[r * WPW + slot][h][t];
        });
    });
    finish_windows<W, ROUNDS, PLANES, false>(p, s, unscale, geo, active, lane, t, [](auto) {}, [&](int r, int row, int col) {
        return lds_p[r * WPW + slot][row][col];
    });
    }
}

template <int DT>
static hipError_t launch_direct_dt(const CorrParams& p, bool planes, hipStream_t stream)
{
    const int blocks = (p.n_windows + 4 * 8 - 1) / (4 * 8);   // 4 waves x 8 windows
    if (planes)
        hipLaunchKernelGGL((corr_direct16_kernel<DT, true>), dim3(blocks), dim3(256), 0, stream, p);
    else
        hipLaunchKernelGGL((corr_direct16_kernel<DT, false>), dim3(blocks), dim3(256), 0, stream, p);
    return hipGetLastError();
}

hipError_t launch_corr_direct(const CorrParams& p, int win, int dtype, bool planes, hipStream_t stream)
{
    if (p.n_windows <= 0) return hipSuccess;
    if (win == 32) return launch_corr_direct32(p, dtype, planes, stream);
    if (win != 16 || p.real_compat) return hipErrorInvalidValue;
    switch (dtype) {
    case 0: return launch_direct_dt<0>(p, planes, stream);
    case 1: return launch_direct_dt<1>(p, planes, stream);
    case 2: return launch_direct_dt<2>(p, planes, stream);
    case 3: return launch_direct_dt<3>(p, planes, stream);
    default: return hipErrorInvalidValue;
    }
}

} // namespace opiv
