// corr_common.hpp -- device code shared by the correlation kernels: frame
// loads, window geometry, and the common tail that turns quadrant-swapped
// correlation planes held as one column per lane into results:
// find_peaks(.., 2, 1) (image_utils.impl.h:37-74), fit_simple_gaussian
// (image_utils.impl.h:76-101) and the vector of
// examples/process/main.cpp:220-234.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdint>

#include "corr_fft.hpp"
#include "xcd.hpp"
#include "fft_reg.hpp"
#include "fit_log.hpp"

namespace opiv {

using fftreg::sfor;

namespace {

// LDS writes of this wave become visible to its later LDS reads; no
// instruction is moved across.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// value of lane+1 / lane-1 (whole-wave DPP shifts; the group-edge lanes
// receive a neighbour group's value or zero, which the scan never uses)
__device__ __forceinline__ float from_right(float v)
{
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
}

// max(value of lane - 1, value of lane + 1, up, down): the left neighbour
// enters through DPP as an operand of the max itself.  Written out because a
// max of a DPP result in compiled code first canonicalises it (one more
// instruction).  s_nop 1 gives the two wait states a DPP read needs after a
// VALU write of its source.  Group-edge lanes take a neighbour group's value
// or zero, as with from_right.
__device__ __forceinline__ float neighbour_max(float c, float up, float down)
{
    const float r = from_right(c);
    float m;
    asm("s_nop 1\n\t"
        "v_max_f32_dpp %0, %1, %2 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:0\n\t"
        "v_max3_f32 %0, %0, %3, %4"
        : "=&v"(m)
        : "v"(c), "v"(r), "v"(up), "v"(down));
    return m;
}

// Partner exchange for reductions over lane groups of W lanes, applied with
// OFF = 1, 2, 4, ... W/2 in that order: lane ^ 1 and lane ^ 2 (quad_perm),
// the mirror lane of the 8-lane half (row_half_mirror), the mirror lane of the
// 16-lane row (row_mirror), then lane ^ OFF.  Each step pairs lanes whose
// accumulated sets are disjoint, and after the last one every lane holds its
// whole group.
template <int OFF>
__device__ __forceinline__ int group_partner(int v)
{
    if constexpr (OFF == 1) return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xf, 0xf, false);
    else if constexpr (OFF == 2) return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xf, 0xf, false);
    else if constexpr (OFF == 4) return __builtin_amdgcn_mov_dpp(v, 0x141, 0xf, 0xf, false);
    else if constexpr (OFF == 8) return __builtin_amdgcn_mov_dpp(v, 0x140, 0xf, 0xf, false);
    else return __shfl_xor(v, OFF, 64);
}
template <int OFF>
__device__ __forceinline__ float group_partner(float v)
{
    return __int_as_float(group_partner<OFF>(__float_as_int(v)));
}

// One step of a max (min) reduction with the partner of group_partner<OFF>.
// Within a 16-lane row the exchange is the DPP operand of the max itself: in
// compiled code a max of a DPP move first canonicalises the moved value, two
// extra instructions of the slow (4+ cycle) class per step.  s_nop 1 gives
// the two wait states a DPP read needs after a VALU write of its source.
#define OPIV_DPP_STEP(OP, CTRL)                                                                        \
    asm("s_nop 1\n\t" OP " %0, %1, %1 " CTRL " row_mask:0xf bank_mask:0xf" : "=v"(r) : "v"(v))
template <int OFF>
__device__ __forceinline__ float max_with_partner(float v)
{
    float r;
    if constexpr (OFF == 1) OPIV_DPP_STEP("v_max_f32_dpp", "quad_perm:[1,0,3,2]");
    else if constexpr (OFF == 2) OPIV_DPP_STEP("v_max_f32_dpp", "quad_perm:[2,3,0,1]");
    else if constexpr (OFF == 4) OPIV_DPP_STEP("v_max_f32_dpp", "row_half_mirror");
    else if constexpr (OFF == 8) OPIV_DPP_STEP("v_max_f32_dpp", "row_mirror");
    else {
        const float o = group_partner<OFF>(v);
        asm("v_max_f32_e32 %0, %1, %2" : "=v"(r) : "v"(v), "v"(o));
    }
    return r;
}
template <int OFF>
__device__ __forceinline__ int min_with_partner(int v)
{
    int r;
    if constexpr (OFF == 1) OPIV_DPP_STEP("v_min_i32_dpp", "quad_perm:[1,0,3,2]");
    else if constexpr (OFF == 2) OPIV_DPP_STEP("v_min_i32_dpp", "quad_perm:[2,3,0,1]");
    else if constexpr (OFF == 4) OPIV_DPP_STEP("v_min_i32_dpp", "row_half_mirror");
    else if constexpr (OFF == 8) OPIV_DPP_STEP("v_min_i32_dpp", "row_mirror");
    else r = min(v, group_partner<OFF>(v));
    return r;
}
#undef OPIV_DPP_STEP

// value of lane + D of the same 16-lane row (row_ror by 16 - D); the caller
// keeps lane % 16 + D below 16
template <int D>
__device__ __forceinline__ double from_row_ahead(double v)
{
    if constexpr (D == 0) return v;
    else {
        const int2 w = __builtin_bit_cast(int2, v);
        const int lo = __builtin_amdgcn_mov_dpp(w.x, 0x120 + (16 - D), 0xf, 0xf, false);
        const int hi = __builtin_amdgcn_mov_dpp(w.y, 0x120 + (16 - D), 0xf, 0xf, false);
        return __builtin_bit_cast(double, make_int2(lo, hi));
    }
}

// frames are global memory: address-space-1 pointers give global (not flat)
// loads even when the pointer comes from the batch table
template <class T>
using gptr = const __attribute__((address_space(1))) T*;

template <int DT>
__device__ __forceinline__ float load_px(const void* base, size_t idx)
{
    if constexpr (DT == 0) return (float)((gptr<uint8_t>)base)[idx];
    else if constexpr (DT == 1) return (float)((gptr<uint16_t>)base)[idx];
    else if constexpr (DT == 2) return ((gptr<float>)base)[idx];
    else return (float)((gptr<double>)base)[idx];
}

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }

} // namespace

// Where one window lives: frames, pitch, bottom-left corner, frame height,
// the index of its result and whether it is a real window of the launch
// (task slots past the end of a pair's range recompute that pair's last
// window and store nothing).
struct WinGeom {
    const void* a;
    const void* b;
    size_t pitch;
    int x0, y0;
    uint32_t img_h;
    int out;
    bool valid;
};

template <class T>
using gptr_t = const __attribute__((address_space(1))) T*;

// g / nx for a grid width nx >= 1 and its multiplier mul = fastdiv_mul(nx):
// one multiply-high and shifts instead of an integer division
__device__ __forceinline__ int div_grid_width(int g, int nx, uint32_t mul)
{
    const uint32_t l = 32u - (uint32_t)__clz(nx - 1);   // ceil(log2 nx); 0 for nx = 1
    return (int)fastdiv((uint32_t)g, l, mul);
}

// Geometry of window `local` of a launch: the single pair of the params, or
// (batch) the pair whose window range [first, first + nx * ny) holds it.
template <bool BATCH>
__device__ __forceinline__ WinGeom window_geom(const CorrParams& p, int local)
{
    WinGeom g;
    g.valid = local < p.n_windows;
    if constexpr (!BATCH) {
        const int l = g.valid ? local : p.n_windows - 1;
        const int gid = p.first + l;
        const int iy = div_grid_width(gid, p.nx, p.nx_mul), ix = gid - iy * p.nx;
        g.a = p.a;
        g.b = p.b;
        g.pitch = p.pitch;
        g.x0 = p.x_start + ix * p.step_x;
        g.y0 = p.y_start + iy * p.step_y;
        g.img_h = p.img_h;
        g.out = l;
    } else {
        // The table is written before the launch and only read here: the
        // constant address space lets uniform reads of it be scalar loads.
        const __attribute__((address_space(4))) BatchPair* pairs =
            (const __attribute__((address_space(4))) BatchPair*)p.pairs;
        const int l = g.valid ? local : p.n_windows - 1;
        // The windows of a wave task lie in one pair (pair ranges are
        // task-aligned): the search runs on the first lane's index, so that
        // the pair record is wave-uniform and stays in scalar registers.
        const int l0 = __builtin_amdgcn_readfirstlane(l);
        int lo = 0, hi = p.npairs;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (pairs[mid].first <= l0) lo = mid;
            else hi = mid;
        }
        const auto& bp = pairs[lo];
        const int n = bp.nx * bp.ny;
        int gid = l - bp.first;
        if (gid < 0 || gid >= n) {               // (gid < 0 only guards the addressing)
            g.valid = false;
            gid = gid < 0 ? 0 : n - 1;
        }
        const int iy = div_grid_width(gid, bp.nx, bp.nx_mul), ix = gid - iy * bp.nx;
        g.a = bp.a;
        g.b = bp.b;
        g.pitch = bp.pitch;
        g.x0 = bp.x_start + ix * bp.step_x;
        g.y0 = bp.y_start + iy * bp.step_y;
        g.img_h = bp.img_h;
        g.out = bp.out_first + gid;
    }
    return g;
}

// max / min over the lanes of each W-lane group, every lane receiving it
template <int W>
__device__ __forceinline__ float group_max(float v)
{
    sfor<ilog2(W)>([&](auto K) { v = max_with_partner<1 << decltype(K)::value>(v); });
    return v;
}
template <int W>
__device__ __forceinline__ int group_min(int v)
{
    sfor<ilog2(W)>([&](auto K) { v = min_with_partner<1 << decltype(K)::value>(v); });
    return v;
}

// find_peaks(.., 2, 1) over one quadrant-swapped plane column held by lane
// t: strict 4-neighbour maxima on rows [hlo, hhi] and columns [wlo, whi]
// (image_utils.impl.h:48-64), reduced to the largest value (and its scan
// index h*W + t; equal values keep scan order) and the second largest value
// and a scan index of it over the W lanes of the window.  own(h) reads the
// lane's plane value at row h from LDS.  Rows go in pairs: of two vertically
// adjacent elements at most one is a strict maximum (each must exceed the
// other), so each pair offers one candidate to the top 2, and the pair's
// number is all the scan keeps of its position.
template <int W, class Own>
__device__ __forceinline__ void scan_top2(const float (&s)[W], int t, int hlo, int hhi, bool col_ok, Own&& own,
                                          float& v1, int& i1, float& v2, int& i2)
{
    v1 = -INFINITY;
    v2 = -INFINITY;
    int p1 = 0;
    // candidate mc of pair P, in scan order: ties keep the first
    auto offer = [&](float mc, int P) {
        const bool gt = mc > v1;
        v2 = __builtin_amdgcn_fmed3f(v2, mc, v1); // = max(v2, min(mc, v1)), as v2 <= v1
        p1 = gt ? P : p1;                         // P is an inline constant of the select
        v1 = gt ? mc : v1;                        // a select is cheaper than a max here
    };
    // rows 1 .. W-3; the region's rows are wave-uniform tests, its columns
    // (col_ok) are applied once after the scan.  The region test is data
    // flow, so the candidate stays a select of canonical values.
    constexpr int NR = W - 3;
    sfor<NR / 2>([&](auto Pp) {
        constexpr int P = decltype(Pp)::value;
        constexpr int h = 2 * P + 1;
        const float c0 = s[h], c1 = s[h + 1];
        const bool m0 = (c0 > neighbour_max(c0, s[h - 1], c1)) & (h >= hlo) & (h <= hhi);
        const bool m1 = (c1 > neighbour_max(c1, c0, s[h + 2])) & (h + 1 >= hlo) & (h + 1 <= hhi);
        offer(m0 ? c0 : (m1 ? c1 : -INFINITY), P);
    });
    if constexpr (NR % 2 == 1) {
        constexpr int h = NR;
        const float c = s[h];
        const bool m = (c > neighbour_max(c, s[h - 1], s[h + 1])) & (h >= hlo) & (h <= hhi);
        offer(m ? c : -INFINITY, NR / 2);
    }
    // keeps the arithmetic below out of the selects above (folded in, it
    // turns every pair number into a non-inline constant held in a register)
    asm("" : "+v"(p1));
    v1 = col_ok ? v1 : -INFINITY;
    v2 = col_ok ? v2 : -INFINITY;
    // the best's row within its pair: the first row holds it exactly when
    // its value is the best (the pair's other row then differs from it);
    // a trailing single row is pair NR / 2's first
    const int hp = 2 * p1 + 1;
    const int h1 = own(hp) == v1 ? hp : hp + 1;
    const int li = h1 * W + t;                    // this lane's best as a scan index
    // top 2 of the group: the largest value; among the lanes holding it the
    // smallest index (scan order; indices of different lanes differ); then
    // the largest of that lane's second value and every other lane's first
    // (a tie with another lane's first gives v2 == v1, as the sort does)
    const float gv1 = group_max<W>(v1);
    i1 = group_min<W>(v1 == gv1 ? li : INT_MAX);
    const bool winner = i1 == li;
    const float cand = winner ? v2 : v1;
    const float gv2 = group_max<W>(cand);
    // where the second peak is (for its own near-tie test): a lane holding
    // it as its best knows its row; the winning lane's second value is
    // searched for in its column, which only waves that need it do
    constexpr int kSearch = INT_MAX - 1;
    i2 = group_min<W>(cand == gv2 ? (winner ? kSearch : li) : INT_MAX);
    if (__builtin_amdgcn_ballot_w64(i2 == kSearch) != 0) {
        int hs = 0x7fff;
        sfor<W - 3>([&](auto Hm) {
            constexpr int h = W - 3 - decltype(Hm)::value;   // descending: the first in scan order wins
            const bool inreg = col_ok & (h >= hlo) & (h <= hhi);
            hs = (inreg & (s[h] == gv2) & (h != h1)) ? h : hs;
        });
        const int found = group_min<W>(winner && hs != 0x7fff ? hs * W + t : INT_MAX);
        i2 = i2 == kSearch ? found : i2;
    }
    v1 = gv1;
    v2 = gv2;
}

// Results of the ROUNDS windows of this lane group from their swapped planes
// s[r][h] (column t of each): find_peaks, the FP64 fit with the 5 * ROUNDS
// logs spread over the group's lanes, and the vector; or, with PLANES, the
// planes themselves.  Plane r is the correlation divided by unscale[r], a
// power of two: the peak order, the tie tests and the log-ratio fit do not
// depend on it, and it is multiplied back (exactly) into the reported peak
// values and planes.  fetch(r, row, col) reads the value of round r's plane
// at (row, col) from LDS (any lane may read any position of its group's
// planes); it supplies the stencil around each peak.  With SEQ the LDS
// holds one plane at a time: stage(r) (a compile-time round) puts plane r
// there, and each lane reads the values it needs of one round before the
// next is staged; otherwise every plane is readable on entry.
template <int W, int ROUNDS, bool PLANES, bool SEQ, class Stage, class Fetch>
__device__ __forceinline__ void finish_windows(const CorrParams& p, const float (&s)[ROUNDS][W],
                                               const float (&unscale)[ROUNDS],
                                               const WinGeom (&geo)[ROUNDS], const bool (&active)[ROUNDS],
                                               int lane, int t, Stage&& stage, Fetch&& fetch)
{
    const int gbase = lane - t;
    if constexpr (PLANES) {
        sfor<ROUNDS>([&](auto R) {
            constexpr int r = decltype(R)::value;
            if (active[r]) {
                double* dst = p.planes + (size_t)geo[r].out * W * W + t;
                sfor<W>([&](auto Hh) {
                    constexpr int h = decltype(Hh)::value;
                    dst[h * W] = (double)(s[r][h] * unscale[r]);
                });
            }
        });
        return;
    } else {
        // search region: whole plane, or the limit-search view rect::dilate(0.5)
        const int r0 = p.limit_search ? W / 4 : 0;
        const int rlen = p.limit_search ? W / 2 : W;
        const int hlo = r0 + 1, hhi = r0 + rlen - 3;
        const bool col_ok = t >= r0 + 1 && t <= r0 + rlen - 2;

        float v1[ROUNDS], v2[ROUNDS];
        int px[ROUNDS], py[ROUNDS];
        int px2[ROUNDS], py2[ROUNDS];          // the second peak (1, 1 when unknown)
        bool two[ROUNDS], at2[ROUNDS];
        // the scan reads one value of the staged plane (its best's row)
        auto scan = [&](auto R) {
            constexpr int r = decltype(R)::value;
            int i1;
            int i2;
            scan_top2<W>(s[r], t, hlo, hhi, col_ok, [&](int h) { return fetch(r, h, t); }, v1[r], i1, v2[r], i2);
            two[r] = v2[r] > -INFINITY;
            px[r] = two[r] ? (i1 & (W - 1)) : 1;
            py[r] = two[r] ? (i1 / W) : 1;
            at2[r] = two[r] & (i2 < W * W);
            px2[r] = at2[r] ? (i2 & (W - 1)) : 1;
            py2[r] = at2[r] ? (i2 / W) : 1;
        };
        if constexpr (!SEQ) sfor<ROUNDS>(scan);
        // stencil value j of round r: (l, c, r, up, down) = plane positions
        // (py, px - 1), (py, px), (py, px + 1), (py - 1, px), (py + 1, px)
        auto stencil = [&](int r, int j) -> float {
            int cx = px[0], cy = py[0];
            sfor<ROUNDS - 1>([&](auto R) {
                constexpr int q = decltype(R)::value + 1;
                cx = r == q ? px[q] : cx;
                cy = r == q ? py[q] : cy;
            });
            cx += j == 0 ? -1 : (j == 2 ? 1 : 0);
            cy += j == 3 ? -1 : (j == 4 ? 1 : 0);
            return fetch(r, cy, cx);
        };

        // ---- FP64 logs of the 5 * ROUNDS stencil values, one per lane:
        // value j of round r sits on lane j * ROUNDS + r of the group ----
        constexpr int NV = 5 * ROUNDS;
        constexpr int NP = (NV + W - 1) / W;
        float sv[NP];          // the stencil value whose log this lane takes
        float nb[4];           // lanes t < ROUNDS: the four neighbours of round t's peak
        float nb2[4];          // and of its second peak
        // neighbour j (left, right, up, down) of round r's second peak
        auto around2 = [&](int r, int j) -> float {
            int cx = px2[0], cy = py2[0];
            sfor<ROUNDS - 1>([&](auto R) {
                constexpr int q = decltype(R)::value + 1;
                cx = r == q ? px2[q] : cx;
                cy = r == q ? py2[q] : cy;
            });
            cx += j == 0 ? -1 : (j == 1 ? 1 : 0);
            cy += j == 2 ? -1 : (j == 3 ? 1 : 0);
            return fetch(r, cy, cx);
        };
        if constexpr (SEQ) {
            sfor<ROUNDS>([&](auto R) {
                constexpr int r = decltype(R)::value;
                stage(R);
                scan(R);
                auto at = [&](int j) {
                    const int cx = px[r] + (j == 0 ? -1 : (j == 2 ? 1 : 0));
                    const int cy = py[r] + (j == 3 ? -1 : (j == 4 ? 1 : 0));
                    return fetch(r, cy, cx);
                };
                sfor<NP>([&](auto K) {
                    constexpr int k = decltype(K)::value;
                    const int L = t + k * W;
                    if (L % ROUNDS == r) sv[k] = at(L < NV ? L / ROUNDS : 1);
                });
                if (t == r) {
                    nb[0] = at(0);
                    nb[1] = at(2);
                    nb[2] = at(3);
                    nb[3] = at(4);
                    sfor<4>([&](auto J) {
                        constexpr int j = decltype(J)::value;
                        nb2[j] = fetch(r, py2[r] + (j == 2 ? -1 : (j == 3 ? 1 : 0)),
                                       px2[r] + (j == 0 ? -1 : (j == 1 ? 1 : 0)));
                    });
                }
            });
        } else {
            sfor<NP>([&](auto K) {
                constexpr int k = decltype(K)::value;
                const int L = t + k * W;
                sv[k] = stencil(L % ROUNDS, L < NV ? L / ROUNDS : 1);
            });
        }
        double lg[NP];
        sfor<NP>([&](auto K) {
            constexpr int k = decltype(K)::value;
            const int L = t + k * W;
            lg[k] = ln_of_float(L < NV ? sv[k] : 1.f);
        });

        // ---- lane t < ROUNDS of the group fits window t ----
        const int rr = t < ROUNDS ? t : 0;
        auto pick = [&](const auto (&arr)[ROUNDS]) {
            auto v = arr[0];
            sfor<ROUNDS - 1>([&](auto R) {
                constexpr int r = decltype(R)::value + 1;
                v = rr == r ? arr[r] : v;
            });
            return v;
        };
        // logs of round rr: value j from lane rr + j * ROUNDS
        double ln[5];
        if constexpr (W >= 16 && NV <= W) {
            sfor<5>([&](auto J) {
                constexpr int j = decltype(J)::value;
                ln[j] = from_row_ahead<j * ROUNDS>(lg[0]);
            });
        } else {
            sfor<5>([&](auto J) {
                constexpr int j = decltype(J)::value;
                const int L = j * ROUNDS + rr;
                double v = 0.0;
                sfor<NP>([&](auto K) {
                    constexpr int k = decltype(K)::value;
                    const double x = __shfl(lg[k], gbase + L % W, 64);
                    v = L / W == k ? x : v;
                });
                ln[j] = v;
            });
        }
        bool act[ROUNDS], tw[ROUNDS], kn2[ROUNDS];
        int gx0[ROUNDS], gy0[ROUNDS], gout[ROUNDS];
        uint32_t gh[ROUNDS];
        sfor<ROUNDS>([&](auto R) {
            constexpr int r = decltype(R)::value;
            act[r] = active[r];
            tw[r] = two[r];
            kn2[r] = at2[r];
            gx0[r] = geo[r].x0;
            gy0[r] = geo[r].y0;
            gout[r] = geo[r].out;
            gh[r] = geo[r].img_h;
        });
        if (t >= ROUNDS || !pick(act)) return;
        // the four stencil neighbours of the peak (near-tie test)
        const float us = pick(unscale);
        if constexpr (!SEQ) {
            nb[0] = stencil(rr, 0);
            nb[1] = stencil(rr, 2);
            nb[2] = stencil(rr, 3);
            nb[3] = stencil(rr, 4);
            sfor<4>([&](auto J) { nb2[decltype(J)::value] = around2(rr, decltype(J)::value); });
        }
        const float nmax = us * fmaxf(fmaxf(nb[0], nb[1]), fmaxf(nb[2], nb[3]));
        const float nmax2 = us * fmaxf(fmaxf(nb2[0], nb2[1]), fmaxf(nb2[2], nb2[3]));
        const float fv1 = us * pick(v1), fv2 = us * pick(v2);
        const int fpx = pick(px), fpy = pick(py);
        const int x0 = pick(gx0), y0 = pick(gy0);
        opiv_window_result res;
        res.reserved = 0;
        if (!pick(tw)) {
            // main.cpp:214-218: the default (all-zero) record stays
            res.x = res.y = res.u = res.v = res.sn = 0.0;
            res.peak0 = res.peak1 = 0.0;
            res.px = res.py = 0;
            res.flags = OPIV_FLAG_FEWER_THAN_2_PEAKS;
        } else {
            // image_utils.impl.h:86-98 on the logs of (l, c, r) and (up, c, down)
            const double lnl = ln[0], lnc = ln[1], lnr = ln[2], lnu = ln[3], lnd = ln[4];
            const double numx = lnl - lnr, denx = 2.0 * (lnl + lnr - 2.0 * lnc);
            const double numy = lnu - lnd, deny = 2.0 * (lnu + lnd - 2.0 * lnc);
            const double plx = (double)fpx + (denx == 0.0 ? 0.0 : numx / denx);
            const double ply = (double)fpy + (deny == 0.0 ? 0.0 : numy / deny);
            uint32_t flags = 0;
            if (!isfinite(plx) || !isfinite(ply)) flags |= OPIV_FLAG_NONFINITE_FIT;
            const float tol = p.tie_tol * fabsf(fv1);
            // the order of the top two, or either one's strict-maximum test,
            // within FP32 reach: the second peak's test decides whether the
            // window has two peaks at all
            if (fv1 - fv2 <= tol || fv1 - nmax <= tol || (pick(kn2) && fv2 - nmax2 <= tol))
                flags |= OPIV_FLAG_NEAR_TIE;
            if (fv1 == fv2) flags |= OPIV_FLAG_EXACT_TIE;
            // main.cpp:220-234
            const int midx = x0 + W / 2, midy = y0 + W / 2;
            res.x = (double)midx;
            res.y = (double)pick(gh) - (double)midy;
            res.u = (double)midx - ((double)x0 + plx);
            res.v = (double)midy - ((double)y0 + ply);
            res.sn = (fv2 > 0.f) ? (double)fv1 / (double)fv2 : 0.0;
            res.peak0 = (double)fv1;
            res.peak1 = (double)fv2;
            res.px = fpx;
            res.py = fpy;
            res.flags = flags;
        }
        p.out[pick(gout)] = res;
    }
}

} // namespace opiv
