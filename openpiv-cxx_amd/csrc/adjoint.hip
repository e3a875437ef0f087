// adjoint.hip -- vector-Jacobian products of each window's displacement
// (u, v) with respect to both frames' pixels, and the per-window variance of
// (u, v) under i.i.d. pixel noise.  No counterpart in the reference; it
// differentiates the reference's vector (examples/process/main.cpp:220-227).
//
// For a window with integer peak (px, py) (held fixed: the argmax has zero
// derivative almost everywhere), u = W/2 - px - dx and
// dx = (ln l - ln r) / (2 (ln l + ln r - 2 ln c)) (image_utils.impl.h:86-94),
// where l, c, r (and up, down for v) are circular correlations
// C[s] = sum_n a[n] b[n+s] at five fixed shifts s_k.  Hence
//   du/da[n] = sum_k du/dC_k * b[n + s_k],   du/db[m] = sum_k du/dC_k * a[m - s_k]
// -- the adjoint of a correlation is again a (5-tap) correlation.
//
// Mapping: W lanes own a window (lane t = column t, as in the forward
// kernel).  Both windows go to LDS; the taps of a lane touch three columns
// per frame, which it caches one at a time, rotated by the peak shift so
// every tap is a register read.  The five C_k are recomputed as direct dot
// products with FP64 accumulation, so the fit's derivatives use FP64 logs
// of accurate values.
//
// The VJP runs in two launches: the window kernel stops at the five tap
// weights (cotangent folded in), and a gather kernel with one thread per
// frame pixel sums, over the windows that cover the pixel, the weighted taps
// of the other frame (wrapped inside the window).  Each gradient element is
// written once: no atomics and no clearing pass.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "adjoint.hpp"
#include "fft_reg.hpp"
#include "xcd.hpp"

namespace opiv {

using fftreg::sfor;

namespace {

__device__ __forceinline__ void wave_sync_adj()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

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

// 1 / x to FP64 precision: the hardware reciprocal and two Newton steps
__device__ __forceinline__ double rcp64(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = fma(fma(-x, r, 1.0), r, r);
    return fma(fma(-x, r, 1.0), r, r);
}

// d(num/den)/d(l, c, r) with num = ln l - ln r, den = 2 (ln l + ln r - 2 ln c)
// (image_utils.impl.h:86-94), from the values and their logs; zero where den
// is zero (the fit's sub-pixel term is then 0)
__device__ __forceinline__ void gauss3_grad(double l, double c, double r, double lnl, double lnc, double lnr,
                                            double& gl, double& gc, double& gr)
{
    const double num = lnl - lnr, den = 2.0 * (lnl + lnr - 2.0 * lnc);
    const double q = den == 0.0 ? 0.0 : rcp64(den * den);
    gl = (den - 2.0 * num) * q * rcp64(l);
    gr = (-den - 2.0 * num) * q * rcp64(r);
    gc = 4.0 * num * q * rcp64(c);
}

} // namespace

template <int W, int DT, bool VARIANCE>
__global__ __launch_bounds__(256) void adjoint_kernel(AdjParams p)
{
    constexpr int WPW = 64 / W;
    constexpr int WAVES = 256 / 64;
    constexpr int H2 = W / 2;
    constexpr int M = W - 1;
    // frame B's window (both frames for the variance, whose gradient pass
    // also reads frame A's columns)
    constexpr int NF = VARIANCE ? 2 : 1;
    __shared__ float lds_all[WAVES][WPW][NF][W][W + 1];

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int slot = lane / W;
    const int t = lane % W;
    float (*lb)[W + 1] = lds_all[wave][slot][0];
    float (*la)[W + 1] = lds_all[wave][slot][NF - 1];

    const int local = (xcd_block((int)blockIdx.x, (int)gridDim.x) * WAVES + wave) * WPW + slot;
    const bool active = local < p.n_windows;
    const int i = active ? local : p.n_windows - 1;
    const int ix = i % p.nx, iy = i / p.nx;
    const int x0 = p.x_start + ix * p.step_x;
    const int y0 = p.y_start + iy * p.step_y;
    const opiv_window_result fw = p.fwd[i];
    const bool valid = active && !(fw.flags & (OPIV_FLAG_FEWER_THAN_2_PEAKS | OPIV_FLAG_NONFINITE_FIT));

    // peak shift in correlation coordinates: swapped (h, w) holds
    // C[(h + W/2) % W][(w + W/2) % W]
    const int sx0 = (fw.px + H2) & M, sy0 = (fw.py + H2) & M;

    // ---- both windows to LDS, column t of frame A kept in registers.  The
    // LDS copies are stored rotated by the peak's row shift, frame B as
    // lb[j][x] = b[(j + sy0 - 1) mod W][x] and frame A (read by the variance
    // gradients) as la[j][x] = a[(j - sy0 - 1) mod W][x], so every tap below
    // reads its rows at compile-time offsets. ----
    float acol[W];
    {
        const size_t idx = (size_t)y0 * p.pitch + (size_t)(x0 + t);
        const int rb = (1 - sy0) & M, ra = (sy0 + 1) & M;   // LDS rows of window row 0
        sfor<W>([&](auto Y) {
            constexpr int y = decltype(Y)::value;
            acol[y] = load_px<DT>(p.a, idx + (size_t)y * p.pitch);
            if constexpr (VARIANCE) la[(ra + y) & M][t] = acol[y];
            lb[(rb + y) & M][t] = load_px<DT>(p.b, idx + (size_t)y * p.pitch);
        });
    }
    wave_sync_adj();

    // taps k = l, c, r, up, down at shifts s_k = (sy0 + dy, sx0 + dx) with
    // (dy, dx) = (0,-1), (0,0), (0,1), (-1,0), (1,0).  The taps touch three
    // columns per frame; one column is cached at a time, rotated so that
    // col[j] = frame[(j + rotation) % W][column] and every tap row is a
    // compile-time register index.
    constexpr int TDY[5] = {0, 0, 0, -1, 1};
    constexpr int TDX[5] = {-1, 0, 1, 0, 0};

    // ---- the five stencil correlations C_k = sum_n a[n] b[n + s_k] (FP64 sums) ----
    double ck[5] = {0, 0, 0, 0, 0};
    sfor<3>([&](auto D) {
        constexpr int d = decltype(D)::value;          // dx = d - 1
        const int col = (t + sx0 + d - 1) & M;
        float bc[W + 2];                                // bc[j] = b[(j + sy0 - 1) % W][col]
        sfor<W + 2>([&](auto J) {
            constexpr int j = decltype(J)::value;
            bc[j] = lb[j & M][col];
        });
        sfor<5>([&](auto K) {
            constexpr int k = decltype(K)::value;
            if constexpr (TDX[k] == d - 1) {
                double s = ck[k];
                sfor<W>([&](auto Y) {
                    constexpr int y = decltype(Y)::value;
                    s = fma((double)acol[y], (double)bc[y + TDY[k] + 1], s);
                });
                ck[k] = s;
            }
        });
    });
#pragma unroll
    for (int off = W / 2; off >= 1; off >>= 1) {
        sfor<5>([&](auto K) {
            constexpr int k = decltype(K)::value;
            ck[k] += __shfl_xor(ck[k], off, 64);
        });
    }

    // ---- du/dC_k and dv/dC_k (u = W/2 - px - dx, v = W/2 - py - dy).  The
    // five logs are taken one per lane (lane k of the group takes ln C_k)
    // and shared through the group. ----
    double lnk[5];
    {
        double mine = ck[0];
        sfor<4>([&](auto K) {
            constexpr int k = decltype(K)::value + 1;
            mine = t == k ? ck[k] : mine;
        });
        const double lg = log(mine);
        const int gbase = lane - t;
        sfor<5>([&](auto K) {
            constexpr int k = decltype(K)::value;
            lnk[k] = __shfl(lg, gbase + k, 64);
        });
    }
    double gu[5] = {0, 0, 0, 0, 0}, gv[5] = {0, 0, 0, 0, 0};
    {
        double gl, gcx, gr, gt, gcy, gbt;
        gauss3_grad(ck[0], ck[1], ck[2], lnk[0], lnk[1], lnk[2], gl, gcx, gr);
        gauss3_grad(ck[3], ck[1], ck[4], lnk[3], lnk[1], lnk[4], gt, gcy, gbt);
        if (valid) {
            gu[0] = -gl;
            gu[1] = -gcx;
            gu[2] = -gr;
            gv[1] = -gcy;
            gv[3] = -gt;
            gv[4] = -gbt;
        }
    }
    float fu[5], fv[5];
    {
        const double cu = (!VARIANCE && active) ? p.cot[2 * (size_t)i] : 1.0;
        const double cv = (!VARIANCE && active) ? p.cot[2 * (size_t)i + 1] : 1.0;
        sfor<5>([&](auto K) {
            constexpr int k = decltype(K)::value;
            if constexpr (VARIANCE) {
                fu[k] = (float)gu[k];
                fv[k] = (float)gv[k];
            } else {
                fu[k] = (float)(cu * gu[k] + cv * gv[k]);   // the VJP folds both cotangents
                fv[k] = 0.f;
            }
        });
    }
    if constexpr (!VARIANCE) {
        if (active && t == 0) {
            AdjCoef c;
            sfor<5>([&](auto K) {
                constexpr int k = decltype(K)::value;
                c.f[k] = valid ? fu[k] : 0.f;
            });
            c.shift = sx0 | (sy0 << 16);
            c.pad_[0] = c.pad_[1] = 0;
            p.coef[local] = c;
        }
        return;
    }

    // ---- gradients: dC_k/da[n] = b[n + s_k], dC_k/db[m] = a[m - s_k] ----
    // g[0][y] (frame A) and g[1][y] (frame B) of column t, for u (or the
    // VJP) and, with VARIANCE, for v
    double su = 0.0, sv = 0.0;
    sfor<2>([&](auto F) {
        constexpr int fr = decltype(F)::value;       // 0: d/da uses frame B taps, 1: d/db uses frame A
        float g_u[W], g_v[W];
        sfor<W>([&](auto Y) {
            constexpr int y = decltype(Y)::value;
            g_u[y] = 0.f;
            g_v[y] = 0.f;
        });
        sfor<3>([&](auto D) {
            constexpr int d = decltype(D)::value;
            // frame B column t + sx0 + dx rotated by sy0 - 1, or frame A
            // column t - sx0 - dx rotated by -sy0 - 1 (the LDS copies are
            // stored with these rotations)
            const int col = fr == 0 ? ((t + sx0 + d - 1) & M) : ((t - sx0 - (d - 1)) & M);
            float (*src)[W + 1] = fr == 0 ? lb : la;
            float cc[W + 2];
            sfor<W + 2>([&](auto J) {
                constexpr int j = decltype(J)::value;
                cc[j] = src[j & M][col];
            });
            sfor<5>([&](auto K) {
                constexpr int k = decltype(K)::value;
                if constexpr (TDX[k] == d - 1) {
                    sfor<W>([&](auto Y) {
                        constexpr int y = decltype(Y)::value;
                        constexpr int row = fr == 0 ? (y + TDY[k] + 1) : (y - TDY[k] + 1);
                        g_u[y] = fmaf(fu[k], cc[row], g_u[y]);
                        if constexpr (VARIANCE) g_v[y] = fmaf(fv[k], cc[row], g_v[y]);
                    });
                }
            });
        });
        sfor<W>([&](auto Y) {
            constexpr int y = decltype(Y)::value;
            su = fma((double)g_u[y], (double)g_u[y], su);
            sv = fma((double)g_v[y], (double)g_v[y], sv);
        });
    });

#pragma unroll
    for (int off = W / 2; off >= 1; off >>= 1) {
        su += __shfl_xor(su, off, 64);
        sv += __shfl_xor(sv, off, 64);
    }
    if (active && t == 0) {
        const double s2 = p.sigma * p.sigma;
        p.var[2 * (size_t)local] = valid ? s2 * su : 0.0;
        p.var[2 * (size_t)local + 1] = valid ? s2 * sv : 0.0;
    }
}

// VJP gather: thread (x, y) of the gradient frames (x < gpitch, y < img_h)
// sums, over the windows (ix, iy) whose square covers it, the five weighted
// taps: grad_a[n] += f_k b[n + s_k], grad_b[n] += f_k a[n - s_k], positions
// wrapped inside the window.  Pixels outside every window get zero.
template <int W, int DT>
__global__ __launch_bounds__(256) void adjoint_gather_kernel(AdjParams p)
{
    constexpr int M = W - 1;
    constexpr int TDY[5] = {0, 0, 0, -1, 1};
    constexpr int TDX[5] = {-1, 0, 1, 0, 0};
    const int x = blockIdx.x * 64 + (int)(threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (int)(threadIdx.x >> 6);
    if (x >= (int)p.gpitch || y >= p.img_h) return;

    // windows ix with x0(ix) <= x < x0(ix) + W, likewise iy
    const int dx = x - p.x_start, dy = y - p.y_start;
    const int ix_lo = dx < W ? 0 : (dx - W) / p.step_x + 1;
    const int ix_hi = dx < 0 ? -1 : min(dx / p.step_x, p.nx - 1);
    const int iy_lo = dy < W ? 0 : (dy - W) / p.step_y + 1;
    const int iy_hi = dy < 0 ? -1 : min(dy / p.step_y, p.ny - 1);

    float ga = 0.f, gb = 0.f;
    for (int iy = iy_lo; iy <= iy_hi; ++iy) {
        const int y0 = p.y_start + iy * p.step_y;
        const int ly = y - y0;
        for (int ix = ix_lo; ix <= ix_hi; ++ix) {
            const AdjCoef c = p.coef[(size_t)iy * p.nx + ix];
            const int x0 = p.x_start + ix * p.step_x;
            const int lx = x - x0;
            const int sx0 = c.shift & 0xffff, sy0 = c.shift >> 16;
            sfor<5>([&](auto K) {
                constexpr int k = decltype(K)::value;
                const int sy = sy0 + TDY[k], sx = sx0 + TDX[k];
                const size_t ib = (size_t)(y0 + ((ly + sy) & M)) * p.pitch + (size_t)(x0 + ((lx + sx) & M));
                const size_t ia = (size_t)(y0 + ((ly - sy) & M)) * p.pitch + (size_t)(x0 + ((lx - sx) & M));
                ga = fmaf(c.f[k], load_px<DT>(p.b, ib), ga);
                gb = fmaf(c.f[k], load_px<DT>(p.a, ia), gb);
            });
        }
    }
    p.grad_a[(size_t)y * p.gpitch + x] = ga;
    p.grad_b[(size_t)y * p.gpitch + x] = gb;
}

// VJP gather through LDS (W <= 32): block = a 64 x 64 tile of gradient
// pixels.  The windows overlapping the tile span a region of at most
// (62 + 2W)^2 pixels; it is staged in LDS one frame at a time (frame B for
// grad_a, then frame A for grad_b), so every tap read is an LDS read, and
// the tap weights of those windows (at most kTileWin) are staged once.  Same
// sums, in the same order, as adjoint_gather_kernel.
constexpr int kTile = 64;
constexpr int kTileWin = 256;
template <int W>
constexpr int region_side() { return kTile - 2 + 2 * W; }
// windows of one axis that can overlap a tile
inline int tile_windows(int win, int step) { return (kTile + win - 2) / step + 1; }

template <int W, int DT>
__global__ __launch_bounds__(1024) void adjoint_gather_tiled_kernel(AdjParams p)
{
    constexpr int M = W - 1;
    constexpr int RS = region_side<W>();
    constexpr int R = kTile / 16;               // rows per thread: column c, rows rg * R .. + R - 1
    __shared__ float reg[RS * RS];
    __shared__ AdjCoef wc[kTileWin];

    const int tx0 = blockIdx.x * kTile, ty0 = blockIdx.y * kTile;
    // windows overlapping the tile: x0 in (tx0 - W, tx0 + kTile), likewise y0
    const int ex = tx0 - p.x_start - W + 1, ey = ty0 - p.y_start - W + 1;
    const int ixa = ex <= 0 ? 0 : (ex + p.step_x - 1) / p.step_x;
    const int iya = ey <= 0 ? 0 : (ey + p.step_y - 1) / p.step_y;
    const int fx = tx0 + kTile - 1 - p.x_start, fy = ty0 + kTile - 1 - p.y_start;
    const int ixb = fx < 0 ? -1 : min(fx / p.step_x, p.nx - 1);
    const int iyb = fy < 0 ? -1 : min(fy / p.step_y, p.ny - 1);
    const bool any = ixa <= ixb && iya <= iyb;
    const int rx0 = p.x_start + ixa * p.step_x, ry0 = p.y_start + iya * p.step_y;
    const int rw = any ? (ixb - ixa) * p.step_x + W : 0;
    const int rh = any ? (iyb - iya) * p.step_y + W : 0;

    // tap weights of the tile's windows, wc[(iy - iya) * nwx + ix - ixa]
    const int nwx = any ? ixb - ixa + 1 : 0, nwy = any ? iyb - iya + 1 : 0;
    for (int i = (int)threadIdx.x; i < nwx * nwy; i += 1024) {
        const int wy = i / nwx, wx = i - wy * nwx;
        wc[i] = p.coef[(size_t)(iya + wy) * p.nx + ixa + wx];
    }

    // rg is the wave index, so every row test below is wave-uniform
    const int c = (int)(threadIdx.x & 63), rg = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int x = tx0 + c, ya = ty0 + rg * R;
    const int dx = x - p.x_start;
    const int ix_lo = dx < W ? 0 : (dx - W) / p.step_x + 1;
    const int ix_hi = dx < 0 ? -1 : min(dx / p.step_x, p.nx - 1);
    // windows iy covering any of rows ya .. ya + R - 1
    const int dya = ya - p.y_start, dyb = ya + R - 1 - p.y_start;
    const int iy_lo = dya < W ? 0 : (dya - W) / p.step_y + 1;
    const int iy_hi = dyb < 0 ? -1 : min(dyb / p.step_y, p.ny - 1);
    // staging: column sc of the region, rows sr, sr + 8, ...
    const int sc = (int)(threadIdx.x & 127), sr = (int)(threadIdx.x >> 7);

    sfor<2>([&](auto F) {
        constexpr int fr = decltype(F)::value;       // 0: grad_a from frame B, 1: grad_b from frame A
        const void* src = fr == 0 ? p.b : p.a;
        float* dst = fr == 0 ? p.grad_a : p.grad_b;
        __syncthreads();                             // previous frame's readers are done
        if (sc < rw) {
            const size_t base = (size_t)ry0 * p.pitch + (size_t)(rx0 + sc);
            for (int r = sr; r < rh; r += 32) {
                float v[4];
                sfor<4>([&](auto U) {
                    constexpr int u = decltype(U)::value;
                    const int rr = r + 8 * u;
                    v[u] = rr < rh ? load_px<DT>(src, base + (size_t)rr * p.pitch) : 0.f;
                });
                sfor<4>([&](auto U) {
                    constexpr int u = decltype(U)::value;
                    const int rr = r + 8 * u;
                    if (rr < rh) reg[rr * RS + sc] = v[u];
                });
            }
        }
        __syncthreads();

        float g[R];
        sfor<R>([&](auto I) { g[decltype(I)::value] = 0.f; });
        for (int iy = iy_lo; iy <= iy_hi; ++iy) {
            const int y0 = p.y_start + iy * p.step_y;
            const int ly0 = ya - y0;                 // local row of the thread's first row
            for (int ix = ix_lo; ix <= ix_hi; ++ix) {
                const AdjCoef cf = wc[(iy - iya) * nwx + ix - ixa];
                const int x0 = p.x_start + ix * p.step_x;
                const int lx = x - x0;
                const int sx0 = cf.shift & 0xffff, sy0 = cf.shift >> 16;
                // columns of the taps l, c, r (up and down use the centre one)
                const int cx = fr == 0 ? ((lx + sx0) & M) : ((lx - sx0) & M);
                const int cl = (cx + (fr == 0 ? -1 : 1)) & M, cr = (cx + (fr == 0 ? 1 : -1)) & M;
                const int ob = (y0 - ry0) * RS + (x0 - rx0);
                // tap rows around the thread's rows: row e (e = 0..R+1) of the
                // window is (ly0 + e - 1 + sy0) mod W (frame B) or with -sy0
                // (frame A); the thread's row i reads l, c, r on row i + 1 and
                // up/down on rows i and i + 2 (swapped for frame A, whose
                // shift is negated)
                const int rb = ly0 - 1 + (fr == 0 ? sy0 : -sy0);
                int rowc[R + 2];
                float vc[R + 2];
                sfor<R + 2>([&](auto E) {
                    constexpr int e = decltype(E)::value;
                    rowc[e] = ob + ((rb + e) & M) * RS;
                    vc[e] = reg[rowc[e] + cx];
                });
                sfor<R>([&](auto I) {
                    constexpr int i = decltype(I)::value;
                    const int ly = ly0 + i;
                    if (ly >= 0 && ly < W) {
                        float t = g[i];
                        t = fmaf(cf.f[0], reg[rowc[i + 1] + cl], t);
                        t = fmaf(cf.f[1], vc[i + 1], t);
                        t = fmaf(cf.f[2], reg[rowc[i + 1] + cr], t);
                        t = fmaf(cf.f[3], vc[fr == 0 ? i : i + 2], t);
                        t = fmaf(cf.f[4], vc[fr == 0 ? i + 2 : i], t);
                        g[i] = t;
                    }
                });
            }
        }
        if (x < (int)p.gpitch) {
            sfor<R>([&](auto I) {
                constexpr int i = decltype(I)::value;
                const int y = ya + i;
                if (y < p.img_h) dst[(size_t)y * p.gpitch + x] = g[i];
            });
        }
    });
}

template <int W, int DT>
static hipError_t launch_adj_w(const AdjParams& p, bool variance, hipStream_t stream)
{
    constexpr int per_block = 4 * (64 / W);
    const int blocks = (p.n_windows + per_block - 1) / per_block;
    if (variance) {
        hipLaunchKernelGGL((adjoint_kernel<W, DT, true>), dim3(blocks), dim3(256), 0, stream, p);
        return hipGetLastError();
    }
    hipLaunchKernelGGL((adjoint_kernel<W, DT, false>), dim3(blocks), dim3(256), 0, stream, p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int twx = tile_windows(W, p.step_x), twy = tile_windows(W, p.step_y);
    if (region_side<W>() * region_side<W>() * 4 <= 65536 && twx * twy <= kTileWin) {
        const dim3 grid((unsigned)((p.gpitch + kTile - 1) / kTile), (unsigned)((p.img_h + kTile - 1) / kTile));
        hipLaunchKernelGGL((adjoint_gather_tiled_kernel<W, DT>), grid, dim3(1024), 0, stream, p);
    } else {
        const dim3 grid((unsigned)((p.gpitch + 63) / 64), (unsigned)((p.img_h + 3) / 4));
        hipLaunchKernelGGL((adjoint_gather_kernel<W, DT>), grid, dim3(256), 0, stream, p);
    }
    return hipGetLastError();
}

template <int DT>
static hipError_t launch_adj_dt(const AdjParams& p, int win, bool variance, hipStream_t stream)
{
    switch (win) {
    case 8: return launch_adj_w<8, DT>(p, variance, stream);
    case 16: return launch_adj_w<16, DT>(p, variance, stream);
    case 32: return launch_adj_w<32, DT>(p, variance, stream);
    case 64: return launch_adj_w<64, DT>(p, variance, stream);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_adjoint(const AdjParams& p, int win, int dtype, bool variance, hipStream_t stream)
{
    if (p.n_windows <= 0) return hipSuccess;
    if (!variance && (p.coef == nullptr || p.nx <= 0 || p.ny <= 0 || p.step_x <= 0 || p.step_y <= 0 ||
                      p.img_h <= 0 || (long long)p.nx * p.ny != p.n_windows))
        return hipErrorInvalidValue;
    switch (dtype) {
    case 0: return launch_adj_dt<0>(p, win, variance, stream);
    case 1: return launch_adj_dt<1>(p, win, variance, stream);
    case 2: return launch_adj_dt<2>(p, win, variance, stream);
    case 3: return launch_adj_dt<3>(p, win, variance, stream);
    default: return hipErrorInvalidValue;
    }
}

} // namespace opiv
