"""The FP64 oracle pinned by mathematics and by the reference's own unit
tests (tests/image_algos_test.cpp, image_utils_test.cpp, grid_test.cpp),
independently of the compiled reference."""
import math

import numpy as np
import pytest

import oracle as orc


def dft2(z, sign):
    h, w = z.shape
    ky = np.exp(sign * 2j * np.pi * np.outer(np.arange(h), np.arange(h)) / h)
    kx = np.exp(sign * 2j * np.pi * np.outer(np.arange(w), np.arange(w)) / w)
    return ky @ z @ kx


@pytest.mark.parametrize("n", [4, 8, 16])
def test_transform_is_the_dft(n):
    rng = np.random.default_rng(n)
    z = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    f = orc.FFT(n)
    np.testing.assert_allclose(f.transform(z, True), dft2(z, -1), rtol=0, atol=1e-10 * n * n)
    np.testing.assert_allclose(f.transform(z, False), dft2(z, +1), rtol=0, atol=1e-10 * n * n)


def test_sinusoid_kat():
    """image_algos_test.cpp:23-62: 128 sin(2 pi (x + y)/8) has bins only at
    (32,32) and (224,224)."""
    y, x = np.mgrid[0:256, 0:256]
    im = 128 * np.sin(2 * np.pi * x / 8 + 2 * np.pi * y / 8)
    out = orc.FFT(256).transform(im.astype(np.complex128), True)
    mag = np.abs(out) ** 2
    mask = np.ones_like(mag, bool)
    mask[32, 32] = mask[224, 224] = False
    assert mag[32, 32] > 1 and mag[224, 224] > 1
    assert mag[mask].max() < 1e-9


def test_two_real_images_kat():
    """image_algos_test.cpp:64-109."""
    y, x = np.mgrid[0:256, 0:256]
    a = 128 * np.sin(2 * np.pi * x / 8 + 2 * np.pi * y / 8)
    b = 128 * np.sin(2 * np.pi * (256 - x) / 8 + 2 * np.pi * y / 8)
    fa, fb = orc.FFT(256).transform_real(a, b)
    for im, pts in ((fa, [(32, 32), (224, 224)]), (fb, [(32, 224), (224, 32)])):
        mag = np.abs(im) ** 2
        mask = np.ones_like(mag, bool)
        for px, py in pts:
            assert mag[py, px] > 1
            mask[py, px] = False
        assert mag[mask].max() < 1e-9


def test_non_pow2_rejected():
    """image_algos_test.cpp:122-130."""
    with pytest.raises(RuntimeError, match="power of 2"):
        orc.FFT(512, 400)
    with pytest.raises(RuntimeError, match="power of 2"):
        orc.FFT(400, 512)


def test_swap_quadrants():
    """image_utils_test.cpp:208-244 (quadrant sums 1,2,4,8 -> 8,4,2,1)."""
    i = np.arange(100 * 100)
    x = np.where(i % 100 < 50, 1, 2)
    y = np.where(i // 100 < 50, 1, 4)
    im = (x * y).reshape(100, 100).astype(np.float64)
    out = orc.swap_quadrants(im)
    q = lambda m, r, c: m[r:r + 50, c:c + 50].sum()  # noqa: E731
    assert (q(im, 0, 0), q(im, 0, 50), q(im, 50, 0), q(im, 50, 50)) == (2500, 5000, 10000, 20000)
    assert (q(out, 0, 0), q(out, 0, 50), q(out, 50, 0), q(out, 50, 50)) == (20000, 10000, 5000, 2500)


def test_find_peaks_in_order():
    """image_utils_test.cpp:246-267 and the empty case :307-314."""
    im = np.zeros((100, 100))
    for v in (20, 30, 40, 50):
        im[v, v] = float(v)
    peaks, tie = orc.find_peaks(im, 3, 1)
    assert peaks == [(50, 50, 50.0), (40, 40, 40.0), (30, 30, 30.0)] and not tie
    assert orc.find_peaks(np.zeros((100, 100)), 3, 1)[0] == []


def test_find_peaks_view_region():
    """A view (image_view) keeps global coordinates: a region starting at 25
    excludes the (20,20) peak and reports the rest at plane positions."""
    im = np.zeros((100, 100))
    for v in (20, 30, 40, 50):
        im[v, v] = float(v)
    peaks, _ = orc.find_peaks(im, 4, 1, region=(25, 25, 50, 50))
    assert peaks == [(50, 50, 50.0), (40, 40, 40.0), (30, 30, 30.0)]


def test_find_peaks_row_range():
    """Rows h in [1, H-2) are scanned: a maximum on row H-2 is never found."""
    im = np.zeros((16, 16))
    im[14, 5] = 9.0   # row H-2
    im[13, 9] = 3.0   # row H-3
    im[5, 14] = 4.0   # column W-2
    im[5, 15] = 7.0   # column W-1 (never a centre)
    peaks, _ = orc.find_peaks(im, 5, 1)
    assert peaks == [(9, 13, 3.0)]


def brute_circular(a, b):
    """C[s] = sum_n a[n] b[n + s] (circular), quadrant-swapped."""
    h, w = a.shape
    c = np.zeros((h, w))
    for sy in range(h):
        for sx in range(w):
            c[sy, sx] = (a * np.roll(np.roll(b, -sy, 0), -sx, 1)).sum()
    return np.fft.fftshift(c)


@pytest.mark.parametrize("n", [8, 16])
def test_cross_correlate_is_circular_correlation(n):
    rng = np.random.default_rng(n + 3)
    a, b = rng.random((n, n)), rng.random((n, n))
    # un-normalised inverse: the reference's plane is n*n times the correlation
    np.testing.assert_allclose(orc.FFT(n).cross_correlate(a, b), n * n * brute_circular(a, b), rtol=1e-11)
    np.testing.assert_allclose(orc.FFT(n).auto_correlate(a), n * n * brute_circular(a, a), rtol=1e-11)


def test_real_variant_zeroes_bins():
    """fft.h:166-183: cross_correlate_real equals the exact correlation of
    spectra with row 0, row H/2 and column 0 removed."""
    n = 16
    rng = np.random.default_rng(5)
    a, b = rng.random((n, n)), rng.random((n, n))
    A, B = np.fft.fft2(a), np.fft.fft2(b)
    P = B * np.conj(A)
    P[0, :] = 0
    P[n // 2, :] = 0
    P[:, 0] = 0
    expect = np.fft.fftshift(np.real(np.fft.ifft2(P))) * n * n
    np.testing.assert_allclose(orc.FFT(n).cross_correlate_real(a, b), expect, rtol=0, atol=1e-9 * np.abs(expect).max())


@pytest.mark.parametrize("d", [(0, 0), (3, -2), (-5, 4)])
def test_shift_theorem(d):
    """b = a circularly shifted by d: the peak is at centre + d, |fit| ~ 0."""
    n = 32
    rng = np.random.default_rng(9)
    a = rng.random((n, n)) ** 4
    b = np.roll(np.roll(a, d[1], 0), d[0], 1)
    plane = orc.FFT(n).cross_correlate(a, b)
    peaks, _ = orc.find_peaks(plane, 2, 1)
    assert peaks[0][:2] == (n // 2 + d[0], n // 2 + d[1])
    fx, fy = orc.fit_simple_gaussian(plane, peaks[0][0], peaks[0][1])
    assert abs(fx - (n // 2 + d[0])) < 1e-9 and abs(fy - (n // 2 + d[1])) < 1e-9


def test_gaussian_fit_exact_on_gaussians():
    """The 3-point estimator is exact for sampled Gaussians."""
    y, x = np.mgrid[0:9, 0:9].astype(np.float64)
    for x0, y0, s in [(4.3, 3.8, 1.3), (3.6, 4.45, 0.9), (4.0, 4.0, 2.0)]:
        g = 1000 * np.exp(-((x - x0) ** 2 + (y - y0) ** 2) / (2 * s * s))
        fx, fy = orc.fit_simple_gaussian(g, 4, 4)
        assert abs(fx - x0) < 1e-12 and abs(fy - y0) < 1e-12


def test_process_vector_convention():
    """main.cpp:220-234: u = W/2 - (px + dx), v = H/2 - (py + dy),
    x = mid_x, y = H_img - mid_y; frame B moved by (+2.3, -1.7) gives
    u ~ -2.3, v ~ +1.7."""
    import openpiv_amd as oa
    a, b = oa.synth_pair(256, 256, seed=3, params=(2.3, -1.7))
    out = orc.process_pair(a, b, 32, 16, 16, threads=2)
    _, xy = orc.grid(256, 256, 32, 32, 16, 16)
    good = (out["flags"] & 1) == 0
    assert abs(np.median(out["u"][good]) + 2.3) < 0.1 and abs(np.median(out["v"][good]) - 1.7) < 0.1
    assert np.array_equal(out["x"][good], (xy[good, 0] + 16).astype(float))
    assert np.array_equal(out["y"][good], (256 - (xy[good, 1] + 16)).astype(float))


def test_grid_closed_form():
    rng = np.random.default_rng(2)
    for _ in range(200):
        iw, ih = rng.integers(8, 400, 2)
        ww, wh = rng.integers(1, 64, 2)
        ox, oy = rng.integers(1, 40, 2)
        rc, xy = orc.grid(int(iw), int(ih), int(ww), int(wh), int(ox), int(oy))
        if ww > iw or wh > ih:
            assert rc == -4
            continue
        nx, ny = 1 + (iw - ww) // ox, 1 + (ih - wh) // oy
        assert len(xy) == nx * ny
        # centred: equal margins up to one pixel, all windows inside
        assert xy[:, 0].min() >= 0 and xy[:, 0].max() + ww <= iw
        assert abs(int(xy[:, 0].min()) - int(iw - (xy[:, 0].max() + ww))) <= 1
        assert (iw - ww - ox * (nx - 1)) // 2 == xy[0, 0]
    assert orc.grid_offsets(32, 32, 0.5)[1:] == (16, 16)
    assert orc.grid_offsets(32, 32, 0.75)[1:] == (24, 24)
    assert orc.grid_offsets(32, 32, 1.01)[0] == -1
    assert math.isclose(0.5 * 32, 16)
