"""The FP64 adjoint of the oracle, pinned by FP64 central finite differences
of the oracle's own forward window processing (step 1e-3 on pixel
intensities), by linearity, and by the closed form of the Gaussian fit."""
import numpy as np
import pytest

import openpiv_amd as oa
import oracle as orc

H = 1e-3


def windows(n, win=32, seed=5):
    a, b = oa.synth_pair(256, 256, seed=seed, params=(1.6, -0.8))
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        x0, y0 = rng.integers(0, 256 - win, 2)
        out.append((a[y0:y0 + win, x0:x0 + win].copy(), b[y0:y0 + win, x0:x0 + win].copy()))
    return out


@pytest.mark.parametrize("seed", [5, 6])
def test_jacobian_matches_central_differences(seed):
    rng = np.random.default_rng(seed + 100)
    checked = 0
    for wa, wb in windows(16, seed=seed):
        j = orc.window_jacobian(wa, wb)
        if j is None:
            continue
        base = orc.window_vector(wa, wb)
        for _ in range(6):
            y, x = rng.integers(0, 32, 2)
            for frame in (0, 1):
                plus, minus = [wa.copy(), wb.copy()], [wa.copy(), wb.copy()]
                plus[frame][y, x] += H
                minus[frame][y, x] -= H
                vp, vm = orc.window_vector(*plus), orc.window_vector(*minus)
                if (vp["px"], vp["py"], vm["px"], vm["py"]) != (base["px"], base["py"], base["px"], base["py"]):
                    continue                      # the integer peak moved inside the step
                fd_u = (vp["u"] - vm["u"]) / (2 * H)
                fd_v = (vp["v"] - vm["v"]) / (2 * H)
                an_u, an_v = j[frame][y, x], j[2 + frame][y, x]
                scale = np.abs(j[frame]).max() + np.abs(j[2 + frame]).max()
                assert abs(fd_u - an_u) <= 1e-5 * scale + 1e-6 * abs(an_u), (frame, y, x, fd_u, an_u)
                assert abs(fd_v - an_v) <= 1e-5 * scale + 1e-6 * abs(an_v), (frame, y, x, fd_v, an_v)
                checked += 1
    assert checked > 100


def test_jacobian_invariances():
    """The fit depends on logarithmic ratios only: scaling frame A by k leaves
    (u, v) unchanged, so sum_n a[n] du/da[n] = 0 (Euler), likewise for B."""
    for wa, wb in windows(6, seed=9):
        j = orc.window_jacobian(wa, wb)
        if j is None:
            continue
        for frame, img in ((0, wa), (1, wb)):
            for comp in (0, 2):
                s = (j[comp + frame] * img).sum()
                assert abs(s) <= 1e-8 * np.abs(j[comp + frame] * img).sum()


def test_vjp_and_variance_assemble_the_jacobians():
    a, b = oa.synth_pair(96, 96, seed=3, params=(1.6, -0.8))
    _, xy = orc.grid(96, 96, 32, 32, 16, 16)
    rng = np.random.default_rng(1)
    cot = rng.normal(size=(len(xy), 2))
    ga, gb = orc.vjp_pair(a, b, 32, 16, 16, cot)
    ea, eb = np.zeros_like(ga), np.zeros_like(gb)
    var = orc.variance_pair(a, b, 32, 16, 16, 2.0)
    for i, (x0, y0) in enumerate(xy):
        j = orc.window_jacobian(a[y0:y0 + 32, x0:x0 + 32], b[y0:y0 + 32, x0:x0 + 32])
        if j is None:
            assert (var[i] == 0).all()
            continue
        ea[y0:y0 + 32, x0:x0 + 32] += cot[i, 0] * j[0] + cot[i, 1] * j[2]
        eb[y0:y0 + 32, x0:x0 + 32] += cot[i, 0] * j[1] + cot[i, 1] * j[3]
        assert np.isclose(var[i, 0], 4.0 * ((j[0] ** 2).sum() + (j[1] ** 2).sum()), rtol=1e-12)
        assert np.isclose(var[i, 1], 4.0 * ((j[2] ** 2).sum() + (j[3] ** 2).sum()), rtol=1e-12)
    np.testing.assert_allclose(ga, ea, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(gb, eb, rtol=1e-12, atol=1e-15)
