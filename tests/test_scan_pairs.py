"""The correlation kernels' peak scan, restated over numpy planes and checked
against the oracle's find_peaks(plane, 2, 1) (image_utils.impl.h:37-74).

Each lane of a kernel holds one column of the quadrant-swapped plane and
scans its rows in pairs: of two vertically adjacent elements at most one is
a strict 4-neighbour maximum, so a pair offers one candidate to the lane's
top 2 and the lane keeps only the pair's number; the row comes back
afterwards from the plane (the pair's first row holds the best exactly when
its value equals the best).  The lanes' results then merge into the group's
top 2 (largest value, smallest scan index among equal values).  This test
runs that procedure on tie-rich integer planes and on continuous planes and
compares the first peak's position and value and the second peak's value
with the oracle.  It pins the procedure; the GPU parity tests pin the
device code against the oracle on real correlation planes.
"""
import numpy as np
import pytest

import oracle as orc


def kernel_scan(plane, limit_search=False):
    """(v1, px, py, v2) as the kernels' scan and group merge compute them,
    or None when the region holds fewer than two strict maxima."""
    h, w = plane.shape
    r0 = w // 4 if limit_search else 0
    rlen = w // 2 if limit_search else w
    hlo, hhi = r0 + 1, r0 + rlen - 3
    nr = w - 3                                   # rows 1 .. w-3
    lanes = []
    for t in range(w):
        col_ok = r0 + 1 <= t <= r0 + rlen - 2
        tc = min(max(t, 1), w - 2)               # (edge lanes scan nothing)
        v1, v2, p1 = -np.inf, -np.inf, 0

        def nmax(hh):
            return max(plane[hh, tc - 1], plane[hh, tc + 1], plane[hh - 1, t], plane[hh + 1, t])

        def offer(mc, pair):
            nonlocal v1, v2, p1
            gt = mc > v1
            v2 = max(v2, min(mc, v1))            # med3(v2, mc, v1)
            if gt:
                p1, v1 = pair, mc

        for pair in range(nr // 2):
            hh = 2 * pair + 1
            c0, c1 = plane[hh, t], plane[hh + 1, t]
            m0 = c0 > nmax(hh) and hlo <= hh <= hhi
            m1 = c1 > nmax(hh + 1) and hlo <= hh + 1 <= hhi
            offer(c0 if m0 else (c1 if m1 else -np.inf), pair)
        if nr % 2:
            hh = nr
            c = plane[hh, t]
            offer(c if (c > nmax(hh) and hlo <= hh <= hhi) else -np.inf, nr // 2)
        if not col_ok:
            v1, v2 = -np.inf, -np.inf
        hp = 2 * p1 + 1
        h1 = hp if plane[hp, t] == v1 else hp + 1
        lanes.append((v1, h1 * w + t, v2))
    gv1 = max(l[0] for l in lanes)
    i1 = min(l[1] for l in lanes if l[0] == gv1)
    cands = [(l[2] if l[1] == i1 else l[0]) for l in lanes]
    gv2 = max(cands)
    if not gv2 > -np.inf:
        return None
    return gv1, i1 % w, i1 // w, gv2


def strict_maxima(plane, limit_search=False):
    """Number of strict 4-neighbour maxima in find_peaks' scan range
    (rows [1, H-2), columns [1, W-1) of the plane or the view)."""
    h, w = plane.shape
    r0, rlen = (w // 4, w // 2) if limit_search else (0, w)
    n = 0
    for y in range(r0 + 1, r0 + rlen - 2):
        for x in range(r0 + 1, r0 + rlen - 1):
            c = plane[y, x]
            n += c > max(plane[y - 1, x], plane[y + 1, x], plane[y, x - 1], plane[y, x + 1])
    return n


def oracle_top2(plane, limit_search=False):
    h, w = plane.shape
    region = (w // 4, w // 4, w // 2, w // 2) if limit_search else None
    peaks, tie = orc.find_peaks(plane, 2, 1, region)
    if len(peaks) < 2:
        return None, tie
    (x1, y1, v1), (_, _, v2) = peaks[0], peaks[1]
    return (v1, x1, y1, v2), tie


@pytest.mark.parametrize("w", [8, 16, 32, 64])
@pytest.mark.parametrize("limit_search", [False, True])
@pytest.mark.parametrize("kind", ["levels4", "levels16", "continuous", "plateaus"])
def test_paired_scan_matches_find_peaks(w, limit_search, kind):
    rng = np.random.default_rng(w * 7 + (3 if limit_search else 0) + len(kind))
    checked = ties = 0
    for _ in range(150 if w <= 16 else 60):
        if kind == "levels4":
            plane = rng.integers(0, 4, size=(w, w)).astype(np.float64)
        elif kind == "levels16":
            plane = rng.integers(0, 16, size=(w, w)).astype(np.float64)
        elif kind == "continuous":
            plane = rng.normal(size=(w, w))
        else:                                     # 2x2 blocks of equal values plus sparse spikes
            plane = np.kron(rng.integers(0, 8, size=(w // 2, w // 2)), np.ones((2, 2)))
            spikes = rng.random((w, w)) < 0.05
            plane = plane + spikes * rng.integers(1, 4, size=(w, w))
        got = kernel_scan(plane, limit_search)
        ref, tie = oracle_top2(plane, limit_search)
        assert (got is None) == (ref is None), (got, ref)
        if got is None:
            continue
        v1, px, py, v2 = got
        assert v1 == ref[0] and v2 == ref[3], (got, ref)
        if v1 == v2:
            # equal top values: the kernels flag the window (EXACT_TIE) and
            # report the first in scan order, which is the reference's order
            # while its sort of the maxima is an insertion sort (16 or fewer)
            ties += 1
            if strict_maxima(plane, limit_search) > 16:
                continue
        assert (px, py) == (ref[1], ref[2]), (got, ref)
        checked += 1
    # tie-rich planes mostly hold equal top values (compared by value above);
    # continuous ones check positions, except the 8 x 8 limit-search view,
    # which never holds two maxima
    if kind == "continuous" and not (w == 8 and limit_search):
        assert checked > 20
