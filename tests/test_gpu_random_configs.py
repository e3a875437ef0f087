"""Seeded random plans against the FP64 oracle.

Each case draws a window size (8-128), a pixel type, a frame size that is
not a multiple of anything in particular, a row pitch with junk in the
padding, a grid step (the reference's fraction, or an explicit step from 1
to beyond the window), a variant (circular FFT, the reference's
cross_correlate_real bins, AUTO, the direct MFMA path where it exists) and
limit-search, then holds every window of the pair to the same bar as the
fixed-shape parity tests (examples/process/main.cpp:189-237 semantics,
grid.impl.h:10-61 windows, find_peaks and the Gaussian fit of
image_utils.impl.h:37-101).  The cases are fixed by their seeds, so a
failure names a reproducible configuration.
"""
import numpy as np
import pytest

import openpiv_amd as oa
import oracle as orc
from test_gpu_parity import compare, say

pytestmark = pytest.mark.gpu

WINDOWS = (8, 16, 32, 64, 128)
DTYPES = (np.uint8, np.uint16, np.float32, np.float64)
N_CASES = 160


def direct_supported(win, dtype):
    """the direct (MFMA) path: 16 x 16 for every pixel type, 32 x 32 for
    8- and 16-bit frames"""
    return win == 16 or (win == 32 and np.dtype(dtype).itemsize <= 2)


def draw_case(seed):
    rng = np.random.default_rng(1000 + seed)
    win = int(rng.choice(WINDOWS, p=(0.2, 0.25, 0.25, 0.2, 0.1)))
    dtype = DTYPES[int(rng.integers(len(DTYPES)))]
    w = int(rng.integers(win, win + (700 if win == 128 else 420)))
    h = int(rng.integers(win, win + (600 if win == 128 else 360)))
    pitch = w if rng.random() < 0.5 else w + int(rng.integers(1, 40))
    if rng.random() < 0.5:
        overlap = float(rng.choice((0.25, 0.5, 0.75, 1.0)))
        step = None
        sx = sy = max(1, int(win * overlap))
    else:
        overlap = 0.5
        sx = sy = int(rng.integers(max(1, win // 8), win + win // 2 + 1))
        step = (sx, sy)
    variants = [oa.CIRCULAR_FFT_FP32, oa.REAL_COMPAT, oa.AUTO]
    if direct_supported(win, dtype):
        variants.append(oa.DIRECT_BF16_SPLIT)
    variant = variants[seed % len(variants)]          # every variant in turn
    limit = bool(rng.random() < 0.3)
    shift = tuple(float(v) for v in rng.uniform(-win / 6, win / 6, size=2))
    return dict(win=win, dtype=dtype, w=w, h=h, pitch=pitch, overlap=overlap, step=step, sx=sx, sy=sy,
                variant=variant, limit=limit, shift=shift, seed=seed)


def frames_for(c):
    dt = np.dtype(c["dtype"])
    if dt == np.uint8:
        i0, noise = (30.0, 250.0), 3.0
    elif dt == np.uint16:
        i0, noise = (1000.0, 60000.0), 50.0
    else:
        i0, noise = (100.0, 255.0), 2.0
    return oa.synth_pair(c["w"], c["h"], seed=50 + c["seed"], i0=i0, noise=noise, params=c["shift"], dtype=dt)


def padded(x, pitch, rng):
    """the frame in the first columns of rows `pitch` pixels apart; the
    padding holds junk that no window may read"""
    h, w = x.shape
    out = np.empty((h, pitch), dtype=x.dtype)
    out[:, :w] = x
    if pitch > w:
        junk = rng.uniform(0, 250, size=(h, pitch - w))
        out[:, w:] = junk.astype(x.dtype)
    return out


@pytest.mark.parametrize("seed", range(N_CASES))
def test_random_plan_against_oracle(gpu_ctx, seed):
    c = draw_case(seed)
    label = (f"case{seed}-w{c['win']}-{np.dtype(c['dtype']).name}-{c['w']}x{c['h']}-p{c['pitch']}"
             f"-s{c['sx']}-v{c['variant']}-l{int(c['limit'])}")
    say(label)
    a, b = frames_for(c)
    rng = np.random.default_rng(7 + seed)
    pa, pb = padded(a, c["pitch"], rng), padded(b, c["pitch"], rng)
    plan = oa.Plan(gpu_ctx, c["w"], c["h"], c["win"], overlap=c["overlap"], step=c["step"], dtype=c["dtype"],
                   pitch=c["pitch"] if c["pitch"] != c["w"] else 0, variant=c["variant"],
                   limit_search=c["limit"])
    res = plan.run(pa, pb)
    # the direct and AUTO paths compute the circular correlation of the
    # complex variant; REAL_COMPAT has its own oracle variant
    ov = oa.REAL_COMPAT if c["variant"] == oa.REAL_COMPAT else oa.CIRCULAR_FFT_FP32
    ref = orc.process_pair(a, b, c["win"], c["sx"], c["sy"], variant=ov, limit_search=c["limit"], threads=16)
    # the window centres (x, and y measured from the bottom as main.cpp:220-234 does)
    assert len(res) == len(ref), (label, len(res), len(ref))
    assert np.array_equal(res["x"], ref["x"]) and np.array_equal(res["y"], ref["y"]), label
    small_int = np.dtype(c["dtype"]) == np.uint8 and c["win"] <= 16
    compare(res, ref, label, max_tie=0.4 if small_int else 0.05,
            max_peak1_diff=max(1, len(res) // 500) if small_int else 0,
            frames=(a, b, c["win"], c["sx"], c["sy"]))
