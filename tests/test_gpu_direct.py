"""The BF16-split MFMA direct-correlation path (16 x 16 windows) held to the
same bar as the FFT path (test_gpu_parity): planes within 1e-5 of their max,
integer peaks equal except flagged near-ties, displacements within 1e-3 px of
the FP64 oracle / reference goldens.  8-bit frames correlate exactly."""
import numpy as np
import pytest

import openpiv_amd as oa
import oracle as orc
from conftest import load_npz, synth_from_meta
from test_gpu_parity import PLANE_RTOL, compare

pytestmark = pytest.mark.gpu
D = oa.DIRECT_BF16_SPLIT


def say(*a):
    print(*a, flush=True)


def test_direct_planes(gpu_ctx):
    rng = np.random.default_rng(16)
    n = 37
    a = rng.random((n, 16, 16)) * 200
    b = np.roll(a, (1, -2), (1, 2)) + rng.random((n, 16, 16)) * 20
    got = gpu_ctx.cross_correlate(a, b, D)
    f = orc.FFT(16)
    for i in range(n):
        ref = f.cross_correlate(a[i], b[i])
        assert np.abs(got[i] - ref).max() <= PLANE_RTOL * np.abs(ref).max(), i


def test_direct_planes_exact_for_8bit_values(gpu_ctx):
    rng = np.random.default_rng(8)
    a = rng.integers(0, 256, (9, 16, 16)).astype(np.float64)
    b = rng.integers(0, 256, (9, 16, 16)).astype(np.float64)
    got = gpu_ctx.cross_correlate(a, b, D)
    f = orc.FFT(16)
    for i in range(9):
        # integer correlation < 2^24 per term sum: exact in FP32 accumulation
        exact = np.rint(f.cross_correlate(a[i], b[i]))
        assert np.array_equal(got[i], exact)


def test_direct_golden_synth512(gpu_ctx):
    g = load_npz("synth512_f64_w16.npz")
    a, b = synth_from_meta("synth512")
    plan = oa.Plan(gpu_ctx, 512, 512, 16, step=(8, 8), dtype=np.float64, variant=D)
    compare(plan.run(a, b), g, "c16")


def test_direct_golden_reference_frames(gpu_ctx):
    g = load_npz("piv_crop512_u8.npz")
    plan = oa.Plan(gpu_ctx, 512, 512, 16, step=(8, 8), dtype=np.uint8, variant=D)
    res = plan.run(g["a"], g["b"])
    compare(res, g, "c16")
    ok = (g["c16_flags"] & 1) == 0
    # exact integer correlation: peak heights equal the FP64 reference's
    assert np.array_equal(res["peak0"][ok], np.rint(g["c16_peak0"][ok]))


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32, np.float64])
def test_direct_dtypes(gpu_ctx, dtype):
    i0 = (30.0, 250.0) if dtype == np.uint8 else (1000.0, 60000.0)
    a, b = oa.synth_pair(768, 512, seed=21, i0=i0, noise=3.0 if dtype == np.uint8 else 50.0,
                         params=(0.4, 1.1), dtype=dtype)
    plan = oa.Plan(gpu_ctx, 768, 512, 16, overlap=0.5, dtype=dtype, variant=D)
    compare(plan.run(a, b), orc.process_pair(a, b, 16, 8, 8, threads=16), f"direct-{np.dtype(dtype).name}")


def test_direct_uint16_4096(gpu_ctx):
    """4096^2 uint16 pair, I0 ~ U(1000, 60000), noise 50, shift (0.4, 1.1):
    16^2 windows at step 8 (261121 windows), every window against the oracle."""
    say("direct uint16 4096 w16")
    a, b = oa.synth_pair(4096, 4096, seed=4, i0=(1000.0, 60000.0), noise=50.0, params=(0.4, 1.1),
                         dtype=np.uint16)
    plan = oa.Plan(gpu_ctx, 4096, 4096, 16, step=(8, 8), dtype=np.uint16, variant=D)
    compare(plan.run(a, b), orc.process_pair(a, b, 16, 8, 8, threads=16), "direct-u16-4096")


def test_direct_limit_search_and_auto(gpu_ctx):
    a, b = oa.synth_pair(640, 480, seed=9, params=(-2.2, 1.4), dtype=np.float32)
    plan = oa.Plan(gpu_ctx, 640, 480, 16, dtype=np.float32, variant=D, limit_search=True)
    compare(plan.run(a, b), orc.process_pair(a, b, 16, 8, 8, limit_search=True, threads=8), "direct-limit")
    # AUTO: the direct path for 8-bit frames at 16 x 16, the FFT kernel otherwise
    for dtype, chosen in ((np.float32, oa.CIRCULAR_FFT_FP32), (np.uint8, D)):
        fa, fb = (a, b) if dtype == np.float32 else (np.clip(a, 0, 255).astype(np.uint8),
                                                       np.clip(b, 0, 255).astype(np.uint8))
        assert oa.direct_chosen(16, dtype) == (chosen == D)
        auto = oa.Plan(gpu_ctx, 640, 480, 16, dtype=dtype, variant=oa.AUTO)
        ref = oa.Plan(gpu_ctx, 640, 480, 16, dtype=dtype, variant=chosen)
        assert auto.run(fa, fb).tobytes() == ref.run(fa, fb).tobytes()


def test_direct_rejects_other_sizes(gpu_ctx):
    with pytest.raises(oa.OpivError) as e:
        oa.Plan(gpu_ctx, 256, 256, 32, variant=D)
    assert e.value.code == 6


def test_direct_uint16_4096_true_75pct_overlap(gpu_ctx):
    """4096^2 uint16 at 16^2 windows, step 4 (true 75% overlap, 1,042,441
    windows) on the direct path, every window against the oracle"""
    say("direct uint16 4096 w16 step 4")
    a, b = oa.synth_pair(4096, 4096, seed=4, i0=(1000.0, 60000.0), noise=50.0, params=(0.4, 1.1),
                         dtype=np.uint16)
    plan = oa.Plan(gpu_ctx, 4096, 4096, 16, step=(4, 4), dtype=np.uint16, variant=D)
    assert plan.count == 1042441
    compare(plan.run(a, b), orc.process_pair(a, b, 16, 4, 4, threads=16), "direct-u16-4096-s4")


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_direct32_dtypes(gpu_ctx, dtype):
    """32 x 32 windows on the int8 MFMA path (8-bit pixels, or 16-bit pixels
    as two exact byte planes), ragged frames, against the oracle"""
    i0 = (30.0, 250.0) if dtype == np.uint8 else (1000.0, 60000.0)
    a, b = oa.synth_pair(777, 519, seed=23, i0=i0, noise=3.0 if dtype == np.uint8 else 50.0,
                         params=(0.4, 1.1), dtype=dtype)
    for step in (16, 8, 32):
        say(f"direct32 {np.dtype(dtype).name} step {step}")
        plan = oa.Plan(gpu_ctx, 777, 519, 32, step=(step, step), dtype=dtype, variant=D)
        compare(plan.run(a, b), orc.process_pair(a, b, 32, step, step, threads=16),
                f"direct32-{np.dtype(dtype).name}-s{step}")


def test_direct32_limit_search_and_pitch(gpu_ctx):
    a, b = oa.synth_pair(640, 480, seed=29, i0=(30.0, 250.0), noise=3.0, params=(-3.1, 2.2), dtype=np.uint8)
    plan = oa.Plan(gpu_ctx, 640, 480, 32, dtype=np.uint8, variant=D, limit_search=True)
    compare(plan.run(a, b), orc.process_pair(a, b, 32, 16, 16, limit_search=True, threads=16), "direct32-limit")
    # rows of 643 bytes: the element-load path
    pa = np.zeros((480, 643), np.uint8)
    pb = np.zeros((480, 643), np.uint8)
    pa[:, :640], pb[:, :640] = a, b
    plan = oa.Plan(gpu_ctx, 640, 480, 32, dtype=np.uint8, variant=D, pitch=643)
    ref = oa.Plan(gpu_ctx, 640, 480, 32, dtype=np.uint8, variant=D)
    assert plan.run(pa, pb).tobytes() == ref.run(a, b).tobytes()


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_direct32_4096(gpu_ctx, dtype):
    """4096^2 pairs at 32^2 / step 16 (65025 windows) on the direct path:
    the uint16 pair of the precision workload (I0 ~ U(1000, 60000), noise 50,
    shift (0.4, 1.1)) and its 8-bit counterpart, every window against the
    oracle"""
    say(f"direct32 4096 {np.dtype(dtype).name}")
    if dtype == np.uint16:
        a, b = oa.synth_pair(4096, 4096, seed=4, i0=(1000.0, 60000.0), noise=50.0, params=(0.4, 1.1), dtype=dtype)
    else:
        a, b = oa.synth_pair(4096, 4096, seed=4, i0=(30.0, 250.0), noise=3.0, params=(0.4, 1.1), dtype=dtype)
    plan = oa.Plan(gpu_ctx, 4096, 4096, 32, step=(16, 16), dtype=dtype, variant=D)
    compare(plan.run(a, b), orc.process_pair(a, b, 32, 16, 16, threads=16), f"direct32-4096-{np.dtype(dtype).name}")


def test_direct32_refuses_float_frames(gpu_ctx):
    for dtype in (np.float32, np.float64):
        with pytest.raises(oa.OpivError) as e:
            oa.Plan(gpu_ctx, 256, 256, 32, dtype=dtype, variant=D)
        assert e.value.code == 6
    with pytest.raises(oa.OpivError) as e:
        oa.Plan(gpu_ctx, 256, 256, 64, dtype=np.uint8, variant=D)
    assert e.value.code == 6
