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
