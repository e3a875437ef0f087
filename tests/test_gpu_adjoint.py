"""GPU adjoint (VJP of each window's (u, v) with respect to both frames, and
the propagated noise variance) against the FP64 CPU adjoint of the oracle and
against FP64 central finite differences (step 1e-3) of the oracle's forward
path.  Bar: gradients within 1e-4 of the window's max |gradient|, variances
within 1e-4 relative."""
import numpy as np
import pytest

import openpiv_amd as oa
import oracle as orc

pytestmark = pytest.mark.gpu

GRAD_TOL = 1e-4
H = 1e-3


def say(*a):
    print(*a, flush=True)


def frames(w=1024, h=1024, seed=5, dtype=np.float32):
    # the differentiation workload: uniform shift (1.6, -0.8)
    return oa.synth_pair(w, h, seed=seed, params=(1.6, -0.8), dtype=dtype)


def isolated_cotangents(nx, ny, px, py, cu, cv):
    """cotangent only on windows (ix, iy) with ix % 2 == px, iy % 2 == py:
    at 50% overlap those windows do not overlap, so the frame gradient inside
    each is that window's own gradient"""
    cot = np.zeros((ny, nx, 2))
    cot[py::2, px::2, 0] = cu
    cot[py::2, px::2, 1] = cv
    return cot.reshape(-1, 2)


@pytest.mark.parametrize("win", [16, 32, 64])
def test_vjp_matches_fp64_adjoint(gpu_ctx, win):
    say(f"vjp vs oracle adjoint win={win}")
    a, b = frames(512, 512)
    step = win // 2
    plan = oa.Plan(gpu_ctx, 512, 512, win, step=(step, step), dtype=np.float32)
    fwd = plan.run(a, b)
    xy = plan.grid()
    nx = 1 + (512 - win) // step
    ny = len(xy) // nx
    checked = 0
    for cu, cv in ((1.0, 0.0), (0.0, 1.0)):
        cot = isolated_cotangents(nx, ny, 1, 1, cu, cv)
        ga, gb = plan.vjp(a, b, cot, fwd)
        for i in np.flatnonzero(cot[:, 0] + cot[:, 1]):
            x0, y0 = xy[i]
            j = orc.window_jacobian(a[y0:y0 + win, x0:x0 + win], b[y0:y0 + win, x0:x0 + win])
            if j is None or fwd["flags"][i] & (oa.FLAG_NEAR_TIE | oa.FLAG_EXACT_TIE):
                continue
            ref_a, ref_b = (j[0], j[1]) if cu else (j[2], j[3])
            got_a = ga[y0:y0 + win, x0:x0 + win]
            got_b = gb[y0:y0 + win, x0:x0 + win]
            scale = max(np.abs(ref_a).max(), np.abs(ref_b).max())
            assert np.abs(got_a - ref_a).max() <= GRAD_TOL * scale, (i, np.abs(got_a - ref_a).max() / scale)
            assert np.abs(got_b - ref_b).max() <= GRAD_TOL * scale, (i, np.abs(got_b - ref_b).max() / scale)
            checked += 1
    assert checked > 20


def test_vjp_random_cotangent_whole_frame(gpu_ctx):
    a, b = frames(256, 256, seed=8)
    plan = oa.Plan(gpu_ctx, 256, 256, 32, dtype=np.float32)
    rng = np.random.default_rng(0)
    cot = rng.normal(size=(plan.count, 2))
    fwd = plan.run(a, b)
    # windows the engine flags as near ties contribute with whichever peak
    # FP32 picked; zero their cotangent so both sides differentiate the same peak
    cot[(fwd["flags"] & (oa.FLAG_NEAR_TIE | oa.FLAG_EXACT_TIE)) != 0] = 0
    ga, gb = plan.vjp(a, b, cot)                   # forward computed inside
    ra, rb = orc.vjp_pair(a, b, 32, 16, 16, cot)
    for got, ref in ((ga, ra), (gb, rb)):
        assert np.abs(got - ref).max() <= GRAD_TOL * np.abs(ref).max()


def test_differentiation_workload_full_size(gpu_ctx):
    """The differentiation workload at its stated size: a 1024^2 f32 pair
    shifted by (1.6, -0.8), 32^2 windows step 16 (3969 windows), a seeded
    cotangent on every window; the whole-frame gradients against the FP64
    CPU adjoint and the variances (sigma 2) against the FP64 variance"""
    a, b = frames(1024, 1024, seed=5)
    plan = oa.Plan(gpu_ctx, 1024, 1024, 32, dtype=np.float32)
    assert plan.count == 3969
    fwd = plan.run(a, b)
    tied = (fwd["flags"] & (oa.FLAG_NEAR_TIE | oa.FLAG_EXACT_TIE)) != 0
    rng = np.random.default_rng(1)
    cot = rng.normal(size=(plan.count, 2))
    cot[tied] = 0
    ga, gb = plan.vjp(a, b, cot, fwd)
    ra, rb = orc.vjp_pair(a, b, 32, 16, 16, cot)
    for got, ref in ((ga, ra), (gb, rb)):
        err = np.abs(got - ref).max() / np.abs(ref).max()
        assert err <= GRAD_TOL, err
    var = plan.variance(a, b, 2.0, fwd)
    ref = orc.variance_pair(a, b, 32, 16, 16, 2.0)
    ok = ~tied & ((fwd["flags"] & oa.FLAG_FEWER_THAN_2_PEAKS) == 0)
    assert ok.sum() > 3900
    assert np.allclose(var[ok], ref[ok], rtol=GRAD_TOL, atol=0)
    say(f"1024^2 vjp vs fp64 ok, {int(tied.sum())} tied windows, variance {int(ok.sum())} windows")


def test_variance_matches_fp64(gpu_ctx):
    a, b = frames(512, 512, seed=11)
    plan = oa.Plan(gpu_ctx, 512, 512, 32, dtype=np.float32)
    var = plan.variance(a, b, 2.0)
    ref = orc.variance_pair(a, b, 32, 16, 16, 2.0)
    fwd = plan.run(a, b)
    ok = (fwd["flags"] & (oa.FLAG_NEAR_TIE | oa.FLAG_EXACT_TIE | oa.FLAG_FEWER_THAN_2_PEAKS)) == 0
    assert np.allclose(var[ok], ref[ok], rtol=GRAD_TOL, atol=0)
    say(f"variance: median std u {np.sqrt(np.median(var[ok, 0])):.4f} px, v {np.sqrt(np.median(var[ok, 1])):.4f} px")


def test_vjp_against_central_differences(gpu_ctx):
    """64 seeded windows of the 1024^2 differentiation pair: GPU gradients
    against (u(+h) - u(-h)) / 2h of the FP64 oracle, h = 1e-3."""
    a, b = frames(1024, 1024, seed=5)
    a64, b64 = a.astype(np.float64), b.astype(np.float64)
    plan = oa.Plan(gpu_ctx, 1024, 1024, 32, dtype=np.float32)
    fwd = plan.run(a, b)
    xy = plan.grid()
    nx = 63
    rng = np.random.default_rng(64)
    picks = rng.choice(np.flatnonzero((np.arange(len(xy)) % nx) % 2 == 0), 64, replace=False)
    # isolate: only even columns of windows carry cotangent, rows every other too
    picks = [i for i in picks if (i // nx) % 2 == 0][:64]
    checked = 0
    for comp in (0, 1):
        cot = np.zeros((len(xy), 2))
        cot[picks, comp] = 1.0
        ga, gb = plan.vjp(a, b, cot, fwd)
        for i in picks:
            if fwd["flags"][i] & (oa.FLAG_NEAR_TIE | oa.FLAG_EXACT_TIE | oa.FLAG_FEWER_THAN_2_PEAKS):
                continue
            x0, y0 = xy[i]
            wa, wb = a64[y0:y0 + 32, x0:x0 + 32], b64[y0:y0 + 32, x0:x0 + 32]
            base = orc.window_vector(wa, wb)
            for frame, grad in ((0, ga), (1, gb)):
                y, x = rng.integers(0, 32, 2)
                plus, minus = [wa.copy(), wb.copy()], [wa.copy(), wb.copy()]
                plus[frame][y, x] += H
                minus[frame][y, x] -= H
                vp, vm = orc.window_vector(*plus), orc.window_vector(*minus)
                if (vp["px"], vp["py"], vm["px"], vm["py"]) != (base["px"], base["py"]) * 2:
                    continue
                fd = (vp["uv"[comp]] - vm["uv"[comp]]) / (2 * H)
                g = grad[y0 + y, x0 + x]
                scale = np.abs(grad[y0:y0 + 32, x0:x0 + 32]).max()
                assert abs(g - fd) <= GRAD_TOL * scale, (i, frame, comp, g, fd, scale)
                checked += 1
    say(f"finite-difference checks: {checked}")
    assert checked >= 100


def test_adjoint_rejects_real_variant(gpu_ctx):
    plan = oa.Plan(gpu_ctx, 128, 128, 32, variant=oa.REAL_COMPAT)
    z = np.zeros((128, 128), np.float32)
    with pytest.raises(oa.OpivError) as e:
        plan.variance(z, z, 1.0)
    assert e.value.code == 6


def test_vjp_async_matches_sync(gpu_ctx):
    """opiv_vjp_pair_async (device buffers, no synchronisation) writes the
    same gradients, bit for bit, as opiv_vjp_pair; missing forward results
    are an error for it."""
    a, b = frames(512, 512, seed=13)
    plan = oa.Plan(gpu_ctx, 512, 512, 32, dtype=np.float32)
    cot = np.random.default_rng(3).normal(size=(plan.count, 2))
    fwd = plan.run(a, b)
    ga, gb = plan.vjp(a, b, cot, fwd=fwd)
    lib = oa.load()
    da, db = oa.DeviceBuffer.from_array(a), oa.DeviceBuffer.from_array(b)
    dfwd, dcot = oa.DeviceBuffer.from_array(fwd), oa.DeviceBuffer.from_array(cot)
    dga, dgb = oa.DeviceBuffer(ga.nbytes), oa.DeviceBuffer(gb.nbytes)
    for _ in range(2):                               # second call reuses the context's tap-weight buffer
        oa._check(lib.opiv_vjp_pair_async(plan.handle, da.ptr, db.ptr, dfwd.ptr, dcot.ptr, dga.ptr, dgb.ptr))
        gpu_ctx.synchronize()
        assert dga.download(np.empty_like(ga)).tobytes() == ga.tobytes()
        assert dgb.download(np.empty_like(gb)).tobytes() == gb.tobytes()
    with pytest.raises(oa.OpivError) as e:
        oa._check(lib.opiv_vjp_pair_async(plan.handle, da.ptr, db.ptr, None, dcot.ptr, dga.ptr, dgb.ptr))
    assert e.value.code == 1


def _random_adjoint_case(seed):
    rng = np.random.default_rng(2000 + seed)
    win = int(rng.choice((8, 16, 32, 64)))
    dtype = (np.uint8, np.uint16, np.float32, np.float64)[seed % 4]
    w = int(rng.integers(win, win + 300))
    h = int(rng.integers(win, win + 260))
    pitch = w if rng.random() < 0.5 else w + int(rng.integers(1, 24))
    step = int(rng.integers(max(1, win // 4), win + win // 2 + 1)) if rng.random() < 0.5 else win // 2
    limit = bool(rng.random() < 0.3)
    shift = tuple(float(v) for v in rng.uniform(-win / 6, win / 6, size=2))
    return win, dtype, w, h, pitch, step, limit, shift


@pytest.mark.parametrize("seed", range(40))
def test_vjp_and_variance_random_plans(gpu_ctx, seed):
    """Seeded random plans (window 8-64, every pixel type, uneven frames with
    junk in the pitch padding, steps from W/4 to 1.5 W, limit-search on some):
    whole-frame gradients of a random cotangent against the FP64 CPU adjoint,
    zero in the padding, and the variances against the FP64 variance"""
    win, dtype, w, h, pitch, step, limit, shift = _random_adjoint_case(seed)
    say(f"adjoint case {seed}: w{win} {np.dtype(dtype).name} {w}x{h} pitch {pitch} step {step} limit {limit}")
    i0, noise = {np.uint8: ((30.0, 250.0), 3.0), np.uint16: ((1000.0, 60000.0), 50.0)}.get(dtype, ((100.0, 255.0), 2.0))
    a, b = oa.synth_pair(w, h, seed=90 + seed, i0=i0, noise=noise, params=shift, dtype=dtype)
    rng = np.random.default_rng(seed)
    pa = np.zeros((h, pitch), dtype)
    pb = np.zeros((h, pitch), dtype)
    pa[:, :w], pb[:, :w] = a, b
    pa[:, w:] = rng.uniform(0, 200, size=(h, pitch - w)).astype(dtype)
    pb[:, w:] = rng.uniform(0, 200, size=(h, pitch - w)).astype(dtype)
    plan = oa.Plan(gpu_ctx, w, h, win, step=(step, step), dtype=dtype, pitch=pitch if pitch != w else 0,
                   limit_search=limit)
    fwd = plan.run(pa, pb)
    skip = (fwd["flags"] & (oa.FLAG_NEAR_TIE | oa.FLAG_EXACT_TIE | oa.FLAG_NONFINITE_FIT)) != 0
    cot = rng.normal(size=(plan.count, 2))
    cot[skip] = 0
    ga, gb = plan.vjp(pa, pb, cot, fwd)
    ra, rb = orc.vjp_pair(a, b, win, step, step, cot, limit_search=limit)
    scale = max(np.abs(ra).max(), np.abs(rb).max())
    if scale > 0:
        for got, ref in ((ga, ra), (gb, rb)):
            err = np.abs(got[:, :w] - ref).max() / scale
            assert err <= GRAD_TOL, (seed, err)
    assert not ga[:, w:].any() and not gb[:, w:].any(), seed
    var = plan.variance(pa, pb, 2.0, fwd)
    ref = orc.variance_pair(a, b, win, step, step, 2.0, limit_search=limit)
    ok = ~skip & ((fwd["flags"] & oa.FLAG_FEWER_THAN_2_PEAKS) == 0)
    assert np.allclose(var[ok], ref[ok], rtol=GRAD_TOL, atol=0), (seed, np.abs(var[ok] / ref[ok] - 1).max())
