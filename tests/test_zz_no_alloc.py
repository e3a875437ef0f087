"""The allocation-free VJP entry point (opiv_vjp_pair_async_noalloc) against
opiv_vjp_pair_async through the C ABI: the same kernels on the same inputs,
so the gradients must be bit-identical.  Two calls in a row, the second on
a smaller plan (the per-GPU scratch keeps the first call's tap weights,
which the second must not read), then a plan over the scratch's limit is
refused before any launch."""
import numpy as np
import pytest

import openpiv_amd as oa

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def vjp_both(ctx, size, win, seed):
    lib = oa.load()
    a, b = oa.synth_pair(size, size, seed=seed, params=(1.6, -0.8), dtype=np.float32)
    plan = oa.Plan(ctx, size, size, win, dtype=np.float32)
    n = plan.count
    da, db = oa.DeviceBuffer.from_array(a), oa.DeviceBuffer.from_array(b)
    dfwd = oa.DeviceBuffer(n * oa.RESULT_DTYPE.itemsize)
    plan.run_async(da.ptr, db.ptr, dfwd.ptr)
    cot = np.random.default_rng(seed).normal(size=(n, 2))
    dcot = oa.DeviceBuffer.from_array(cot)
    grads = []
    for fn in (lib.opiv_vjp_pair_async, lib.opiv_vjp_pair_async_noalloc):
        dga, dgb = oa.DeviceBuffer(size * size * 4), oa.DeviceBuffer(size * size * 4)
        dga.zero()
        dgb.zero()
        oa._check(fn(plan.handle, da.ptr, db.ptr, dfwd.ptr, dcot.ptr, dga.ptr, dgb.ptr))
        ctx.synchronize()
        grads.append((dga.download(np.empty((size, size), np.float32)),
                      dgb.download(np.empty((size, size), np.float32))))
    (ga, gb), (na, nb) = grads
    assert np.abs(ga).max() > 0 and np.abs(gb).max() > 0
    return (ga, gb), (na, nb)


@pytest.mark.parametrize("win", [16, 32])
def test_noalloc_vjp_matches_allocating_entry_point(gpu_ctx, win):
    for k, size in enumerate((1024, 384)):        # second call smaller than the first
        say(f"noalloc vjp win {win} size {size}")
        (ga, gb), (na, nb) = vjp_both(gpu_ctx, size, win, seed=40 + k)
        assert np.array_equal(ga, na) and np.array_equal(gb, nb), (win, size)


def test_noalloc_vjp_refuses_plans_over_its_limit(gpu_ctx):
    lib = oa.load()
    # 8 x 8 windows at step 4 over 2064 x 2064: 515 x 515 = 265225 > 262144
    plan = oa.Plan(gpu_ctx, 2064, 2064, 8, step=(4, 4), dtype=np.float32)
    assert plan.count > 262144
    buf = oa.DeviceBuffer(1 << 20)
    rc = lib.opiv_vjp_pair_async_noalloc(plan.handle, buf.ptr, buf.ptr, buf.ptr, buf.ptr, buf.ptr, buf.ptr)
    assert rc == 6 and "scratch" in oa.last_error()
