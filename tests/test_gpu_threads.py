"""Several host threads, each with its own context (one HIP stream each), on
one GPU at once: single pairs, batches (whose creation writes the tuning
cache) and the allocation-free VJP's per-device scratch lookup.  Every
result is bit-identical to the same call made alone, and each thread's
error text stays its own (opiv_last_error is per thread)."""
import threading

import numpy as np
import pytest

import openpiv_amd as oa

pytestmark = pytest.mark.gpu


def frames(k):
    return oa.synth_pair(512, 384, seed=700 + k, params=(1.0 + 0.3 * k, -0.7), dtype=np.float32)


def run_threads(fns):
    errors = []

    def wrap(fn):
        try:
            fn()
        except BaseException as e:   # reported in the main thread
            errors.append(e)

    ts = [threading.Thread(target=wrap, args=(f,)) for f in fns]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in ts), "a worker thread did not finish"
    if errors:
        raise errors[0]


def test_contexts_on_threads_match_sequential(gpu_ctx, tmp_path, monkeypatch):
    monkeypatch.setenv("OPIV_TUNE_CACHE", str(tmp_path / "tune.txt"))
    pairs = [frames(k) for k in range(4)]
    alone = []
    for a, b in pairs:
        plan = oa.Plan(gpu_ctx, 512, 384, 32, dtype=np.float32)
        alone.append(plan.run(a, b))
    # batches of two pairs each, alone
    bufs = [(oa.DeviceBuffer.from_array(a), oa.DeviceBuffer.from_array(b)) for a, b in pairs]
    descs = [[(bufs[2 * j + i][0].ptr, bufs[2 * j + i][1].ptr, 512, 384) for i in range(2)] for j in range(2)]

    def batch_results(ctx, d):
        batch = oa.Batch(ctx, d, (64, 32, 16))
        out = oa.DeviceBuffer(batch.count * oa.RESULT_DTYPE.itemsize)
        batch.run(out.ptr)
        ctx.synchronize()
        res = out.download(np.zeros(batch.count, oa.RESULT_DTYPE))
        batch.close()
        return res

    batch_alone = [batch_results(gpu_ctx, d) for d in descs]
    got = {}

    def worker(k):
        def fn():
            with oa.Context(0) as ctx:
                for rep in range(5):
                    for i in (k, k + 2):
                        a, b = pairs[i]
                        plan = oa.Plan(ctx, 512, 384, 32, dtype=np.float32)
                        r = plan.run(a, b)
                        assert r.tobytes() == alone[i].tobytes(), (k, rep, i)
                    got[k] = batch_results(ctx, descs[k])
                    # an error on this thread leaves its own text; the other
                    # thread's calls succeed meanwhile
                    with pytest.raises(oa.OpivError):
                        oa.Plan(ctx, 512, 384, 24 + k, dtype=np.float32)
                    assert f"[{24 + k},{24 + k}]" in oa.last_error()
        return fn

    run_threads([worker(0), worker(1)])
    for k in range(2):
        assert got[k].tobytes() == batch_alone[k].tobytes(), k


def test_noalloc_vjp_on_two_threads(gpu_ctx):
    """the allocation-free VJP looks up its per-device scratch array on first
    use; two threads with their own contexts on one GPU share that array,
    so their calls are made in turn (the documented contract) and each
    matches the allocating entry point"""
    lib = oa.load()
    lock = threading.Lock()
    out = {}

    def worker(k):
        def fn():
            a, b = frames(10 + k)
            with oa.Context(0) as ctx:
                plan = oa.Plan(ctx, 512, 384, 32, dtype=np.float32)
                fwd = plan.run(a, b)
                cot = np.random.default_rng(k).normal(size=(plan.count, 2))
                ga, gb = plan.vjp(a, b, cot, fwd)
                da, db = oa.DeviceBuffer.from_array(a), oa.DeviceBuffer.from_array(b)
                dfwd, dcot = oa.DeviceBuffer.from_array(fwd), oa.DeviceBuffer.from_array(cot)
                dga, dgb = oa.DeviceBuffer(ga.nbytes), oa.DeviceBuffer(gb.nbytes)
                for _ in range(3):
                    with lock:   # calls that overlap on one GPU share the scratch array
                        oa._check(lib.opiv_vjp_pair_async_noalloc(plan.handle, da.ptr, db.ptr, dfwd.ptr, dcot.ptr,
                                                                  dga.ptr, dgb.ptr))
                        ctx.synchronize()
                    assert dga.download(np.empty_like(ga)).tobytes() == ga.tobytes()
                    assert dgb.download(np.empty_like(gb)).tobytes() == gb.tobytes()
                out[k] = True
        return fn

    run_threads([worker(0), worker(1)])
    assert out == {0: True, 1: True}
