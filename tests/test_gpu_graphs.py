"""A caller's own HIP graph around the asynchronous entry points: the forward
pass (opiv_run_pair_async) and the allocation-free VJP
(opiv_vjp_pair_async_noalloc, the entry point for callers that cannot
allocate, such as code under graph capture) captured together on the
context's stream, then replayed several times with eager work of the same
context and cleared outputs between the replays.  Every replay must equal
the eager calls bit for bit."""
import ctypes

import numpy as np
import pytest

import openpiv_amd as oa

pytestmark = pytest.mark.gpu

vp = ctypes.c_void_p


def hip_graph_api():
    h = oa._hip_runtime()
    h.hipStreamBeginCapture.argtypes = [vp, ctypes.c_int]
    h.hipStreamEndCapture.argtypes = [vp, ctypes.POINTER(vp)]
    h.hipGraphInstantiate.argtypes = [ctypes.POINTER(vp), vp, vp, ctypes.c_char_p, ctypes.c_size_t]
    h.hipGraphLaunch.argtypes = [vp, vp]
    h.hipGraphExecDestroy.argtypes = [vp]
    h.hipGraphDestroy.argtypes = [vp]
    for f in (h.hipStreamBeginCapture, h.hipStreamEndCapture, h.hipGraphInstantiate, h.hipGraphLaunch,
              h.hipGraphExecDestroy, h.hipGraphDestroy):
        f.restype = ctypes.c_int
    return h


@pytest.mark.parametrize("win", [16, 32])
def test_forward_and_noalloc_vjp_in_a_caller_graph(gpu_ctx, win):
    lib = oa.load()
    h = hip_graph_api()
    a, b = oa.synth_pair(768, 512, seed=41, params=(1.6, -0.8), dtype=np.float32)
    plan = oa.Plan(gpu_ctx, 768, 512, win, dtype=np.float32)
    n = plan.count
    cot = np.random.default_rng(win).normal(size=(n, 2))
    fwd = plan.run(a, b)
    ga, gb = plan.vjp(a, b, cot, fwd)
    da, db = oa.DeviceBuffer.from_array(a), oa.DeviceBuffer.from_array(b)
    dcot = oa.DeviceBuffer.from_array(cot)
    dout = oa.DeviceBuffer(n * oa.RESULT_DTYPE.itemsize)
    dga, dgb = oa.DeviceBuffer(ga.nbytes), oa.DeviceBuffer(gb.nbytes)
    stream = gpu_ctx.stream
    graph, exe = vp(), vp()
    assert h.hipStreamBeginCapture(vp(stream), 2) == 0            # relaxed mode
    try:
        oa._check(lib.opiv_run_pair_async(plan.handle, da.ptr, db.ptr, dout.ptr, n))
        oa._check(lib.opiv_vjp_pair_async_noalloc(plan.handle, da.ptr, db.ptr, dout.ptr, dcot.ptr, dga.ptr, dgb.ptr))
    finally:
        assert h.hipStreamEndCapture(vp(stream), ctypes.byref(graph)) == 0
    assert h.hipGraphInstantiate(ctypes.byref(exe), graph, None, None, 0) == 0
    try:
        other_a, other_b = oa.synth_pair(768, 512, seed=42, params=(-0.5, 2.2), dtype=np.float32)
        for rep in range(4):
            for buf in (dout, dga, dgb):
                buf.zero()
            assert h.hipGraphLaunch(exe, vp(stream)) == 0
            gpu_ctx.synchronize()
            got = dout.download(np.zeros(n, oa.RESULT_DTYPE))
            assert got.tobytes() == fwd.tobytes(), rep
            assert dga.download(np.empty_like(ga)).tobytes() == ga.tobytes(), rep
            assert dgb.download(np.empty_like(gb)).tobytes() == gb.tobytes(), rep
            # eager work of the same context between replays, through the
            # allocating paths
            plan.run(other_a, other_b)
            plan.vjp(other_a, other_b, cot)
    finally:
        h.hipGraphExecDestroy(exe)
        h.hipGraphDestroy(graph)
