"""Argument checks at the boundary, on the GPU box: the C ABI refuses host
pointers where only device memory is accepted (a host pointer reaching a
kernel would fault), plans whose wave tasks would span more than 32-bit
frame offsets, and the Python layer refuses host frames or result arrays
smaller than the plan reads or writes.  No kernel is launched with a bad
argument."""
import numpy as np
import pytest

import openpiv_amd as oa

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_async_entry_points_refuse_host_memory(gpu_ctx):
    say("host pointers at the async entry points")
    plan = oa.Plan(gpu_ctx, 256, 256, 32, dtype=np.float32)
    a = np.zeros((256, 256), np.float32)
    out = np.zeros(plan.count, oa.RESULT_DTYPE)
    da = oa.DeviceBuffer.from_array(a)
    dout = oa.DeviceBuffer(plan.count * oa.RESULT_DTYPE.itemsize)
    with pytest.raises(oa.OpivError) as e:
        plan.run_async(a.ctypes.data, da.ptr, dout.ptr)
    assert e.value.code == 1 and "device memory" in str(e.value)
    with pytest.raises(oa.OpivError) as e:
        plan.run_async(da.ptr, da.ptr, out.ctypes.data)
    assert e.value.code == 1
    lib = oa.load()
    cot = np.zeros((plan.count, 2))
    dcot = oa.DeviceBuffer.from_array(cot)
    dg = oa.DeviceBuffer(256 * 256 * 4)
    rc = lib.opiv_vjp_pair_async(plan.handle, da.ptr, da.ptr, dout.ptr, cot.ctypes.data, dg.ptr, dg.ptr)
    assert rc == 1 and "device memory" in oa.last_error()
    # the same calls with device memory go through
    plan.run_async(da.ptr, da.ptr, dout.ptr)
    gpu_ctx.synchronize()


def test_batch_refuses_host_memory(gpu_ctx):
    a = np.zeros((128, 128), np.float32)
    da = oa.DeviceBuffer.from_array(a)
    with pytest.raises(oa.OpivError) as e:
        oa.Batch(gpu_ctx, [(a.ctypes.data, da.ptr, 128, 128)], (32,))
    assert e.value.code == 1 and "device memory" in str(e.value)
    batch = oa.Batch(gpu_ctx, [(da.ptr, da.ptr, 128, 128)], (32,))
    host_out = np.zeros(batch.count, oa.RESULT_DTYPE)
    with pytest.raises(oa.OpivError) as e:
        batch.run(host_out.ctypes.data)
    assert e.value.code == 1


def test_plan_refuses_task_spans_past_32_bits(gpu_ctx):
    """4 Mi-pixel f32 rows (16 MiB, the widest accepted) with an explicit row
    step of 256: a wave task of two 64 x 64 windows in consecutive grid rows
    spans 320 rows = 5 GiB of offsets"""
    with pytest.raises(oa.OpivError) as e:
        oa.Plan(gpu_ctx, 1 << 22, 400, 64, step=(1 << 22, 256), dtype=np.float32)
    assert e.value.code == 6 and "4 GiB" in str(e.value)
    # the same frames at the default step are fine (no frames are touched)
    oa.Plan(gpu_ctx, 1 << 22, 400, 64, step=(1 << 21, 32), dtype=np.float32)


def test_python_layer_checks_host_sizes(gpu_ctx):
    plan = oa.Plan(gpu_ctx, 128, 96, 32, dtype=np.float32, pitch=130)
    full = np.zeros((96, 130), np.float32)
    with pytest.raises(ValueError):
        plan.run(np.zeros((96, 128), np.float32), np.zeros((96, 128), np.float32))   # pitch 130 needed
    with pytest.raises(ValueError):
        plan.run(full, np.zeros((95, 130), np.float32))
    with pytest.raises(ValueError):
        plan.run(full, full, out=np.zeros(plan.count - 1, oa.RESULT_DTYPE))
    with pytest.raises(ValueError):
        plan.vjp(full, full, np.zeros(plan.count))
    res = plan.run(full, full)
    assert len(res) == plan.count
    # device buffers refuse host arrays larger than themselves (not asserts:
    # they hold under python -O)
    buf = oa.DeviceBuffer(64)
    with pytest.raises(ValueError):
        buf.upload(np.zeros(17, np.float32))
    with pytest.raises(ValueError):
        buf.download(np.zeros(17, np.float32))
    with pytest.raises(ValueError):
        buf.download(np.zeros((4, 8), np.float32)[:, ::2])
