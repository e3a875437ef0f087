"""GPU background subtraction (examples/average_subtract/main.cpp:98-159)
against the reference's own pipeline (golden fixtures from its loader, image
arithmetic, find_image_range and PNM writer) and the FP64 oracle: the mean
and every saved 16-bit image bit-identical."""
from pathlib import Path

import numpy as np
import pytest

import openpiv_amd as oa
import oracle as orc

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).parent / "golden" / "avgsub_cases.npz"


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("case", ["u16", "u8", "pair", "same", "piv"])
def test_average_subtract_matches_reference(gpu_ctx, case):
    g = np.load(GOLDEN)
    out, mean = gpu_ctx.average_subtract(g[f"{case}_frames"], return_mean=True)
    assert np.array_equal(mean, g[f"{case}_mean"])
    assert np.array_equal(out, g[f"{case}_out"])


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32, np.float64])
def test_average_subtract_matches_oracle(gpu_ctx, dtype):
    """17 frames (two chunks of 8 plus one) of 333 x 257 pixels."""
    rng = np.random.default_rng(17)
    hi = {np.uint8: 255, np.uint16: 65535}.get(dtype, 1.0)
    frames = (rng.random((17, 257, 333)) * hi).astype(dtype)
    if dtype in (np.float32, np.float64):
        frames = frames - 0.25                                     # negative values too
    out, mean = gpu_ctx.average_subtract(frames, return_mean=True)
    rmean, rout = orc.average_subtract(frames)
    assert np.array_equal(mean, rmean)
    assert np.array_equal(out, rout)


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32, np.float64])
def test_average_subtract_runs_of_8_match_oracle(gpu_ctx, dtype):
    """17 frames of 256 x 200 pixels (widths and pitches that are multiples
    of 8 take the run-of-8 kernels), host frames and device frames with a
    wider pitch"""
    rng = np.random.default_rng(23)
    hi = {np.uint8: 255, np.uint16: 65535}.get(dtype, 1.0)
    n, h, w, pitch = 17, 200, 256, 264
    full = (rng.random((n, h, pitch)) * hi).astype(dtype)
    if dtype in (np.float32, np.float64):
        full = full - 0.25
    frames = np.ascontiguousarray(full[:, :, :w])
    rmean, rout = orc.average_subtract(frames)
    out, mean = gpu_ctx.average_subtract(frames, return_mean=True)
    assert np.array_equal(mean, rmean)
    assert np.array_equal(out, rout)
    dframes = [oa.DeviceBuffer.from_array(full[i]) for i in range(n)]
    douts = [oa.DeviceBuffer(h * w * 2) for _ in range(n)]
    dmean = oa.DeviceBuffer(h * w * 8)
    gpu_ctx.average_subtract_device([d.ptr for d in dframes], [d.ptr for d in douts], w, h, dtype,
                                    pitch=pitch, mean_ptr=dmean.ptr)
    assert np.array_equal(dmean.download(np.zeros((h, w))), rmean)
    for i in range(n):
        assert np.array_equal(douts[i].download(np.zeros((h, w), np.uint16)), rout[i])


def test_average_subtract_device_pointers_and_pitch(gpu_ctx):
    """device frames with pitch > width, device outputs and mean"""
    rng = np.random.default_rng(3)
    n, h, w, pitch = 5, 200, 300, 320
    full = rng.integers(0, 4096, (n, h, pitch)).astype(np.uint16)
    dframes = [oa.DeviceBuffer.from_array(full[i]) for i in range(n)]
    douts = [oa.DeviceBuffer(h * w * 2) for _ in range(n)]
    dmean = oa.DeviceBuffer(h * w * 8)
    gpu_ctx.average_subtract_device([d.ptr for d in dframes], [d.ptr for d in douts], w, h, np.uint16,
                                    pitch=pitch, mean_ptr=dmean.ptr)
    rmean, rout = orc.average_subtract(np.ascontiguousarray(full[:, :, :w]))
    assert np.array_equal(dmean.download(np.zeros((h, w))), rmean)
    for i in range(n):
        assert np.array_equal(douts[i].download(np.zeros((h, w), np.uint16)), rout[i])


def test_average_subtract_4096_u16(gpu_ctx):
    """8 synthetic 4096^2 uint16 frames (particles on a background)"""
    say("average_subtract 8 x 4096^2 u16")
    frames = np.stack([oa.synth_pair(4096, 4096, seed=40 + i, i0=(1000.0, 60000.0), noise=50.0,
                                     dtype=np.uint16)[0] for i in range(8)])
    out, mean = gpu_ctx.average_subtract(frames, return_mean=True)
    rmean, rout = orc.average_subtract(frames)
    assert np.array_equal(mean, rmean)
    assert np.array_equal(out, rout)


def test_average_subtract_needs_two_frames(gpu_ctx):
    with pytest.raises(oa.OpivError) as e:
        gpu_ctx.average_subtract(np.zeros((1, 8, 8), np.uint8))
    assert e.value.code == 1 and "two or more" in str(e.value)
