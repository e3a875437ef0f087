"""Benchmark of the PIV window-correlation hot path on one MI355X.

A step is one pass of the fused engine over one synthetic frame pair that is
already resident in HBM: 4096 x 4096 float32 frames (particles d = 2.5 px,
density 0.05 /px, I0 ~ U(100, 255), noise sigma 2, solid-body rotation with
|d| = 6 px at the farthest window), 32 x 32 windows at step 16 (50%
overlap): 65025 windows -> 65025 vectors per step.

Timing: W untimed warm-up steps, then K launches bracketed by HIP events on
the engine's own stream (opiv_ctx_stream) with device synchronisation on
both sides.  Accuracy: a bounded sample of windows is compared with the FP64
oracle: max and RMS displacement error, integer-peak mismatches outside the
flagged windows, and the count of each flag.  The CPU baseline is the
reference's own FP64 path (oracle/_ref/ref_piv, the reference's headers and
thread pool) timed on this host's cores over the same frame pair, or the
oracle restatement when the reference harness is absent.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--win 32]
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "openpiv-cxx_amd"))

HBM_PEAK_GBS = 8000.0
# MI355X published dense peaks (TFLOP/s or TOP/s): FP32 vector, bf16 and int8 MFMA
FP32_PEAK_TFLOPS = 157.3
BF16_MFMA_PEAK_TFLOPS = 2500.0
I8_MFMA_PEAK_TOPS = 5000.0
METRIC = "vectors/sec at 32² win 50% overlap, chosen vs plain GPU path at ≤1e-3 px"
RESULT_BYTES = 72          # one opiv_window_result as stored
# algorithmic result bytes per window: the vector (u, v), its position and
# s/n as the reference's output needs them, 16 B (the record itself is 72 B)
ALG_RESULT_BYTES = 16
PREWARM_S = 0.3            # untimed launches before the warm-up steps


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class HipEvents:
    """hipEvent timing on an arbitrary stream through libamdhip64."""

    def __init__(self):
        import openpiv_amd as oa
        self.hip = oa._hip_runtime()
        self.hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def create(self):
        e = ctypes.c_void_p()
        assert self.hip.hipEventCreate(ctypes.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.hip.hipEventRecord(e, ctypes.c_void_p(stream)) == 0

    def elapsed_ms(self, e0, e1):
        assert self.hip.hipEventSynchronize(e1) == 0
        ms = ctypes.c_float()
        assert self.hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1) == 0
        return ms.value

    def graph_of(self, fn, stream):
        """fn's work on stream, captured once (relaxed mode) into an
        instantiated graph; replays drop the host's per-call launch cost."""
        h = self.hip
        h.hipStreamBeginCapture.argtypes = [ctypes.c_void_p, ctypes.c_int]
        h.hipStreamEndCapture.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
        h.hipGraphInstantiate.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_char_p, ctypes.c_size_t]
        h.hipGraphLaunch.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        h.hipGraphExecDestroy.argtypes = [ctypes.c_void_p]
        h.hipGraphDestroy.argtypes = [ctypes.c_void_p]
        graph, exe = ctypes.c_void_p(), ctypes.c_void_p()
        assert h.hipStreamBeginCapture(ctypes.c_void_p(stream), 2) == 0
        try:
            fn()
        finally:
            assert h.hipStreamEndCapture(ctypes.c_void_p(stream), ctypes.byref(graph)) == 0
        assert h.hipGraphInstantiate(ctypes.byref(exe), graph, None, None, 0) == 0
        return graph, exe

    def graph_launch(self, exe, stream):
        assert self.hip.hipGraphLaunch(exe, ctypes.c_void_p(stream)) == 0

    def graph_destroy(self, graph, exe):
        self.hip.hipGraphExecDestroy(exe)
        self.hip.hipGraphDestroy(graph)


def workload(win: int, dtype: str = "f32"):
    """4096^2 frame pair of the benchmark workloads: f32 with solid-body
    rotation (|d| = 6 px at the farthest window), or uint16 (I0 ~ U(1000,
    60000), noise 50, shift (0.4, 1.1)), or uint8 (I0 ~ U(30, 250), noise 3,
    same shift)."""
    import openpiv_amd as oa
    w = 4096
    step = win // 2
    if dtype == "f32":
        r = float(np.hypot(w / 2 - win / 2, w / 2 - win / 2))
        a, b = oa.synth_pair(w, w, seed=2, field=oa.FIELD_ROTATION, params=(6.0, r), dtype=np.float32)
    elif dtype == "u16":
        a, b = oa.synth_pair(w, w, seed=4, i0=(1000.0, 60000.0), noise=50.0, params=(0.4, 1.1), dtype=np.uint16)
    else:
        a, b = oa.synth_pair(w, w, seed=4, i0=(30.0, 250.0), noise=3.0, params=(0.4, 1.1), dtype=np.uint8)
    return a, b, w, step


def host_cores():
    """Cores this process may run on (affinity mask) and the CPU model."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    model = "unknown"
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
                break
    except (OSError, subprocess.SubprocessError):
        pass
    return n, model


def cpu_loop_times(a, b, win, step, threads, timed=5):
    """Processing-loop times (s) of the reference's FP64 path on `threads`
    pool threads: one warm-up loop, then `timed` loops (main.cpp:239,304-309
    time only the loop).  Returns (times, kind)."""
    ref_bin = ROOT / "oracle" / "_ref" / "ref_piv"
    h, w = a.shape
    if ref_bin.exists():
        with tempfile.TemporaryDirectory() as d:
            pa, pb = Path(d) / "a.bin", Path(d) / "b.bin"
            a.astype(np.float64).tofile(pa)
            b.astype(np.float64).tofile(pb)
            out = subprocess.run([str(ref_bin), "time", str(pa), str(pb), str(w), str(h), str(win), str(step),
                                  str(step), str(threads), str(timed + 1)], capture_output=True, text=True,
                                 check=True, timeout=600).stdout.split("\n")
        # lines in run order; the first loop is the warm-up
        us = [float(l.split()[1]) for l in out if l.strip()][1:]
        return [u * 1e-6 for u in us], "reference"
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc
    times = []
    for k in range(timed + 1):
        t0 = time.perf_counter()
        orc.process_pair(a, b, win, step, step, threads=threads)
        if k:
            times.append(time.perf_counter() - t0)
    return times, "port"


def cpu_baseline(a, b, win, step):
    """Reference FP64 path on this host's cores, vectors/s over the whole
    pair: nproc - 1 pool threads (the reference default, main.cpp:46) and
    nproc threads, median of 5 loops after 1 warm-up each.  `value` is the
    nproc - 1 figure."""
    h, w = a.shape
    nwin = (1 + (w - win) // step) * (1 + (h - win) // step)
    ncores, model = host_cores()
    by_threads = {}
    kind = "port"
    for threads in sorted({max(1, ncores - 1), ncores}):
        log(f"[bench] cpu baseline, {threads} threads")
        times, kind = cpu_loop_times(a, b, win, step, threads)
        by_threads[str(threads)] = round(nwin / float(np.median(times)), 1)
    default_threads = max(1, ncores - 1)
    what = ("reference FP64 algos::FFT path + its ThreadPool 'pool' strategy (examples/process/main.cpp:189-276), "
            "built from the reference sources") if kind == "reference" else "FP64 oracle restatement"
    return {"value": by_threads[str(default_threads)], "unit": "vectors/s", "cores": default_threads, "kind": kind,
            "sample": f"whole {w}x{h} pair, {nwin} windows, {what}; median of 5 timed loops after 1 warm-up",
            "by_threads": by_threads, "host_cores": ncores, "cpu_model": model}


def accuracy(res, a, b, win, step, sample=4096, seed=0):
    """Max |d(u,v)| against the FP64 oracle on a bounded window sample."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle as orc
    import openpiv_amd as oa
    n = len(res)
    rng = np.random.default_rng(seed)
    first = int(rng.integers(0, max(1, n - sample)))
    ref = orc.process_pair(a, b, win, step, step, threads=min(16, os.cpu_count() or 1), first=first, count=sample)
    got = res[first:first + len(ref)]
    tie = (got["flags"] & (oa.FLAG_NEAR_TIE | oa.FLAG_EXACT_TIE | oa.FLAG_FEWER_THAN_2_PEAKS)) != 0
    ok = ~tie & np.isfinite(ref["u"]) & np.isfinite(ref["v"])
    mism = int(((got["px"] != ref["px"]) | (got["py"] != ref["py"]))[ok].sum())
    du, dv = got["u"][ok] - ref["u"][ok], got["v"][ok] - ref["v"][ok]
    err = float(max(np.abs(du).max(), np.abs(dv).max()))
    rms = float(np.sqrt(np.mean(np.concatenate([du, dv]) ** 2)))
    counts = {name: int(((got["flags"] & bit) != 0).sum())
              for name, bit in (("near_tie", oa.FLAG_NEAR_TIE), ("exact_tie", oa.FLAG_EXACT_TIE),
                                ("fewer_than_2_peaks", oa.FLAG_FEWER_THAN_2_PEAKS),
                                ("nonfinite_fit", oa.FLAG_NONFINITE_FIT))}
    return {"max_err_px": err, "rms_err_px": rms, "checked_windows": int(len(ref)), "peak_mismatches": mism,
            "near_ties": int(tie.sum()), "flags": counts}


def copy_calibration(mib: int, reps: int):
    """Device-to-device copies of a known size: the runtime's copy kernel
    reads and writes exactly mib MiB per launch, which scales the FETCH_SIZE /
    WRITE_SIZE counters to bytes."""
    import openpiv_amd as oa
    n = mib << 20
    src, dst = oa.DeviceBuffer(n), oa.DeviceBuffer(n)
    src.zero()
    hip = oa._hip_runtime()
    for _ in range(reps):
        assert hip.hipMemcpy(dst.ptr, src.ptr, n, 3) == 0      # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    print(json.dumps({"copy_calibration_mib": mib, "reps": reps}), flush=True)


def bench_adjoint(steps: int, warmup: int, size: int = 1024):
    """size^2 f32 pair, uniform shift (1.6, -0.8), 32^2 windows at step 16:
    device time of the forward pass, of the VJP (both frames' gradients of
    sum cot . (u, v)) and of the variance propagation (sigma = 2)."""
    import openpiv_amd as oa
    a, b = oa.synth_pair(size, size, seed=5, params=(1.6, -0.8), dtype=np.float32)
    ctx = oa.Context(0)
    plan = oa.Plan(ctx, size, size, 32, dtype=np.float32)
    n = plan.count
    da, db = oa.DeviceBuffer.from_array(a), oa.DeviceBuffer.from_array(b)
    dfwd = oa.DeviceBuffer(n * RESULT_BYTES)
    dcot = oa.DeviceBuffer.from_array(np.random.default_rng(0).normal(size=(n, 2)))
    dga, dgb = oa.DeviceBuffer(size * size * 4), oa.DeviceBuffer(size * size * 4)
    dvar = oa.DeviceBuffer(n * 16)
    lib = oa.load()
    ev = HipEvents()

    def prewarm(fn):
        # clock settling, as in the default mode: PREWARM_S of untimed calls
        t_pre = time.perf_counter()
        while time.perf_counter() - t_pre < PREWARM_S:
            for _ in range(20):
                fn()
            ctx.synchronize()

    def timed(fn):
        prewarm(fn)
        for _ in range(warmup):
            fn()
        ctx.synchronize()
        e0, e1 = ev.create(), ev.create()
        ev.record(e0, ctx.stream)
        for _ in range(steps):
            fn()
        ev.record(e1, ctx.stream)
        ctx.synchronize()
        return ev.elapsed_ms(e0, e1) / steps

    def replayed(fn):
        # the same call captured once and replayed: when the host enqueues
        # slower than the device runs (small frames, two launches per VJP
        # call), the eager event time is the host's, not the device's
        graph, exe = ev.graph_of(fn, ctx.stream)
        try:
            prewarm(lambda: ev.graph_launch(exe, ctx.stream))
            for _ in range(warmup):
                ev.graph_launch(exe, ctx.stream)
            ctx.synchronize()
            e0, e1 = ev.create(), ev.create()
            ev.record(e0, ctx.stream)
            for _ in range(steps):
                ev.graph_launch(exe, ctx.stream)
            ev.record(e1, ctx.stream)
            ctx.synchronize()
            return ev.elapsed_ms(e0, e1) / steps
        finally:
            ev.graph_destroy(graph, exe)

    fwd = lambda: plan.run_async(da.ptr, db.ptr, dfwd.ptr)
    vjp = lambda: oa._check(lib.opiv_vjp_pair_async(plan.handle, da.ptr, db.ptr, dfwd.ptr, dcot.ptr,
                                                     dga.ptr, dgb.ptr))
    # both passes timed the same way: asynchronous calls on device buffers,
    # device time between HIP events, eager and graph-replayed
    fwd_ms = timed(fwd)
    vjp_ms = timed(vjp)
    ga_eager = dga.download(np.empty(size * size, np.float32))
    gb_eager = dgb.download(np.empty(size * size, np.float32))
    fwd_graph_ms = replayed(fwd)
    dga.zero()
    dgb.zero()
    vjp_graph_ms = replayed(vjp)
    # the replays wrote every gradient element again, bit for bit
    assert np.array_equal(dga.download(np.empty(size * size, np.float32)), ga_eager)
    assert np.array_equal(dgb.download(np.empty(size * size, np.float32)), gb_eager)
    var_ms = timed(lambda: oa._check(lib.opiv_variance_pair(plan.handle, da.ptr, db.ptr, dfwd.ptr, 2.0,
                                                             dvar.ptr)))
    # the figure is the eager ratio; the replayed one shows that the host's
    # enqueue rate does not set it (graph launches add their own overhead)
    f_ms, v_ms = fwd_ms, vjp_ms
    line = {"metric": f"adjoint cost relative to the forward pass ({size}^2 f32, 32^2 windows, step 16)",
            "value": round(v_ms / f_ms, 3), "unit": "x forward", "n_gpus": 1, "steps": steps,
            "higher_is_better": False, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"{size}x{size} f32 pair, uniform shift (1.6, -0.8), 32x32 windows step 16",
                       "windows": n},
            "forward_ms": round(f_ms, 4), "vjp_ms": round(v_ms, 4),
            "graph_replay": {"forward_ms": round(fwd_graph_ms, 4), "vjp_ms": round(vjp_graph_ms, 4),
                             "ratio": round(vjp_graph_ms / fwd_graph_ms, 3)},
            "variance_ms": round(var_ms, 4), "variance_over_forward": round(var_ms / f_ms, 3),
            "note": "forward and VJP: asynchronous calls on device buffers, device time between HIP events, "
                    "eager (the figure) and as a replayed graph of one call (graph_replay); "
                    "the variance call is synchronous and includes its stream synchronisation"}
    print(json.dumps(line), flush=True)


def bench_background(steps: int, warmup: int, nframes: int = 16):
    """examples/average_subtract/main.cpp:98-159 on nframes device-resident
    4096^2 uint16 frames (particle images), device outputs: wall time of the
    C-ABI call (it synchronises), HIP events around it."""
    import openpiv_amd as oa
    w = h = 4096
    frames = [oa.synth_pair(w, h, seed=60 + i, i0=(1000.0, 60000.0), noise=50.0, dtype=np.uint16)[0]
              for i in range(nframes)]
    ctx = oa.Context(0)
    dframes = [oa.DeviceBuffer.from_array(f) for f in frames]
    douts = [oa.DeviceBuffer(w * h * 2) for _ in range(nframes)]
    dmean = oa.DeviceBuffer(w * h * 8)
    fp, op = [d.ptr for d in dframes], [d.ptr for d in douts]
    ev = HipEvents()

    def call():
        ctx.average_subtract_device(fp, op, w, h, np.uint16, mean_ptr=dmean.ptr)

    for _ in range(warmup):
        call()
    ctx.synchronize()
    e0, e1 = ev.create(), ev.create()
    ev.record(e0, ctx.stream)
    for _ in range(steps):
        call()
    ev.record(e1, ctx.stream)
    ctx.synchronize()
    ms = ev.elapsed_ms(e0, e1) / steps
    # passes: mean (frames 1..n-1 read, mean written), range (every frame and
    # the mean once per 8-frame chunk read), scale (the same reads, outputs written)
    px = w * h
    chunks = (nframes + 7) // 8
    alg = (nframes - 1) * 2 * px + 8 * px + 2 * (nframes * 2 * px + chunks * 8 * px) + nframes * 2 * px
    line = {"metric": f"background subtraction of {nframes} 4096^2 uint16 frames (average_subtract)",
            "value": round(nframes / (ms * 1e-3), 1), "unit": "frames/s", "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": f"{nframes} x 4096x4096 uint16 frames, mean of frames 1..n-1, subtract, "
                                   "16-bit range scaling (bit-identical to the reference)", "frames": nframes},
            "roofline": {"bound": "hbm", "achieved": round(alg / (ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                         "algorithmic_bytes_per_step": alg,
                         "note": "three passes (mean, range, scale); the call's own synchronisation is inside"}}
    print(json.dumps(line), flush=True)


def fft_flops(win: int) -> float:
    """Algorithmic flops of one window on the FFT path: 5 N log2 N per complex
    N-point transform, halved for real input (two forward transforms, the
    half-spectrum product, one inverse)."""
    return 7.5 * win * win * np.log2(win * win) + 3 * win * win


def batch_compute(per_pass, wins, ms):
    flops = sum(n * fft_flops(w) for n, w in zip(per_pass, wins))
    rate = flops / (ms * 1e-3) / 1e12
    return {"bound": "fp32_valu", "achieved": round(rate, 2), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(rate / FP32_PEAK_TFLOPS, 4), "flops_per_step": round(flops)}


def bench_batch(steps: int, warmup: int):
    """24 frame pairs with seeded sizes from oa.MULTIPASS_SIZES, each with a seeded
    Lamb-Oseen vortex (core radius 40 px, peak 5 px), passes 64 -> 32 -> 16 at
    50% overlap, all windows in one graph launch per step."""
    import openpiv_amd as oa
    bufs, pairs = [], []
    for a, b in oa.multipass_pairs(24, seed=24):
        da, db = oa.DeviceBuffer.from_array(a), oa.DeviceBuffer.from_array(b)
        bufs.append((da, db))
        pairs.append((da.ptr, db.ptr, a.shape[1], a.shape[0]))
    log("[bench] 24 pairs:", [(w, h) for (_, _, w, h) in pairs])
    ctx = oa.Context(0)
    t0 = time.perf_counter()
    batch = oa.Batch(ctx, pairs, (64, 32, 16))
    tune_s = time.perf_counter() - t0
    dout = oa.DeviceBuffer(batch.count * RESULT_BYTES)
    for _ in range(warmup):
        batch.run(dout.ptr)
    ctx.synchronize()
    ev = HipEvents()
    e0, e1 = ev.create(), ev.create()
    ev.record(e0, ctx.stream)
    for _ in range(steps):
        batch.run(dout.ptr)
    ev.record(e1, ctx.stream)
    ctx.synchronize()
    ms = ev.elapsed_ms(e0, e1) / steps
    npx = sum(w * h for (_, _, w, h) in pairs)
    per_pass = [int(batch.offsets[(p + 1) * len(pairs)] - batch.offsets[p * len(pairs)]) for p in range(3)]
    line = {"metric": "vectors/sec, 24-pair multipass batch (64/32/16, 50% overlap), one graph launch",
            "value": round(batch.count / (ms * 1e-3), 1), "unit": "vectors/s", "n_gpus": 1, "steps": steps,
            "warmup": warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": "24 f32 pairs, sizes seeded from 512^2/1024^2/2048^2/4096^2/1280x800, "
                                   "Lamb-Oseen vortex r_c 40 px peak 5 px, passes 64/32/16 at 50%",
                       "pairs": len(pairs), "pixels_per_frame_total": npx, "windows": batch.count,
                       "windows_per_pass": per_pass,
                       "shapes": [batch.shape(p) for p in range(3)], "single_launch": batch.form(),
                       "create_and_tune_s": round(tune_s, 2)},
            "roofline": {"bound": "hbm", "achieved": round((3 * 2 * 4 * npx + batch.count * ALG_RESULT_BYTES)
                                                           / (ms * 1e-3) / 1e9, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "note": "algorithmic bytes: every pass reads both frames once"},
            "compute_roofline": batch_compute(per_pass, (64, 32, 16), ms)}
    print(json.dumps(line), flush=True)


def load_traffic(key: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE +
    WRITE_SIZE, copy-kernel calibrated), committed under profiles/."""
    f = ROOT / "profiles" / "traffic.json"
    if f.exists():
        d = json.loads(f.read_text())
        if key in d:
            return d[key]
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # defaults: ~10 ms of warm-up launches, then ~60 ms timed, so the clock has
    # settled and launch-to-launch jitter averages out
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--win", type=int, default=32)
    ap.add_argument("--step", type=int, default=0, help="grid step (default: half the window, 50%% overlap)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--path", choices=["auto", "fft", "direct"], default="auto",
                    help="correlation path: fused FP32 FFT, direct MFMA (16^2; 32^2 for 8/16-bit frames), "
                         "or the faster")
    ap.add_argument("--dtype", choices=["f32", "u16", "u8"], default="f32")
    ap.add_argument("--adjoint", action="store_true",
                    help="f32 differentiation workload: forward vs VJP vs variance (prints its own line)")
    ap.add_argument("--adjoint-size", type=int, default=1024, help="frame side of the --adjoint workload")
    ap.add_argument("--batch", action="store_true",
                    help="24 uneven frame pairs x passes 64/32/16 in one graph launch (prints its own line)")
    ap.add_argument("--background", action="store_true",
                    help="background subtraction of 16 4096^2 uint16 frames (prints its own line)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs the hot path on a single MI355X "
              "(multi-GPU sharding is a design note, not built); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import openpiv_amd as oa

    if args.copy_calibration:
        copy_calibration(args.copy_calibration, max(1, args.steps))
        return
    if args.batch:
        bench_batch(args.steps, args.warmup)
        return
    if args.background:
        bench_background(args.steps, args.warmup)
        return
    if args.adjoint:
        bench_adjoint(args.steps, args.warmup, args.adjoint_size)
        return

    win = args.win
    log(f"[bench] generating 4096x4096 {args.dtype} pair, win={win}")
    a, b, w, step = workload(win, args.dtype)
    if args.step:
        step = args.step
    ctx = oa.Context(0)
    np_dtype = {"f32": np.float32, "u16": np.uint16, "u8": np.uint8}[args.dtype]
    variants = {"auto": oa.AUTO, "fft": oa.CIRCULAR_FFT_FP32, "direct": oa.DIRECT_BF16_SPLIT}
    # frames resident in HBM before timing (device buffers from the engine's
    # own HIP runtime)
    da, db = oa.DeviceBuffer.from_array(a), oa.DeviceBuffer.from_array(b)
    ev = HipEvents()
    hip = oa._hip_runtime()

    def measure(variant):
        plan = oa.Plan(ctx, w, w, win, step=(step, step), dtype=np_dtype, variant=variant)
        n = plan.count
        dout = oa.DeviceBuffer(n * RESULT_BYTES)
        # clock settling: an idle GPU ramps its clocks up over the first tens
        # of milliseconds of work, so the launches run for PREWARM_S of wall
        # time before the W warm-up steps (none of it is timed)
        t_pre = time.perf_counter()
        while time.perf_counter() - t_pre < PREWARM_S:
            for _ in range(20):
                plan.run_async(da.ptr, db.ptr, dout.ptr)
            ctx.synchronize()
        for _ in range(args.warmup):
            plan.run_async(da.ptr, db.ptr, dout.ptr)
        ctx.synchronize()
        e0, e1 = ev.create(), ev.create()
        hip.hipDeviceSynchronize()
        t0 = time.perf_counter()
        ev.record(e0, ctx.stream)
        for _ in range(args.steps):
            plan.run_async(da.ptr, db.ptr, dout.ptr)
        ev.record(e1, ctx.stream)
        ctx.synchronize()
        hip.hipDeviceSynchronize()
        wall = time.perf_counter() - t0
        ms = ev.elapsed_ms(e0, e1) / args.steps
        res = dout.download(np.zeros(n, oa.RESULT_DTYPE))
        return n, ms, wall, res

    log(f"[bench] timing {args.steps} steps (plain FFT path)")
    n, plain_ms, wall, res = measure(oa.CIRCULAR_FFT_FP32)
    chosen = args.path
    if args.path == "auto":
        chosen = "direct" if oa.direct_chosen(win, np_dtype) else "fft"
    ms_per_step = plain_ms
    if chosen == "direct":
        log(f"[bench] timing {args.steps} steps (direct MFMA path)")
        n, ms_per_step, wall, res = measure(oa.DIRECT_BF16_SPLIT)
    value = n / (ms_per_step * 1e-3)

    log("[bench] accuracy sample vs FP64 oracle")
    acc = accuracy(res, a, b, win, step)

    frame_bytes = 2 * w * w * np.dtype(np_dtype).itemsize
    alg_bytes = frame_bytes + n * ALG_RESULT_BYTES
    achieved = alg_bytes / (ms_per_step * 1e-3) / 1e9
    # measured traffic exists for the default grid step (half the window) only
    key = f"{'direct_bf16' if chosen == 'direct' else 'fft_fp32'}_w{win}_4096_{args.dtype}"
    if step != win // 2:
        key += f"_s{step}"
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": load_traffic(key),
                "traffic_calibration": "rocprofv3 --pmc FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000 (KiB per dispatch; "
                                       "factors from a device-to-device copy of known size)",
                "algorithmic_bytes_per_launch": alg_bytes,
                "bytes_per_window": round(alg_bytes / n, 1)}

    # compute side of the roofline.  FFT path: F(W) = 7.5 W^2 log2(W^2) + 3 W^2
    # flops per window (5 N log2 N per complex N-point transform, halved for
    # real input: two forward transforms, the half-spectrum product, one
    # inverse) against the FP32 vector peak.  Direct path: the window's GEMM,
    # 2 W^4 operations per product against the MFMA peak of its type: int8
    # for 8-bit frames (one product) and for 16-bit frames at 32 x 32 (four
    # byte-plane products), split bf16 otherwise (4 products for u16, 6 for
    # float frames).
    direct_i8 = chosen == "direct" and (np_dtype == np.uint8 or (np_dtype == np.uint16 and win == 32))
    if chosen == "fft":
        per_win = fft_flops(win)
        peak, unit, kind = FP32_PEAK_TFLOPS, "TFLOP/s", "fp32_valu"
    elif direct_i8:
        per_win = 2.0 * win ** 4 * (1 if np_dtype == np.uint8 else 4)
        peak, unit, kind = I8_MFMA_PEAK_TOPS, "TOP/s", "mfma_i8"
    else:
        per_win = 2.0 * win ** 4 * (4 if np_dtype == np.uint16 else 6)
        peak, unit, kind = BF16_MFMA_PEAK_TFLOPS, "TFLOP/s", "mfma_bf16"
    rate = per_win * n / (ms_per_step * 1e-3) / 1e12
    floor_us = per_win * n / (peak * 1e12) * 1e6
    hbm_floor_us = alg_bytes / (HBM_PEAK_GBS * 1e9) * 1e6
    compute = {"bound": kind, "achieved": round(rate, 2), "peak": peak, "unit": unit,
               "frac": round(rate / peak, 4), "ops_per_window": round(per_win, 1),
               "floor_us": round(floor_us, 2), "hbm_floor_us": round(hbm_floor_us, 2),
               "binding_floor": kind if floor_us > hbm_floor_us else "hbm"}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(a, b, win, step)

    path_name = {"fft": "fft_fp32_fused",
                 "direct": "direct_i8_mfma" if direct_i8 else "direct_bf16_split_mfma"}[chosen]
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "vectors/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        # arithmetic of the correlation: FP32 FFT, int8 products accumulated in
        # int32 (8-bit frames on the direct path), or bf16 products accumulated in FP32
        "dtype": "i8" if direct_i8 else ("bf16" if chosen == "direct" else "f32"), "data": "synthetic",
        "config": {"workload": f"4096x4096 {args.dtype} frame pair, {win}x{win} windows step {step} ({n} windows), "
                               f"correlation + find_peaks + Gaussian fit fused, chosen path {path_name}",
                   "windows": n, "window": win, "step": step, "path": path_name, "frame_dtype": args.dtype,
                   "frames_resident_in_hbm": True},
        "plain_gpu": {"path": "fft_fp32_fused", "ms_per_step": round(plain_ms, 5),
                      "value": round(n / (plain_ms * 1e-3), 1)},
        "roofline": roofline, "compute_roofline": compute, "cpu_baseline": cpu, "accuracy": acc,
        "wall_s_timed_region": round(wall, 4), "prewarm_s": PREWARM_S,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
