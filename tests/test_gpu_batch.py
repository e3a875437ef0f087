"""Batches: many uneven frame pairs x passes 64 -> 32 -> 16 in one graph
launch.  Every window of the batch must be bit-identical to the single-pair
path (same kernel arithmetic, only the addressing differs), whatever launch
shape the tuner picks; the single-pair path itself is held to the oracle bar
in test_gpu_parity.  A sample of pairs is also checked against the oracle."""
import os

import numpy as np
import pytest

import openpiv_amd as oa
import oracle as orc
from test_gpu_parity import compare

pytestmark = pytest.mark.gpu

SIZES = [(512, 512), (1280, 800), (384, 256), (1024, 1024), (700, 420), (256, 256)]
WINS = (64, 32, 16)


def say(*a):
    print(*a, flush=True)


def make_pairs(dtype=np.float32, seed0=100):
    frames = []
    for k, (w, h) in enumerate(SIZES):
        rng = np.random.default_rng(seed0 + k)
        cx, cy = rng.uniform(0.3, 0.7) * w, rng.uniform(0.3, 0.7) * h
        a, b = oa.synth_pair(w, h, seed=seed0 + k, field=oa.FIELD_LAMB_OSEEN, params=(cx, cy, 40.0, 5.0),
                             dtype=dtype)
        frames.append((a, b))
    return frames


@pytest.fixture(scope="module")
def batch_setup(gpu_ctx):
    frames = make_pairs()
    bufs = [(oa.DeviceBuffer.from_array(a), oa.DeviceBuffer.from_array(b)) for a, b in frames]
    pairs = [(da.ptr, db.ptr, a.shape[1], a.shape[0]) for (a, b), (da, db) in zip(frames, bufs)]
    return frames, bufs, pairs


def eager(ctx, frames, win):
    out = []
    for a, b in frames:
        plan = oa.Plan(ctx, a.shape[1], a.shape[0], win, overlap=0.5, dtype=a.dtype)
        out.append(plan.run(a, b))
    return out


def run_batch(batch):
    dout = oa.DeviceBuffer(batch.count * oa.RESULT_DTYPE.itemsize)
    batch.run(dout.ptr)
    batch.ctx.synchronize()
    return dout, dout.download(np.zeros(batch.count, oa.RESULT_DTYPE))


def test_batch_equals_single_pair_path(gpu_ctx, batch_setup):
    frames, bufs, pairs = batch_setup
    say("batch vs single-pair")
    batch = oa.Batch(gpu_ctx, pairs, WINS)
    dout, res = run_batch(batch)
    for p, win in enumerate(WINS):
        ref = eager(gpu_ctx, frames, win)
        for i in range(len(frames)):
            assert batch.block(res, p, i).tobytes() == ref[i].tobytes(), (win, i)
    say("shapes", [batch.shape(p) for p in range(len(WINS))])


def test_batch_against_oracle_sample(gpu_ctx, batch_setup):
    frames, bufs, pairs = batch_setup
    batch = oa.Batch(gpu_ctx, pairs, WINS)
    _, res = run_batch(batch)
    for p, win in enumerate(WINS):
        for i in (1, 4):
            a, b = frames[i]
            ref = orc.process_pair(a, b, win, win // 2, win // 2, threads=16)
            compare(batch.block(res, p, i), ref, f"batch-w{win}-pair{i}")


def test_graph_replays_with_eager_work_between(gpu_ctx, batch_setup):
    frames, bufs, pairs = batch_setup
    batch = oa.Batch(gpu_ctx, pairs, WINS)
    dout = oa.DeviceBuffer(batch.count * oa.RESULT_DTYPE.itemsize)
    first = None
    plan = oa.Plan(gpu_ctx, 512, 512, 32, dtype=np.float32)
    for k in range(4):
        dout.zero()
        batch.run(dout.ptr)
        gpu_ctx.synchronize()
        got = dout.download(np.zeros(batch.count, oa.RESULT_DTYPE))
        first = got if first is None else first
        assert got.tobytes() == first.tobytes(), k
        plan.run(frames[0][0], frames[0][1])          # eager work between replays


def test_set_pairs_swaps_frames(gpu_ctx, batch_setup):
    frames, bufs, pairs = batch_setup
    batch = oa.Batch(gpu_ctx, pairs, (32,))
    swapped = [(pb, pa, w, h) for (pa, pb, w, h) in pairs]
    batch.set_pairs(swapped)
    _, res = run_batch(batch)
    for i, (a, b) in enumerate(frames):
        plan = oa.Plan(gpu_ctx, a.shape[1], a.shape[0], 32, dtype=a.dtype)
        assert batch.block(res, 0, i).tobytes() == plan.run(b, a).tobytes()
    with pytest.raises(oa.OpivError):
        batch.set_pairs(swapped[:-1])


def test_results_do_not_depend_on_launch_shape(gpu_ctx, batch_setup, tmp_path, monkeypatch):
    frames, bufs, pairs = batch_setup
    outs = []
    for nt, bpc4 in ((256, 0), (64, 4), (256, 8), (64, 64)):
        cache = tmp_path / f"tune_{nt}_{bpc4}.txt"
        n32 = sum((1 + (w - 32) // 16) * (1 + (h - 32) // 16) for w, h in SIZES)
        key = f"gfx950 32 2 {int(round(np.log2(n32)))}"
        cache.write_text(f"{key} {nt} {bpc4}\n")
        monkeypatch.setenv("OPIV_TUNE_CACHE", str(cache))
        monkeypatch.setenv("OPIV_TUNE", "0")
        batch = oa.Batch(gpu_ctx, pairs, (32,))
        t, blocks = batch.shape(0)
        assert t == nt, (t, nt)
        outs.append(run_batch(batch)[1])
        say(f"shape nt={t} blocks={blocks}")
    for o in outs[1:]:
        assert o.tobytes() == outs[0].tobytes()


def test_batch_errors(gpu_ctx, batch_setup):
    frames, bufs, pairs = batch_setup
    with pytest.raises(oa.OpivError, match="power of 2"):
        oa.Batch(gpu_ctx, pairs, (24,))
    with pytest.raises(oa.OpivError, match="bigger than image"):
        oa.Batch(gpu_ctx, [(pairs[0][0], pairs[0][1], 40, 40)], (64,))


def test_multipass_workload_full_size(gpu_ctx, monkeypatch):
    """The benchmark's batch exactly: 24 seeded pairs of 512^2 .. 4096^2 and
    1280 x 800 with Lamb-Oseen vortices (r_c 40 px, peak 5 px), passes
    64 / 32 / 16 at 50% overlap in one graph launch.  Every window of every
    pair and pass is bit-identical to the single-pair launch, in both launch
    forms; the largest pair and a 1280 x 800 pair of each pass are checked
    against the FP64 oracle window by window."""
    frames = list(oa.multipass_pairs(24, seed=24))
    sizes = [a.shape for a, _ in frames]
    say("multipass 24 pairs", sizes)
    bufs = [(oa.DeviceBuffer.from_array(a), oa.DeviceBuffer.from_array(b)) for a, b in frames]
    pairs = [(da.ptr, db.ptr, a.shape[1], a.shape[0]) for (a, _), (da, db) in zip(frames, bufs)]
    batch = oa.Batch(gpu_ctx, pairs, WINS)
    _, res = run_batch(batch)
    say("batch windows", batch.count, "shapes", [batch.shape(p) for p in range(3)], "form", batch.form())
    monkeypatch.setenv("OPIV_BATCH_FORM", "single" if not batch.form()[0] else "perpass")
    other = oa.Batch(gpu_ctx, pairs, WINS)
    assert other.form()[0] != batch.form()[0]
    assert run_batch(other)[1].tobytes() == res.tobytes()
    say("both launch forms bit-identical")
    big = max(range(len(frames)), key=lambda i: frames[i][0].size)
    wide = [i for i, s in enumerate(sizes) if s == (800, 1280)]
    for p, win in enumerate(WINS):
        for i, (a, _) in enumerate(frames):
            h, w = a.shape
            plan = oa.Plan(gpu_ctx, w, h, win, overlap=0.5, dtype=np.float32)
            one = plan.run(bufs[i][0].ptr, bufs[i][1].ptr)
            assert batch.block(res, p, i).tobytes() == one.tobytes(), (win, i)
        say(f"pass {win}: all {len(frames)} pairs bit-identical")
        for i in [big] + wide[:1]:
            a, b = frames[i]
            ref = orc.process_pair(a, b, win, win // 2, win // 2, threads=16)
            compare(batch.block(res, p, i), ref, f"multipass-w{win}-pair{i}-{a.shape[1]}x{a.shape[0]}",
                    frames=(a, b, win, win // 2, win // 2))
            say(f"pass {win} pair {i} {a.shape[1]}x{a.shape[0]} vs oracle ok")


def test_batch_with_128_pass(gpu_ctx, batch_setup):
    """a 128 x 128 pass (one workgroup per window) in a batch beside 32 x 32:
    bit-identical to the single-pair launches; pairs smaller than the window
    are refused by the grid as in the reference"""
    frames, bufs, pairs = batch_setup
    keep = [i for i, (a, _) in enumerate(frames) if min(a.shape) >= 128]
    sub = [pairs[i] for i in keep]
    batch = oa.Batch(gpu_ctx, sub, (128, 32))
    _, res = run_batch(batch)
    assert batch.shape(0) == (256, 0)
    for p, win in enumerate((128, 32)):
        for j, i in enumerate(keep):
            a, b = frames[i]
            plan = oa.Plan(gpu_ctx, a.shape[1], a.shape[0], win, overlap=0.5, dtype=a.dtype)
            assert batch.block(res, p, j).tobytes() == plan.run(a, b).tobytes(), (win, i)


@pytest.mark.parametrize("dtype", [np.float32, np.uint8, np.uint16, np.float64])
def test_single_launch_equals_per_pass(gpu_ctx, monkeypatch, dtype):
    """every pass in one persistent launch (the multi-pass kernel) against one
    launch per pass, passes 8 / 64 / 16 / 32 over the uneven pairs: the same
    per-window code runs either way, so the results are bit-identical"""
    frames = make_pairs(dtype, seed0=300)
    bufs = [(oa.DeviceBuffer.from_array(a), oa.DeviceBuffer.from_array(b)) for a, b in frames]
    pairs = [(da.ptr, db.ptr, a.shape[1], a.shape[0]) for (a, b), (da, db) in zip(frames, bufs)]
    wins = (8, 64, 16, 32)
    res = {}
    for form in ("perpass", "single"):
        monkeypatch.setenv("OPIV_BATCH_FORM", form)
        batch = oa.Batch(gpu_ctx, pairs, wins, dtype=dtype)
        single, blocks = batch.form()
        assert single == (form == "single") and (blocks > 0) == single, (form, single, blocks)
        res[form] = run_batch(batch)[1]
        say(f"{np.dtype(dtype).name} {form}: {batch.count} windows, blocks {blocks}")
    assert res["single"].tobytes() == res["perpass"].tobytes()


def test_single_launch_options_and_cache(gpu_ctx, batch_setup, tmp_path, monkeypatch):
    """limit_search and REAL_COMPAT reach the single launch; the cached form
    line (single launch, blocks per CU x 4) is obeyed without timing"""
    frames, bufs, pairs = batch_setup
    for kw in ({"limit_search": True}, {"variant": oa.REAL_COMPAT}):
        res = {}
        for form in ("perpass", "single"):
            monkeypatch.setenv("OPIV_BATCH_FORM", form)
            res[form] = run_batch(oa.Batch(gpu_ctx, pairs, WINS, **kw))[1]
        assert res["single"].tobytes() == res["perpass"].tobytes(), kw
    monkeypatch.setenv("OPIV_BATCH_FORM", "perpass")
    plain = run_batch(oa.Batch(gpu_ctx, pairs, WINS))[1]
    monkeypatch.setenv("OPIV_BATCH_FORM", "single")
    forced = oa.Batch(gpu_ctx, pairs, WINS).form()
    monkeypatch.delenv("OPIV_BATCH_FORM")
    n = sum((1 + (w - win) // (win // 2)) * (1 + (h - win) // (win // 2)) for win in WINS for w, h in SIZES)
    cache = tmp_path / "tune.txt"
    cache.write_text(f"gfx950 form 2 64/32/16 {int(round(np.log2(n)))} 1 16\n")
    monkeypatch.setenv("OPIV_TUNE_CACHE", str(cache))
    monkeypatch.setenv("OPIV_TUNE", "0")
    cached = oa.Batch(gpu_ctx, pairs, WINS)
    assert cached.form() == (True, 2 * forced[1]), (cached.form(), forced)
    assert run_batch(cached)[1].tobytes() == plain.tobytes()


def test_128_pass_stays_per_pass(gpu_ctx, batch_setup, monkeypatch):
    """a 128 x 128 pass runs one workgroup per window: the batch keeps one
    launch per pass even when the single launch is asked for"""
    frames, bufs, pairs = batch_setup
    sub = [p for p, (a, _) in zip(pairs, frames) if min(a.shape) >= 128]
    monkeypatch.setenv("OPIV_BATCH_FORM", "single")
    assert oa.Batch(gpu_ctx, sub, (128, 32)).form() == (False, 0)


@pytest.mark.parametrize("seed", range(24))
def test_random_batches_equal_single_pair_path(gpu_ctx, monkeypatch, seed):
    """Seeded random batches: 1-6 pairs of uneven sizes, 1-4 passes of any
    window size 8-128 with their own step fractions, either variant,
    limit-search on some, the launch form forced or tuned.  Every window of
    every pass and pair is bit-identical to the single-pair launch of the
    same plan."""
    rng = np.random.default_rng(3000 + seed)
    dtype = (np.uint8, np.uint16, np.float32, np.float64)[seed % 4]
    npasses = int(rng.integers(1, 5))
    wins = tuple(int(w) for w in rng.choice((8, 16, 32, 64, 128), size=npasses, p=(0.15, 0.3, 0.3, 0.15, 0.1)))
    overlaps = [float(o) for o in rng.choice((0.25, 0.5, 0.75, 1.0), size=npasses)]
    variant = (oa.CIRCULAR_FFT_FP32, oa.REAL_COMPAT)[int(rng.integers(2))]
    limit = bool(rng.random() < 0.3)
    form = ("perpass", "single", None)[seed % 3]
    lo = max(wins)
    sizes = [(int(rng.integers(lo, lo + 500)), int(rng.integers(lo, lo + 400))) for _ in range(int(rng.integers(1, 7)))]
    say(f"random batch {seed}: {np.dtype(dtype).name} wins {wins} overlaps {overlaps} variant {variant} "
        f"limit {limit} form {form} sizes {sizes}")
    frames = [oa.synth_pair(w, h, seed=400 + seed * 8 + k, params=tuple(rng.uniform(-3, 3, size=2)), dtype=dtype)
              for k, (w, h) in enumerate(sizes)]
    bufs = [(oa.DeviceBuffer.from_array(a), oa.DeviceBuffer.from_array(b)) for a, b in frames]
    pairs = [(da.ptr, db.ptr, a.shape[1], a.shape[0]) for (a, _), (da, db) in zip(frames, bufs)]
    if form:
        monkeypatch.setenv("OPIV_BATCH_FORM", form)
    else:
        monkeypatch.delenv("OPIV_BATCH_FORM", raising=False)
    batch = oa.Batch(gpu_ctx, pairs, wins, overlaps, dtype=dtype, variant=variant, limit_search=limit)
    _, res = run_batch(batch)
    for p, (win, ov) in enumerate(zip(wins, overlaps)):
        for i, (a, b) in enumerate(frames):
            plan = oa.Plan(gpu_ctx, a.shape[1], a.shape[0], win, overlap=ov, dtype=dtype, variant=variant,
                           limit_search=limit)
            one = plan.run(bufs[i][0].ptr, bufs[i][1].ptr)
            assert batch.block(res, p, i).tobytes() == one.tobytes(), (seed, p, win, i)
