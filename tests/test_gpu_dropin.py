"""GPU tests of the C++ drop-in layer: HipFFT (algos::FFT's call surface)
against the FP64 oracle, and the process front-end against the reference's
golden vectors and the oracle.  Same bar as test_gpu_parity: planes within
1e-5 of their max, displacements within 1e-3 px (plus the 6-significant-digit
print rounding of main.cpp:312-313)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

import openpiv_amd as oa
import oracle as orc
from conftest import load_npz
from dropin_util import AVGSUB_BIN, PROCESS_BIN, build_dropin_check, parse_process_output, write_pgm, write_tiff

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def dropin(tmp_path_factory):
    return build_dropin_check(tmp_path_factory.mktemp("dropin"))


@pytest.mark.parametrize("variant", ["complex", "real", "auto"])
def test_hipfft_matches_oracle(dropin, tmp_path, variant):
    say(f"HipFFT {variant}")
    rng = np.random.default_rng(4)
    a, b = rng.random((32, 32)) * 100, rng.random((32, 32)) * 100
    a.tofile(tmp_path / "a.bin")
    b.tofile(tmp_path / "b.bin")
    subprocess.run([str(dropin), "correlate", "32", variant, str(tmp_path / "a.bin"), str(tmp_path / "b.bin"),
                    str(tmp_path / "o.bin")], check=True, timeout=60)
    got = np.fromfile(tmp_path / "o.bin", dtype=np.float64).reshape(32, 32)
    f = orc.FFT(32)
    ref = {"complex": lambda: f.cross_correlate(a, b), "real": lambda: f.cross_correlate_real(a, b),
           "auto": lambda: f.auto_correlate(a)}[variant]()
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()


def test_hipfft_pow2_message(dropin):
    out = subprocess.run([str(dropin), "pow2", "24", "32"], capture_output=True, text=True, timeout=60).stdout
    assert "dimensions must be power of 2: [24,32]" in out


def check_output(text, ref_xy_u_v_sn, few):
    got = parse_process_output(text)
    assert got.shape == (len(few), 5)
    x, y, u, v, sn = ref_xy_u_v_sn
    ok = ~few
    assert np.array_equal(got[ok, 0], x[ok]) and np.array_equal(got[ok, 1], y[ok])
    tol = 1e-3 + 1e-5 * np.abs(u[ok])
    assert (np.abs(got[ok, 2] - u[ok]) <= tol).all()
    assert (np.abs(got[ok, 3] - v[ok]) <= 1e-3 + 1e-5 * np.abs(v[ok])).all()
    assert np.allclose(got[ok, 4], sn[ok], rtol=1e-4, atol=1e-5)
    assert (got[few] == 0).all()


@pytest.mark.parametrize("label,size,ftype", [("c32", 32, "complex"), ("r32", 32, "real"),
                                              ("c16", 16, "pocket"), ("c64", 64, "pocket_real")])
def test_process_on_reference_example_frames(tmp_path, label, size, ftype):
    say(f"process {label} -f {ftype}")
    g = load_npz("piv_crop512_u8.npz")
    write_pgm(tmp_path / "a.pgm", g["a"])
    write_pgm(tmp_path / "b.pgm", g["b"], comment=True)
    r = subprocess.run([str(PROCESS_BIN), "-s", str(size), "-f", ftype, str(tmp_path / "a.pgm"),
                        str(tmp_path / "b.pgm")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    few = (g[f"{label}_flags"] & 1) != 0
    check_output(r.stdout, [g[f"{label}_{k}"] for k in ("x", "y", "u", "v", "sn")], few)


def test_process_16bit_limit_search_and_overlap(tmp_path):
    a, b = oa.synth_pair(640, 480, seed=31, i0=(1000.0, 60000.0), noise=50.0, params=(1.3, -2.2), dtype=np.uint16)
    write_pgm(tmp_path / "a.pgm", a)
    write_pgm(tmp_path / "b.pgm", b)
    for args, limit, off in ((["-o", "0.75"], False, 24), (["-l"], True, 16), (["--limit-search", "-s", "32"], True, 16)):
        say(f"process 16-bit {args}")
        r = subprocess.run([str(PROCESS_BIN)] + args + [str(tmp_path / "a.pgm"), str(tmp_path / "b.pgm")],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        ref = orc.process_pair(a, b, 32, off, off, limit_search=limit, threads=8)
        check_output(r.stdout, [ref[k] for k in ("x", "y", "u", "v", "sn")], (ref["flags"] & 1) != 0)


def test_cpp_plan_adjoint_matches_bindings(dropin, tmp_path, gpu_ctx):
    """Plan::vjp / Plan::variance of the C++ layer give the C ABI's results
    (the same calls the ctypes bindings make)."""
    w = h = 256
    a, b = oa.synth_pair(w, h, seed=12, params=(1.6, -0.8), dtype=np.float32)
    plan = oa.Plan(gpu_ctx, w, h, 32, dtype=np.float32)
    cot = np.random.default_rng(3).normal(size=(plan.count, 2))
    for name, arr in (("a", a), ("b", b)):
        arr.tofile(tmp_path / f"{name}.bin")
    cot.tofile(tmp_path / "cot.bin")
    subprocess.run([str(dropin), "adjoint", str(w), str(h), "32", str(tmp_path / "a.bin"), str(tmp_path / "b.bin"),
                    str(tmp_path / "cot.bin"), str(tmp_path / "o.bin")], check=True, timeout=60)
    raw = (tmp_path / "o.bin").read_bytes()
    ga = np.frombuffer(raw[:w * h * 4], np.float32).reshape(h, w)
    gb = np.frombuffer(raw[w * h * 4:2 * w * h * 4], np.float32).reshape(h, w)
    var = np.frombuffer(raw[2 * w * h * 4:], np.float64).reshape(-1, 2)
    fwd = plan.run(a, b)
    ra, rb = plan.vjp(a, b, cot, fwd)
    assert np.array_equal(ga, ra) and np.array_equal(gb, rb)
    assert np.array_equal(var, plan.variance(a, b, 2.0, fwd))


def test_cpp_batch_matches_single_pair(dropin, tmp_path, gpu_ctx):
    """Batch of the C++ layer (one pair, passes 32 and 16) against single-pair
    plans: bit-identical records."""
    w, h = 320, 256
    a, b = oa.synth_pair(w, h, seed=13, params=(-1.2, 2.1), dtype=np.float32)
    a.tofile(tmp_path / "a.bin")
    b.tofile(tmp_path / "b.bin")
    subprocess.run([str(dropin), "batch", str(w), str(h), str(tmp_path / "a.bin"), str(tmp_path / "b.bin"),
                    str(tmp_path / "o.bin")], check=True, timeout=60)
    got = np.frombuffer((tmp_path / "o.bin").read_bytes(), oa.RESULT_DTYPE)
    ref = np.concatenate([oa.Plan(gpu_ctx, w, h, win, dtype=np.float32).run(a, b) for win in (32, 16)])
    assert got.tobytes() == ref.tobytes()


@pytest.mark.parametrize("direction", ["forward", "reverse"])
def test_hipfft_transform_matches_oracle(dropin, tmp_path, direction):
    rng = np.random.default_rng(7)
    z = rng.normal(size=(32, 32)) + 1j * rng.normal(size=(32, 32))
    z.astype(np.complex128).tofile(tmp_path / "z.bin")
    subprocess.run([str(dropin), "transform", "32", direction, str(tmp_path / "z.bin"), str(tmp_path / "o.bin")],
                   check=True, timeout=60)
    got = np.fromfile(tmp_path / "o.bin", dtype=np.complex128).reshape(32, 32)
    ref = orc.FFT(32).transform(z, forward=direction == "forward")
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()


@pytest.mark.parametrize("case", ["u16", "u8", "piv"])
def test_average_subtract_front_end_files(tmp_path, case):
    """opiv_average_subtract writes <input>.avg_sub.pnm files byte-identical to
    those of the reference's examples/average_subtract (golden outputs of its
    own loader, arithmetic and PNM writer); the second input is a TIFF."""
    g = np.load(Path(__file__).parent / "golden" / "avgsub_cases.npz")
    frames, outs = g[f"{case}_frames"], g[f"{case}_out"]
    paths = []
    for i, f in enumerate(frames):
        p = tmp_path / (f"f{i}.tif" if i == 1 else f"f{i}.pgm")
        (write_tiff(p, [f]) if i == 1 else write_pgm(p, f))
        paths.append(p)
    r = subprocess.run([str(AVGSUB_BIN)] + [str(p) for p in paths], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    h, w = frames.shape[1:]
    for p, o in zip(paths, outs):
        expect = b"P5\n# created by pnm_image_loader\n%d %d\n65535\n" % (w, h) + o.astype(">u2").tobytes()
        assert Path(str(p) + ".avg_sub.pnm").read_bytes() == expect


def test_average_subtract_front_end_needs_two_inputs(tmp_path):
    write_pgm(tmp_path / "a.pgm", np.zeros((8, 8), np.uint8))
    r = subprocess.run([str(AVGSUB_BIN), str(tmp_path / "a.pgm")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "require two or more input images" in r.stderr


def test_process_reads_tiff_like_pgm(tmp_path):
    """the process front-end on TIFF frames (big-endian, strips) gives the
    same output as on the same frames as PGM"""
    a, b = oa.synth_pair(320, 256, seed=33, i0=(1000.0, 60000.0), noise=50.0, params=(0.7, -1.9), dtype=np.uint16)
    write_pgm(tmp_path / "a.pgm", a)
    write_pgm(tmp_path / "b.pgm", b)
    write_tiff(tmp_path / "a.tif", [a], big_endian=True, rows_per_strip=16)
    write_tiff(tmp_path / "b.tif", [b], rows_per_strip=5)
    outs = []
    for ext in ("pgm", "tif"):
        r = subprocess.run([str(PROCESS_BIN), str(tmp_path / f"a.{ext}"), str(tmp_path / f"b.{ext}")],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        outs.append(r.stdout)
    assert outs[0] == outs[1] and len(outs[0]) > 0


def test_process_on_reference_example_tiffs():
    """process on the reference's own example pair, examples/process/
    exp1_001_a.tiff (= tests/corr_a.tiff) and exp1_001_b.tiff: RGB,
    Deflate-compressed, read as grey by the reference's integer rule; the
    output against the FP64 oracle on the same grey frames"""
    from conftest import GOLDEN
    fa, fb = GOLDEN / "corr_a.tiff", GOLDEN / "exp1_001_b.tiff"
    a, b = oa.read_image(fa), oa.read_image(fb)
    assert a.shape == b.shape == (369, 511) and a.dtype == np.uint8
    for size in (16, 32, 64):
        say(f"process example tiffs -s {size}")
        r = subprocess.run([str(PROCESS_BIN), "-s", str(size), str(fa), str(fb)], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stderr
        ref = orc.process_pair(a, b, size, size // 2, size // 2, threads=8)
        check_output(r.stdout, [ref[k] for k in ("x", "y", "u", "v", "sn")], (ref["flags"] & 1) != 0)


def test_read_image_into_device_memory(tmp_path, gpu_ctx):
    rng = np.random.default_rng(5)
    a = rng.integers(0, 65536, (90, 70)).astype(np.uint16)
    write_tiff(tmp_path / "a.tif", [a, a[::-1].copy()], rows_per_strip=9)
    pitch = 80
    buf = oa.DeviceBuffer(90 * pitch * 2)
    gpu_ctx.read_image_device(tmp_path / "a.tif", buf.ptr, pitch=pitch, index=1)
    got = buf.download(np.zeros((90, pitch), np.uint16))[:, :70]
    assert np.array_equal(got, a[::-1])


DROPIN_CORE = Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "dropin_core"


@pytest.mark.skipif(not DROPIN_CORE.exists(), reason="oracle/_ref/dropin_core is built where openpiv-cxx is mounted")
@pytest.mark.parametrize("win", [16, 32, 64])
def test_core_types_registry_dropin(tmp_path, win):
    """algos::HipFFT (include/openpiv_amd_core.hpp) in examples/process's
    correlator registry on the reference's own core::gf_image frames, with
    the reference's extract / find_peaks / fit_simple_gaussian and ThreadPool,
    against algos::FFT in the same binary: integer peaks equal, vectors within
    1e-3 px (complex and real variants); transform and auto_correlate within
    1e-5 of the largest magnitude; a size mismatch raises the reference's text"""
    say(f"dropin_core win {win}")
    a, b = oa.synth_pair(256, 192, seed=31 + win, params=(2.3, -1.7), dtype=np.float64)
    a.tofile(tmp_path / "a.bin")
    b.tofile(tmp_path / "b.bin")
    r = subprocess.run([str(DROPIN_CORE), "run", str(tmp_path / "a.bin"), str(tmp_path / "b.bin"), "256", "192",
                        str(win), str(win // 2), "4"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    f = r.stdout.split()
    vals = {f[i]: float(f[i + 1]) for i in range(0, len(f) - 1, 2)}
    say(r.stdout.strip())
    assert vals["windows"] > 0 and vals["mismatches"] == 0
    assert vals["max_err"] <= 1e-3 and vals["real_err"] <= 1e-3
    assert vals["transform_err"] <= 1e-5 and vals["auto_err"] <= 1e-5
    assert vals["size_error"] == 1
