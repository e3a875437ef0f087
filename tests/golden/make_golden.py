"""Write the golden vectors under tests/golden/ from the compiled reference.

Runs oracle/_ref/ref_piv (the reference's own headers, built by
oracle/Makefile from /root/reference) on seeded synthetic frames and on a
crop of the reference's bundled example frames, and stores inputs (or the
recipe + SHA-256 of synthetic inputs) and outputs as small .npz / .json
fixtures.  Only needed where /root/reference is mounted; the fixtures are
committed so tests never need the reference at run time.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "openpiv-cxx_amd"))
sys.path.insert(0, str(ROOT / "oracle"))
import openpiv_amd as oa  # noqa: E402  (synthetic frames only)
import oracle as orc  # noqa: E402

OUT = Path(__file__).resolve().parent
REF_FRAMES = Path("/root/reference/examples/process")

# config 1 of the benchmark definition: 1024^2 float64, particles d = 2.5 px,
# density 0.05 /px, I0 ~ U(100, 255), noise sigma 2, uniform shift (2.3, -1.7)
SYNTH_1024 = dict(w=1024, h=1024, seed=1, density=0.05, diameter=2.5, i0=(100.0, 255.0),
                  noise=2.0, field=oa.FIELD_UNIFORM, params=(2.3, -1.7), dtype="float64")
SYNTH_512 = dict(SYNTH_1024, w=512, h=512, seed=7, params=(-0.6, 3.4))

FIELDS = ["x", "y", "u", "v", "sn", "peak0", "peak1", "px", "py", "flags"]


def synth(spec):
    s = dict(spec)
    dt = np.dtype(s.pop("dtype"))
    w, h = s.pop("w"), s.pop("h")
    return oa.synth_pair(w, h, dtype=dt, threads=8, **s)


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def ref_process(tmp, a, b, win, step, variant="complex", limit=0):
    h, w = a.shape
    pa, pb, po = tmp / "a.bin", tmp / "b.bin", tmp / "o.bin"
    a.astype(np.float64).tofile(pa)
    b.astype(np.float64).tofile(pb)
    orc.run_reference(["process", pa, pb, w, h, win, step, step, variant, limit, 1, po])
    return np.fromfile(po, dtype=orc.REF_DTYPE)


def ref_planes(tmp, a, b, win, step, variant, idx):
    h, w = a.shape
    pa, pb, po = tmp / "a.bin", tmp / "b.bin", tmp / "p.bin"
    a.astype(np.float64).tofile(pa)
    b.astype(np.float64).tofile(pb)
    orc.run_reference(["planes", pa, pb, w, h, win, step, step, variant, po] + list(idx))
    return np.fromfile(po, dtype=np.float64).reshape(len(idx), win, win)


def records(prefix, rec):
    return {f"{prefix}_{k}": np.asarray(rec[k]) for k in FIELDS}


def grid_out(text: str) -> dict:
    """Grid output lines; long grids are kept as count + SHA-256 + ends."""
    lines = [l for l in text.split("\n") if l]
    if len(lines) <= 64:
        return {"out": lines}
    return {"count": len(lines), "sha256": hashlib.sha256(text.encode()).hexdigest(),
            "head": lines[:4], "tail": lines[-4:]}


def read_pgm(path: Path) -> np.ndarray:
    data = path.read_bytes()
    assert data[:2] == b"P5"
    parts = data.split(b"\n", 3)
    w, h = map(int, parts[1].split())
    assert int(parts[2]) == 255
    return np.frombuffer(parts[3][: w * h], dtype=np.uint8).reshape(h, w)


def main():
    if not orc.have_reference_harness():
        sys.exit("oracle/_ref/ref_piv is missing: run `make -C oracle` where /root/reference is mounted")
    tmp = Path(tempfile.mkdtemp())

    # ---- synthetic 1024^2 float64 pair (config 1) ----
    a, b = synth(SYNTH_1024)
    arrays = {}
    for name, win, step, var, lim in [("c32", 32, 16, "complex", 0), ("r32", 32, 16, "real", 0),
                                      ("l32", 32, 16, "complex", 1), ("c64", 64, 32, "complex", 0)]:
        arrays.update(records(name, ref_process(tmp, a, b, win, step, var, lim)))
    idx = [0, 1, 62, 63, 1984, 2000, 3905, 3968]
    arrays["plane_idx"] = np.array(idx, np.int32)
    arrays["planes_c32"] = ref_planes(tmp, a, b, 32, 16, "complex", idx)
    arrays["planes_r32"] = ref_planes(tmp, a, b, 32, 16, "real", idx)
    np.savez_compressed(OUT / "synth1024_f64.npz", **arrays)
    meta = {"synth": SYNTH_1024, "sha256_a": sha(a), "sha256_b": sha(b)}

    # ---- synthetic 512^2 pair at 16^2 windows ----
    a2, b2 = synth(SYNTH_512)
    np.savez_compressed(OUT / "synth512_f64_w16.npz",
                        **records("c16", ref_process(tmp, a2, b2, 16, 8, "complex", 0)),
                        **records("c8", ref_process(tmp, a2, b2, 8, 4, "complex", 0)))
    meta["synth512"] = SYNTH_512
    meta["synth512_sha256_a"] = sha(a2)
    meta["synth512_sha256_b"] = sha(b2)
    (OUT / "synth_meta.json").write_text(json.dumps(meta, indent=1))

    # ---- the reference's example frames (examples/process/piv1.pgm, piv2.pgm):
    # central 512^2 crop, 8-bit, as data ----
    f1 = read_pgm(REF_FRAMES / "piv1.pgm")[256:768, 256:768].copy()
    f2 = read_pgm(REF_FRAMES / "piv2.pgm")[256:768, 256:768].copy()
    np.savez_compressed(OUT / "piv_crop512_u8.npz", a=f1, b=f2,
                        **records("c32", ref_process(tmp, f1, f2, 32, 16, "complex", 0)),
                        **records("r32", ref_process(tmp, f1, f2, 32, 16, "real", 0)),
                        **records("c16", ref_process(tmp, f1, f2, 16, 8, "complex", 0)),
                        **records("c64", ref_process(tmp, f1, f2, 64, 32, "complex", 0)))

    # ---- FFT::transform known answers on seeded complex input ----
    rng = np.random.default_rng(11)
    kat = {}
    for n in (8, 16, 32):
        z = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        zi = tmp / "z.bin"
        np.ascontiguousarray(np.stack([z.real, z.imag], -1)).tofile(zi)
        kat[f"in{n}"] = z
        for d in (0, 1):
            orc.run_reference(["transform", n, d, zi, tmp / "t.bin"])
            t = np.fromfile(tmp / "t.bin", dtype=np.float64).reshape(n, n, 2)
            kat[f"out{n}_{d}"] = t[..., 0] + 1j * t[..., 1]
    np.savez_compressed(OUT / "fft_transform_kat.npz", **kat)

    # ---- generate_cartesian_grid, incl. error cases ----
    grid_cases = []
    for img, win, off in [((100, 50), (32, 32), (16, 16)), ((64, 64), (32, 32), (16, 16)),
                          ((65, 65), (32, 32), (16, 16)), ((66, 66), (32, 32), (16, 16)),
                          ((64, 64), (32, 32), (8, 8)), ((64, 64), (32, 32), (16, 8)),
                          ((64, 64), (32, 16), (16, 8)), ((4096, 4096), (32, 32), (16, 16)),
                          ((1280, 800), (64, 64), (32, 32)), ((1280, 800), (16, 16), (8, 8)),
                          ((31, 64), (32, 32), (16, 16)), ((64, 64), (32, 32), (0, 16)),
                          ((0, 64), (32, 32), (16, 16))]:
        text = orc.run_reference(["grid", *img, *win, *off]).stdout
        grid_cases.append({"img": img, "win": win, "off": off, **grid_out(text)})
    for img, win, frac in [((1024, 1024), (32, 32), 0.5), ((1024, 1024), (32, 32), 0.75),
                           ((100, 50), (32, 32), 0.5), ((64, 64), (32, 32), 1.5),
                           ((64, 64), (32, 32), -0.1), ((64, 64), (32, 32), 0.01)]:
        text = orc.run_reference(["gridf", *img, *win, frac]).stdout
        grid_cases.append({"img": img, "win": win, "frac": frac, **grid_out(text)})
    (OUT / "grid_cases.json").write_text(json.dumps(grid_cases))

    # ---- find_peaks on seeded planes (incl. exact ties) ----
    peak_cases = []
    for seed, (w, h) in enumerate([(32, 32), (16, 16), (40, 24), (100, 100)]):
        rng = np.random.default_rng(100 + seed)
        p = np.round(rng.random((h, w)) * 20.0)       # integer values: plenty of ties
        p.astype(np.float64).tofile(tmp / "pk.bin")
        case = {"w": w, "h": h, "plane": p.astype(int).tolist(), "out": {}}
        for num in (1, 2, 3, 5):
            r = orc.run_reference(["peaks", w, h, num, 1, tmp / "pk.bin"]).stdout.split("\n")
            case["out"][str(num)] = [l for l in r if l]
        peak_cases.append(case)
    (OUT / "find_peaks_cases.json").write_text(json.dumps(peak_cases))
    print("golden vectors written to", OUT)


if __name__ == "__main__":
    main()
