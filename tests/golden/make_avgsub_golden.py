"""Write tests/golden/avgsub_cases.npz from the compiled reference.

Runs the average-subtract pipeline of examples/average_subtract/main.cpp:98-159
through oracle/_ref/ref_piv avgsub (the reference's PNM loader, image
arithmetic, algos::find_image_range and PNM writer) on seeded frame stacks
and on a crop of the reference's example frames, and stores the inputs, the
mean and the 16-bit images the writer saves.  Only needed where
/root/reference is mounted; the fixture is committed.

    python tests/golden/make_avgsub_golden.py
"""
from __future__ import annotations

import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "tests"))
import oracle as orc  # noqa: E402
from dropin_util import write_pgm  # noqa: E402

OUT = Path(__file__).resolve().parent
REF_FRAMES = Path("/root/reference/examples/process")
HEADER = b"P5\n# created by pnm_image_loader\n%d %d\n65535\n"


def read_pgm(path: Path) -> np.ndarray:
    data = path.read_bytes()
    parts, pos = [], 0
    while len(parts) < 4:
        while data[pos:pos + 1].isspace():
            pos += 1
        if data[pos:pos + 1] == b"#":
            pos = data.index(b"\n", pos) + 1
            continue
        end = pos
        while not data[end:end + 1].isspace():
            end += 1
        parts.append(data[pos:end])
        pos = end
    pos += 1
    w, h, maxval = int(parts[1]), int(parts[2]), int(parts[3])
    dt = np.uint8 if maxval < 256 else np.dtype(">u2")
    return np.frombuffer(data[pos:], dtype=dt, count=w * h).reshape(h, w).astype(
        np.uint8 if maxval < 256 else np.uint16)


def run_case(tmp: Path, name: str, frames: np.ndarray) -> dict:
    paths = []
    for i, f in enumerate(frames):
        p = tmp / f"{name}_{i}.pgm"
        write_pgm(p, f)
        paths.append(p)
    mean, pnms = orc.reference_average_subtract(paths, str(tmp / f"{name}_out_"))
    n, h, w = frames.shape
    head = HEADER % (w, h)
    outs = []
    for b in pnms:
        assert b.startswith(head), b[:64]
        outs.append(np.frombuffer(b[len(head):], dtype=">u2").reshape(h, w).astype(np.uint16))
    return {f"{name}_frames": frames, f"{name}_mean": mean.reshape(h, w), f"{name}_out": np.stack(outs)}


def main():
    if not orc.have_reference_harness():
        sys.exit("oracle/_ref/ref_piv is missing: run `make -C oracle` where /root/reference is mounted")
    tmp = Path(tempfile.mkdtemp())
    rng = np.random.default_rng(2024)
    y, x = np.mgrid[0:48, 0:64]
    background = 3000 + 40 * x + 25 * y
    u16 = np.stack([np.clip(background + rng.normal(0, 900, (48, 64)) * (rng.random((48, 64)) < 0.3), 0, 65535)
                    for _ in range(5)]).astype(np.uint16)
    u8 = rng.integers(0, 256, (4, 40, 56)).astype(np.uint8)
    pair = rng.integers(0, 65536, (2, 24, 32)).astype(np.uint16)                # n = 2: mean = frame 1
    same = np.repeat(rng.integers(0, 256, (1, 16, 16)).astype(np.uint8), 3, 0)  # d = 0: range 1
    f1 = read_pgm(REF_FRAMES / "piv1.pgm")[200:264, 300:380]
    f2 = read_pgm(REF_FRAMES / "piv2.pgm")[200:264, 300:380]
    arrays = {}
    for name, frames in (("u16", u16), ("u8", u8), ("pair", pair), ("same", same),
                         ("piv", np.stack([f1, f2, f1]))):
        arrays.update(run_case(tmp, name, frames))
    np.savez_compressed(OUT / "avgsub_cases.npz", **arrays)
    print("wrote", OUT / "avgsub_cases.npz", sum(v.nbytes for v in arrays.values()), "bytes")


if __name__ == "__main__":
    main()
