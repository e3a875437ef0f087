"""Writes tests/golden/loader_cases.npz: PNM files (P5 and P6, 8 and 16 bits,
comment lines) and the pixels the reference's own loader registry
(loaders/image_loader.cpp:23-57, pnm_image_loader.cpp) gives for them when it
loads them into a gf_image as examples/process does.  The reference is run
through oracle/_ref/ref_piv (built by oracle/Makefile from the reference
sources); run this script in the container where /root/reference exists:

    python3 tests/golden/make_loader_golden.py
"""
from __future__ import annotations

import subprocess
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF_BIN = HERE.parent.parent / "oracle" / "_ref" / "ref_piv"


def pnm_bytes(kind: int, pixels: np.ndarray, comment: bool) -> bytes:
    h, w = pixels.shape[:2]
    wide = pixels.dtype == np.uint16
    head = f"P{kind}\n" + ("# a comment line\n# and another\n" if comment else "") + f"{w} {h}\n"
    head += "65535\n" if wide else "255\n"
    data = pixels.astype(">u2").tobytes() if wide else pixels.astype(np.uint8).tobytes()
    return head.encode() + data


def main() -> None:
    rng = np.random.default_rng(11)
    cases = {
        "p5_u8": (5, rng.integers(0, 256, (13, 17)).astype(np.uint8), False),
        "p5_u16": (5, rng.integers(0, 65536, (9, 11)).astype(np.uint16), True),
        # bytes across the whole range: the reference reads P6 8-bit samples
        # through a signed char
        "p6_u8": (6, rng.integers(0, 256, (12, 19, 3)).astype(np.uint8), True),
        "p6_u16": (6, rng.integers(0, 65536, (7, 10, 3)).astype(np.uint16), False),
    }
    out = {}
    with tempfile.TemporaryDirectory() as d:
        for name, (kind, px, comment) in cases.items():
            f = Path(d) / f"{name}.pnm"
            raw = pnm_bytes(kind, px, comment)
            f.write_bytes(raw)
            res = subprocess.run([str(REF_BIN), "load", str(f), str(Path(d) / "o.bin")], capture_output=True,
                                 text=True, check=True)
            w, h = map(int, res.stdout.split())
            ref = np.fromfile(Path(d) / "o.bin", dtype=np.float64).reshape(h, w)
            out[f"{name}_file"] = np.frombuffer(raw, dtype=np.uint8)
            out[f"{name}_ref"] = ref
    np.savez_compressed(HERE / "loader_cases.npz", **out)
    print("wrote", HERE / "loader_cases.npz", sorted(out))


if __name__ == "__main__":
    main()
