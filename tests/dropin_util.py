"""Helpers for the C++ drop-in tests: build the test driver, write PGM frames."""
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
PROCESS_BIN = ROOT / "openpiv-cxx_amd" / "lib" / "opiv_process"
AVGSUB_BIN = ROOT / "openpiv-cxx_amd" / "lib" / "opiv_average_subtract"


def build_dropin_check(tmp: Path) -> Path:
    exe = tmp / "dropin_check"
    lib = ROOT / "openpiv-cxx_amd" / "lib"
    rocm = Path("/opt/rocm")
    subprocess.run(["g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", "-I", str(ROOT / "include"),
                    "-I", str(rocm / "include"),
                    str(ROOT / "tests" / "cpp" / "dropin_check.cpp"), "-o", str(exe),
                    f"-L{lib}", "-lopenpiv_amd", f"-Wl,-rpath,{lib}",
                    f"-L{rocm / 'lib'}", "-lamdhip64", f"-Wl,-rpath,{rocm / 'lib'}"], check=True)
    return exe


def write_pgm(path: Path, a: np.ndarray, comment: bool = False) -> None:
    h, w = a.shape
    maxval = 255 if a.dtype == np.uint8 else 65535
    head = "P5\n" + ("# written by the drop-in tests\n" if comment else "") + f"{w} {h}\n{maxval}\n"
    data = a.astype(">u2").tobytes() if a.dtype == np.uint16 else a.astype(np.uint8).tobytes()
    path.write_bytes(head.encode() + data)


def parse_process_output(text: str) -> np.ndarray:
    rows = [list(map(float, l.split(","))) for l in text.strip().split("\n") if l.strip()]
    return np.array(rows, dtype=np.float64)


def write_tiff(path: Path, pages, big_endian: bool = False, rows_per_strip: int = 0) -> None:
    """Baseline TIFF writer for the loader tests: uncompressed greyscale
    uint8/uint16 pages, strips of rows_per_strip rows (0 = one strip), in
    either byte order, every page its own IFD."""
    import struct
    e = ">" if big_endian else "<"
    out = bytearray(b"MM\x00\x2a" if big_endian else b"II\x2a\x00")
    out += struct.pack(e + "I", 0)                      # first IFD offset, patched below
    prev_link = 4
    for img in pages:
        img = np.ascontiguousarray(img)
        h, w = img.shape
        bps = 16 if img.dtype == np.uint16 else 8
        rps = rows_per_strip or h
        strips = [img[r:r + rps] for r in range(0, h, rps)]
        offsets, counts = [], []
        for s in strips:
            data = s.astype(e + ("u2" if bps == 16 else "u1")).tobytes()
            if len(out) % 2:
                out += b"\x00"
            offsets.append(len(out))
            counts.append(len(data))
            out += data
        # arrays of offsets/counts when there is more than one strip
        def arr(vals):
            nonlocal out
            if len(vals) == 1:
                return vals[0]
            if len(out) % 2:
                out += b"\x00"
            at = len(out)
            out += struct.pack(e + "%dI" % len(vals), *vals)
            return at
        off_v, cnt_v = arr(offsets), arr(counts)
        entries = [(256, 4, 1, w), (257, 4, 1, h), (258, 3, 1, bps), (259, 3, 1, 1), (262, 3, 1, 1),
                   (273, 4, len(offsets), off_v), (277, 3, 1, 1), (278, 4, 1, rps), (279, 4, len(counts), cnt_v)]
        if len(out) % 2:
            out += b"\x00"
        ifd = len(out)
        struct.pack_into(e + "I", out, prev_link, ifd)
        out += struct.pack(e + "H", len(entries))
        for tag, typ, cnt, val in entries:
            if typ == 3 and cnt == 1:
                out += struct.pack(e + "HHIHH", tag, typ, cnt, val, 0)
            else:
                out += struct.pack(e + "HHII", tag, typ, cnt, val)
        prev_link = len(out)
        out += struct.pack(e + "I", 0)
    Path(path).write_bytes(bytes(out))
