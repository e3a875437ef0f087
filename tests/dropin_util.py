"""Helpers for the C++ drop-in tests: build the test driver, write PGM frames."""
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
PROCESS_BIN = ROOT / "openpiv-cxx_amd" / "lib" / "opiv_process"


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
