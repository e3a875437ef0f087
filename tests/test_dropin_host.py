"""C++ drop-in layer without a GPU: the header compiles against an image type
with the reference's accessors, and the front-end's argument and loader
errors behave like examples/process (main.cpp:85-132)."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import ROOT
from dropin_util import PROCESS_BIN, build_dropin_check, write_pgm


def run(args):
    return subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=60)


def test_header_compiles_and_reports_missing_device(tmp_path):
    exe = build_dropin_check(tmp_path)
    out = run([exe, "nodevice"]).stdout
    assert out.strip() == "ok" or out.startswith("error 7 ")


def test_process_requires_two_inputs(tmp_path):
    r = run([PROCESS_BIN, "-s", "32", tmp_path / "a.pgm"])
    assert r.returncode == 1 and "require two input images" in r.stderr


def test_process_missing_file(tmp_path):
    r = run([PROCESS_BIN, tmp_path / "a.pgm", tmp_path / "b.pgm"])
    assert r.returncode == 1 and "failed to open" in r.stderr


def test_process_rejects_size_mismatch_and_bad_type(tmp_path):
    write_pgm(tmp_path / "a.pgm", np.zeros((64, 64), np.uint8))
    write_pgm(tmp_path / "b.pgm", np.zeros((64, 32), np.uint8), comment=True)
    r = run([PROCESS_BIN, tmp_path / "a.pgm", tmp_path / "b.pgm"])
    assert r.returncode == 1 and "image sizes don't match" in r.stderr
    write_pgm(tmp_path / "b.pgm", np.zeros((64, 64), np.uint8), comment=True)
    r = run([PROCESS_BIN, "-f", "fftw", tmp_path / "a.pgm", tmp_path / "b.pgm"])
    assert r.returncode == 1 and "unknown fft type" in r.stderr
    r = run([PROCESS_BIN, "--bogus", tmp_path / "a.pgm", tmp_path / "b.pgm"])
    assert r.returncode == 1 and "error parsing options" in r.stderr


REF_DIR = Path("/root/reference")
DROPIN_CORE = ROOT / "oracle" / "_ref" / "dropin_core"


@pytest.mark.skipif(not (REF_DIR / "openpiv" / "core" / "image.h").exists(),
                    reason="openpiv-cxx headers are not mounted here")
def test_core_types_dropin_compiles_and_checks_sizes():
    """include/openpiv_amd_core.hpp compiled against the reference's own
    core::size / core::gf_image / core::cf_image, with algos::HipFFT in the
    correlator registry of main.cpp:153-186 beside algos::FFT; a
    non-power-of-two size raises the reference's message before any device
    is opened"""
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "_ref/dropin_core"], check=True, timeout=600)
    r = run([DROPIN_CORE, "check"])
    assert r.returncode == 0 and r.stdout.strip() == "check ok", r.stdout + r.stderr


def test_front_ends_fail_cleanly_on_huge_declared_frames(tmp_path):
    """A 90-byte Deflate TIFF that declares a 2^31 x 2^31 16-bit frame: the
    loader refuses the size, and both front-ends report it once ("failed to
    load image: ...", as examples/average_subtract/main.cpp:123 words it)
    and exit 1; a frame buffer that cannot be allocated would end the same
    way (both catch std::exception)"""
    import struct
    import zlib
    from dropin_util import AVGSUB_BIN
    w = h = 2 ** 31
    data = zlib.compress(bytes(64))
    entries = [(256, 4, 1, w), (257, 4, 1, h), (258, 3, 1, 16), (259, 3, 1, 8), (262, 3, 1, 1),
               (273, 4, 1, 8 + 2 + 9 * 12 + 4), (277, 3, 1, 1), (278, 4, 1, h), (279, 4, 1, len(data))]
    ifd = struct.pack("<H", len(entries))
    for t, ty, c, v in entries:
        ifd += struct.pack("<HHIHH", t, ty, c, v, 0) if ty == 3 else struct.pack("<HHII", t, ty, c, v)
    ifd += struct.pack("<I", 0)
    path = tmp_path / "huge.tif"
    path.write_bytes(b"II\x2a\x00" + struct.pack("<I", 8) + ifd + data)
    for args in ([PROCESS_BIN, path, path], [AVGSUB_BIN, path, path]):
        r = run(args)
        assert r.returncode == 1, (args[0].name, r.returncode, r.stderr[-300:])
        assert r.stderr.count("failed to load image: ") == 1 and "too large" in r.stderr, r.stderr
