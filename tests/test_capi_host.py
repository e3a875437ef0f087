"""C ABI checks that need no GPU: the library loads, exports exactly what
include/openpiv_amd.h declares, its struct layouts match the bindings, and
its host-only grid functions agree with the reference's grid goldens."""
import ctypes
import re
import subprocess
import tempfile
from pathlib import Path

import numpy as np
import pytest

import openpiv_amd as oa
from conftest import ROOT, load_json

HEADER = ROOT / "include" / "openpiv_amd.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(opiv_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = oa.load()
    names = declared_functions()
    assert names, "no declarations parsed"
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(oa.EXPORTS) == names
    assert oa.version().startswith("openpiv-amd")


def test_struct_layouts_match_bindings():
    src = r"""
    #include <stdio.h>
    #include <stddef.h>
    #include "openpiv_amd.h"
    int main(void) {
      printf("%zu %zu %zu %zu %zu %zu\n", sizeof(opiv_window_result), offsetof(opiv_window_result, px),
             offsetof(opiv_window_result, flags), sizeof(opiv_plan_desc), offsetof(opiv_plan_desc, overlap),
             offsetof(opiv_plan_desc, limit_search));
      return 0; }
    """
    with tempfile.TemporaryDirectory() as d:
        c = Path(d) / "layout.c"
        c.write_text(src)
        exe = Path(d) / "layout"
        subprocess.run(["gcc", "-I", str(ROOT / "include"), str(c), "-o", str(exe)], check=True)
        vals = list(map(int, subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()))
    assert vals[0] == oa.RESULT_DTYPE.itemsize
    assert vals[1] == oa.RESULT_DTYPE.fields["px"][1]
    assert vals[2] == oa.RESULT_DTYPE.fields["flags"][1]
    assert vals[3] == ctypes.sizeof(oa.PlanDesc)
    assert vals[4] == oa.PlanDesc.overlap.offset
    assert vals[5] == oa.PlanDesc.limit_search.offset


def test_grid_matches_reference_goldens():
    for case in load_json("grid_cases.json"):
        (iw, ih), (ww, wh) = case["img"], case["win"]
        try:
            if "frac" in case:
                ox, oy = oa.grid_offsets(ww, wh, case["frac"])
            else:
                ox, oy = case["off"]
            xy = oa.grid(iw, ih, ww, wh, ox, oy)
        except oa.OpivError as e:
            assert e.code == 4
            assert case["out"] == [f"error {e}"], case
            continue
        lines = [f"{x} {y}" for x, y in xy]
        if "out" in case:
            assert lines == case["out"]
        else:
            assert len(lines) == case["count"] and lines[:4] == case["head"] and lines[-4:] == case["tail"]


def test_context_without_device_fails_cleanly():
    """On a host without a gfx950 device, creating a context reports
    OPIV_E_NO_DEVICE instead of crashing; with one, it succeeds."""
    try:
        ctx = oa.Context(0)
    except oa.OpivError as e:
        assert e.code == 7 and "device" in str(e)
    else:
        ctx.close()


def test_oracle_is_not_linked_by_the_product():
    """The product library must not depend on the test oracle."""
    out = subprocess.run(["ldd", str(oa.library_path())], capture_output=True, text=True).stdout
    assert "oracle" not in out
    text = " ".join(p.read_text() for p in (ROOT / "openpiv-cxx_amd" / "openpiv_amd").glob("*.py"))
    assert "import oracle" not in text and "liboracle" not in text


def test_batch_entry_points_reject_null_handles():
    """the batch entry points check their arguments before any HIP call, so
    a null handle is an OPIV_E_INVALID error with a message, on any host"""
    lib = oa.load()
    i32 = ctypes.c_int32
    a, b = i32(), i32()
    n = ctypes.c_size_t()
    for rc, what in [(lib.opiv_batch_form(None, ctypes.byref(a), ctypes.byref(b)), "opiv_batch_form"),
                     (lib.opiv_batch_shape(None, 0, ctypes.byref(a), ctypes.byref(b)), "opiv_batch_shape"),
                     (lib.opiv_batch_window_count(None, ctypes.byref(n)), "opiv_batch_window_count"),
                     (lib.opiv_batch_run(None, None, 0), "opiv_batch_run")]:
        assert rc == 1, (what, rc)
        assert oa.last_error(), what
    assert lib.opiv_batch_destroy(None) == 0


def test_fit_logarithm_against_extended_precision(tmp_path):
    """The Gaussian fit's FP64 logarithm of FP32 plane values (fit_log.hpp,
    compiled into the kernels) against the x87 extended-precision logl:
    every float of [1/8, 8), a stride over the whole positive range and the
    special values; at most 1.5 ulp."""
    exe = tmp_path / "fit_log_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-I", str(ROOT / "openpiv-cxx_amd" / "csrc"),
                    str(ROOT / "tests" / "cpp" / "fit_log_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and out.stdout.strip().endswith("ok"), out.stdout


def test_grid_width_division_matches_divide(tmp_path):
    """The kernels find a window's grid row with geom.hpp's multiplier
    division; checked against the divide instruction on the host."""
    exe = tmp_path / "fastdiv_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-I", str(ROOT / "openpiv-cxx_amd" / "csrc"),
                    str(ROOT / "tests" / "cpp" / "fastdiv_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout
