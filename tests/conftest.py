"""Shared test set-up: import paths, the `gpu` marker and golden fixtures."""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(ROOT / "openpiv-cxx_amd"))
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def load_npz(name: str) -> dict:
    with np.load(GOLDEN / name, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def load_json(name: str):
    return json.loads((GOLDEN / name).read_text())


def synth_from_meta(key: str = "synth"):
    """Regenerate a golden synthetic pair from its recorded recipe."""
    import openpiv_amd as oa

    meta = load_json("synth_meta.json")
    s = dict(meta[key])
    dt = np.dtype(s.pop("dtype"))
    w, h = s.pop("w"), s.pop("h")
    s["i0"] = tuple(s["i0"])
    s["params"] = tuple(s["params"])
    return oa.synth_pair(w, h, dtype=dt, **s)


@pytest.fixture(scope="session")
def gpu_ctx():
    import openpiv_amd as oa

    ctx = oa.Context(0)
    yield ctx
    ctx.close()
