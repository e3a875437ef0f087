"""ctypes access to the FP64 oracle (liboracle.so) and the reference harness.

TEST INFRASTRUCTURE ONLY: the checker the HIP engine is compared against.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
LIB_PATH = ORACLE_DIR / "liboracle.so"
REF_BIN = ORACLE_DIR / "_ref" / "ref_piv"

U8, U16, F32, F64 = 0, 1, 2, 3
CORR_COMPLEX, CORR_REAL = 0, 1
FLAG_FEWER_THAN_2_PEAKS, FLAG_NONFINITE_FIT, FLAG_EXACT_TIE = 1, 2, 8

VECTOR_DTYPE = np.dtype([
    ("x", "<f8"), ("y", "<f8"), ("u", "<f8"), ("v", "<f8"), ("sn", "<f8"),
    ("px", "<i4"), ("py", "<i4"), ("peak0", "<f8"), ("peak1", "<f8"),
    ("flags", "<u4"), ("pad", "<u4"),
])
# record written by ref_piv process
REF_DTYPE = np.dtype([
    ("x", "<f8"), ("y", "<f8"), ("u", "<f8"), ("v", "<f8"), ("sn", "<f8"),
    ("peak0", "<f8"), ("peak1", "<f8"),
    ("px", "<i4"), ("py", "<i4"), ("flags", "<u4"), ("pad", "<i4"),
])
_DT = {np.dtype(np.uint8): U8, np.dtype(np.uint16): U16, np.dtype(np.float32): F32, np.dtype(np.float64): F64}


class Job(ctypes.Structure):
    _fields_ = [
        ("a", ctypes.c_void_p), ("b", ctypes.c_void_p), ("dtype", ctypes.c_int),
        ("img_w", ctypes.c_uint32), ("img_h", ctypes.c_uint32), ("pitch", ctypes.c_size_t),
        ("win", ctypes.c_uint32), ("off_x", ctypes.c_uint32), ("off_y", ctypes.c_uint32),
        ("variant", ctypes.c_int), ("limit_search", ctypes.c_int), ("nthreads", ctypes.c_int),
        ("first", ctypes.c_size_t), ("count", ctypes.c_size_t),
    ]


class Peak(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int32), ("y", ctypes.c_int32), ("value", ctypes.c_double)]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise OSError(f"{LIB_PATH} missing: run `make -C {ORACLE_DIR}`")
        lib = ctypes.CDLL(str(LIB_PATH))
        vp = ctypes.c_void_p
        u32 = ctypes.c_uint32
        lib.oref_grid_offsets_from_fraction.argtypes = [u32, u32, ctypes.c_double, ctypes.POINTER(u32), ctypes.POINTER(u32)]
        lib.oref_grid.argtypes = [u32, u32, u32, u32, u32, u32, ctypes.POINTER(u32), ctypes.POINTER(u32), vp]
        lib.oref_fft_create.argtypes = [u32, u32]
        lib.oref_fft_create.restype = vp
        lib.oref_fft_destroy.argtypes = [vp]
        lib.oref_fft_transform.argtypes = [vp, vp, vp, ctypes.c_int]
        lib.oref_fft_transform_real.argtypes = [vp, vp, vp, vp, vp]
        lib.oref_cross_correlate.argtypes = [vp, vp, vp, vp]
        lib.oref_cross_correlate_real.argtypes = [vp, vp, vp, vp]
        lib.oref_auto_correlate.argtypes = [vp, vp, vp]
        lib.oref_swap_quadrants.argtypes = [vp, u32, u32]
        lib.oref_find_peaks.argtypes = [vp, u32, u32, u32, u32, u32, u32, ctypes.c_uint16, u32,
                                        ctypes.POINTER(Peak), ctypes.POINTER(ctypes.c_int)]
        lib.oref_find_peaks.restype = ctypes.c_size_t
        lib.oref_fit_simple_gaussian.argtypes = [vp, u32, ctypes.c_int32, ctypes.c_int32,
                                                 ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        lib.oref_process_pair.argtypes = [ctypes.POINTER(Job), vp]
        lib.oref_process_pair.restype = ctypes.c_long
        lib.oref_window_jacobian.argtypes = [vp, vp, vp, u32, ctypes.c_int, vp, vp, vp, vp]
        lib.oref_window_jacobian.restype = ctypes.c_int
        lib.oref_vjp_pair.argtypes = [ctypes.POINTER(Job), vp, vp, vp]
        lib.oref_vjp_pair.restype = ctypes.c_long
        lib.oref_variance_pair.argtypes = [ctypes.POINTER(Job), ctypes.c_double, vp]
        lib.oref_average_subtract.argtypes = [vp, ctypes.c_size_t, u32, u32, vp, vp]
        lib.oref_variance_pair.restype = ctypes.c_long
        lib.oref_process_window.argtypes = [vp, vp, vp, u32, ctypes.c_int32, ctypes.c_int32, u32, ctypes.c_int,
                                            ctypes.c_int, vp, vp]
        _lib = lib
    return _lib


def grid_offsets(win_w, win_h, fraction):
    ox, oy = ctypes.c_uint32(), ctypes.c_uint32()
    rc = load().oref_grid_offsets_from_fraction(win_w, win_h, fraction, ctypes.byref(ox), ctypes.byref(oy))
    return rc, ox.value, oy.value


def grid(img_w, img_h, win_w, win_h, off_x, off_y):
    """(rc, xy[n, 2])"""
    lib = load()
    nx, ny = ctypes.c_uint32(), ctypes.c_uint32()
    rc = lib.oref_grid(img_w, img_h, win_w, win_h, off_x, off_y, ctypes.byref(nx), ctypes.byref(ny), None)
    if rc:
        return rc, np.zeros((0, 2), np.int32)
    xy = np.zeros((nx.value * ny.value, 2), np.int32)
    lib.oref_grid(img_w, img_h, win_w, win_h, off_x, off_y, ctypes.byref(nx), ctypes.byref(ny), xy.ctypes.data)
    return 0, xy


class FFT:
    """algos::FFT restated in FP64 (fft.h)."""

    def __init__(self, w, h=None):
        h = w if h is None else h
        self.w, self.h = w, h
        self.handle = load().oref_fft_create(w, h)
        if not self.handle:
            raise RuntimeError(f"dimensions must be power of 2: [{w},{h}]")

    def __del__(self):
        if getattr(self, "handle", None):
            load().oref_fft_destroy(self.handle)
            self.handle = None

    def transform(self, z: np.ndarray, forward: bool = True) -> np.ndarray:
        z = np.ascontiguousarray(z, dtype=np.complex128)
        if z.shape != (self.h, self.w):
            raise RuntimeError(f"image size is different from expected: [{z.shape[1]},{z.shape[0]}], [{self.w},{self.h}]")
        out = np.empty_like(z)
        load().oref_fft_transform(self.handle, z.ctypes.data, out.ctypes.data, 0 if forward else 1)
        return out

    def transform_real(self, a, b):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        oa = np.empty(a.shape, np.complex128)
        ob = np.empty(a.shape, np.complex128)
        load().oref_fft_transform_real(self.handle, a.ctypes.data, b.ctypes.data, oa.ctypes.data, ob.ctypes.data)
        return oa, ob

    def cross_correlate(self, a, b):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        out = np.empty_like(a)
        load().oref_cross_correlate(self.handle, a.ctypes.data, b.ctypes.data, out.ctypes.data)
        return out

    def cross_correlate_real(self, a, b):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        out = np.empty_like(a)
        load().oref_cross_correlate_real(self.handle, a.ctypes.data, b.ctypes.data, out.ctypes.data)
        return out

    def auto_correlate(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        out = np.empty_like(a)
        load().oref_auto_correlate(self.handle, a.ctypes.data, out.ctypes.data)
        return out


def swap_quadrants(im: np.ndarray) -> np.ndarray:
    im = np.ascontiguousarray(im, dtype=np.float64).copy()
    load().oref_swap_quadrants(im.ctypes.data, im.shape[1], im.shape[0])
    return im


def find_peaks(plane, num_peaks=2, radius=1, region=None):
    """[(x, y, value)], exact_tie  -- region = (rx, ry, rw, rh) view or None."""
    plane = np.ascontiguousarray(plane, dtype=np.float64)
    h, w = plane.shape
    rx, ry, rw, rh = region if region else (0, 0, w, h)
    out = (Peak * max(1, num_peaks))()
    tie = ctypes.c_int()
    n = load().oref_find_peaks(plane.ctypes.data, w, h, rx, ry, rw, rh, num_peaks, radius, out, ctypes.byref(tie))
    return [(out[i].x, out[i].y, out[i].value) for i in range(n)], bool(tie.value)


def fit_simple_gaussian(plane, cx, cy):
    plane = np.ascontiguousarray(plane, dtype=np.float64)
    fx, fy = ctypes.c_double(), ctypes.c_double()
    load().oref_fit_simple_gaussian(plane.ctypes.data, plane.shape[1], cx, cy, ctypes.byref(fx), ctypes.byref(fy))
    return fx.value, fy.value


def process_pair(a, b, win, off_x, off_y, *, variant=CORR_COMPLEX, limit_search=False, threads=0,
                 first=0, count=0):
    """Per-window results of examples/process/main.cpp:189-237 in grid order (FP64)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, dtype=a.dtype)
    h, w = a.shape
    rc, xy = grid(w, h, win, win, off_x, off_y)
    if rc:
        raise ValueError(f"grid error {rc}")
    total = len(xy)
    n = (min(count, total - first) if count else total - first)
    out = np.zeros(max(n, 0), dtype=VECTOR_DTYPE)
    job = Job(a.ctypes.data, b.ctypes.data, _DT[a.dtype], w, h, w, win, off_x, off_y, variant,
              1 if limit_search else 0, threads or (os.cpu_count() or 1), first, count)
    r = load().oref_process_pair(ctypes.byref(job), out.ctypes.data)
    if r < 0:
        raise ValueError(f"oref_process_pair failed: {r}")
    return out


# ---------------------------------------------------------------- reference harness

def have_reference_harness() -> bool:
    return REF_BIN.exists()


def run_reference(args, **kw):
    return subprocess.run([str(REF_BIN)] + [str(x) for x in args], check=True,
                          capture_output=True, text=True, **kw)


# ---------------------------------------------------------------- adjoint (FP64)

def window_vector(wa, wb, limit_search=False):
    """(u, v, flags) of one window at origin 0 (main.cpp:189-237 in FP64)."""
    wa = np.ascontiguousarray(wa, dtype=np.float64)
    wb = np.ascontiguousarray(wb, dtype=np.float64)
    w = wa.shape[0]
    f = FFT(w)
    plane = np.empty_like(wa)
    out = np.zeros(1, dtype=VECTOR_DTYPE)
    load().oref_process_window(f.handle, wa.ctypes.data, wb.ctypes.data, w, 0, 0, w, 0,
                               1 if limit_search else 0, plane.ctypes.data, out.ctypes.data)
    return out[0]


def window_jacobian(wa, wb, limit_search=False):
    """(ju_a, ju_b, jv_a, jv_b) FP64 W x W arrays, or None without a valid fit."""
    wa = np.ascontiguousarray(wa, dtype=np.float64)
    wb = np.ascontiguousarray(wb, dtype=np.float64)
    w = wa.shape[0]
    f = FFT(w)
    j = np.zeros((4, w, w))
    rc = load().oref_window_jacobian(f.handle, wa.ctypes.data, wb.ctypes.data, w, 1 if limit_search else 0,
                                     j[0].ctypes.data, j[1].ctypes.data, j[2].ctypes.data, j[3].ctypes.data)
    return None if rc else j


def _job(a, b, win, off_x, off_y, limit_search):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, dtype=a.dtype)
    h, w = a.shape
    return a, b, Job(a.ctypes.data, b.ctypes.data, _DT[a.dtype], w, h, w, win, off_x, off_y, 0,
                     1 if limit_search else 0, 1, 0, 0)


def vjp_pair(a, b, win, off_x, off_y, cot, limit_search=False):
    a, b, job = _job(a, b, win, off_x, off_y, limit_search)
    cot = np.ascontiguousarray(cot, dtype=np.float64)
    ga = np.zeros(a.shape)
    gb = np.zeros(a.shape)
    if load().oref_vjp_pair(ctypes.byref(job), cot.ctypes.data, ga.ctypes.data, gb.ctypes.data) < 0:
        raise ValueError("oref_vjp_pair failed")
    return ga, gb


def variance_pair(a, b, win, off_x, off_y, sigma, limit_search=False):
    a, b, job = _job(a, b, win, off_x, off_y, limit_search)
    _, xy = grid(a.shape[1], a.shape[0], win, win, off_x, off_y)
    var = np.zeros((len(xy), 2))
    if load().oref_variance_pair(ctypes.byref(job), sigma, var.ctypes.data) < 0:
        raise ValueError("oref_variance_pair failed")
    return var


# ---------------------------------------------------------------- background subtraction

def average_subtract(frames):
    """examples/average_subtract/main.cpp:98-159 on n stacked frames (any
    numeric dtype, values as loaded): (mean f64 (h, w), out uint16 (n, h, w))."""
    f = np.ascontiguousarray(frames, dtype=np.float64)
    n, h, w = f.shape
    mean = np.zeros((h, w))
    out = np.zeros((n, h, w), np.uint16)
    if load().oref_average_subtract(f.ctypes.data, n, w, h, mean.ctypes.data, out.ctypes.data) != 0:
        raise ValueError("average_subtract needs two or more frames")
    return mean, out


def reference_average_subtract(pgm_paths, out_prefix):
    """The reference's own pipeline (loader, image arithmetic, find_image_range,
    PNM writer) through oracle/_ref/ref_piv avgsub: (mean, [PNM bytes per frame])."""
    run_reference(["avgsub", out_prefix] + [str(p) for p in pgm_paths])
    mean = np.fromfile(f"{out_prefix}mean.bin", dtype=np.float64)
    return mean, [Path(f"{out_prefix}{i}.pnm").read_bytes() for i in range(len(pgm_paths))]
