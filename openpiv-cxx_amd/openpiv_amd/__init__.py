"""Python bindings (ctypes) of the openpiv-amd C ABI (include/openpiv_amd.h).

The engine itself is the HIP library ``lib/libopenpiv_amd.so``; this module
only marshals arguments.  There is no CPU fallback: every compute call goes
through the gfx950 kernels, and loading fails loudly when the library has not
been built.

Reference interface mirrored (ErichZimmer/openpiv-cxx 0.2.0):
  * ``grid`` / ``grid_offsets``   -> core::generate_cartesian_grid (core/detail/grid.impl.h:10-61)
  * ``Plan.run``                  -> the per-window loop of examples/process/main.cpp:189-302
  * ``Context.cross_correlate``   -> algos::FFT::cross_correlate / cross_correlate_real (algos/fft.h:188-227)
  * ``Context.auto_correlate``    -> algos::FFT::auto_correlate (algos/fft.h:229-242)
"""
from __future__ import annotations

import ctypes
import os
import weakref
from pathlib import Path

import numpy as np

PKG_ROOT = Path(__file__).resolve().parent.parent
LIB_DIR = PKG_ROOT / "lib"
LIB_PATH = LIB_DIR / "libopenpiv_amd.so"
SYNTH_PATH = LIB_DIR / "libopiv_synth.so"

# opiv_dtype
U8, U16, F32, F64 = 0, 1, 2, 3
DTYPE_CODES = {np.dtype(np.uint8): U8, np.dtype(np.uint16): U16,
               np.dtype(np.float32): F32, np.dtype(np.float64): F64}
# opiv_variant
CIRCULAR_FFT_FP32 = 0
REAL_COMPAT = 1
DIRECT_BF16_SPLIT = 2
AUTO = 3
# opiv_flag
FORWARD = 0
REVERSE = 1
FLAG_FEWER_THAN_2_PEAKS = 1
FLAG_NONFINITE_FIT = 2
FLAG_NEAR_TIE = 4
FLAG_EXACT_TIE = 8

RESULT_DTYPE = np.dtype([
    ("x", "<f8"), ("y", "<f8"), ("u", "<f8"), ("v", "<f8"), ("sn", "<f8"),
    ("peak0", "<f8"), ("peak1", "<f8"),
    ("px", "<i4"), ("py", "<i4"), ("flags", "<u4"), ("reserved", "<u4"),
])
assert RESULT_DTYPE.itemsize == 72

# every symbol include/openpiv_amd.h declares
EXPORTS = (
    "opiv_last_error", "opiv_version",
    "opiv_ctx_create", "opiv_ctx_destroy", "opiv_ctx_synchronize", "opiv_ctx_stream",
    "opiv_grid_offsets", "opiv_grid",
    "opiv_plan_create", "opiv_plan_destroy", "opiv_plan_window_count", "opiv_plan_grid",
    "opiv_run_pair", "opiv_run_pair_async",
    "opiv_correlate_windows", "opiv_auto_correlate_windows", "opiv_fft_transform_windows",
    "opiv_batch_create", "opiv_batch_destroy", "opiv_batch_window_count", "opiv_batch_offsets",
    "opiv_batch_set_pairs", "opiv_batch_shape", "opiv_batch_form", "opiv_batch_run",
    "opiv_vjp_pair", "opiv_vjp_pair_async", "opiv_vjp_pair_async_noalloc", "opiv_variance_pair",
    "opiv_average_subtract",
    "opiv_image_probe", "opiv_image_read",
)


class OpivError(RuntimeError):
    """A non-zero opiv_status; ``code`` holds it, the message is opiv_last_error()."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class PlanDesc(ctypes.Structure):
    _fields_ = [
        ("img_w", ctypes.c_uint32), ("img_h", ctypes.c_uint32),
        ("pitch", ctypes.c_size_t), ("dtype", ctypes.c_int32),
        ("win", ctypes.c_uint32),
        ("step_x", ctypes.c_uint32), ("step_y", ctypes.c_uint32),
        ("overlap", ctypes.c_double),
        ("variant", ctypes.c_int32), ("limit_search", ctypes.c_int32),
    ]


class PairDesc(ctypes.Structure):
    _fields_ = [("a", ctypes.c_void_p), ("b", ctypes.c_void_p),
                ("img_w", ctypes.c_uint32), ("img_h", ctypes.c_uint32), ("pitch", ctypes.c_size_t)]


class ImageInfo(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("dtype", ctypes.c_int32),
                ("count", ctypes.c_uint32)]


class BatchDesc(ctypes.Structure):
    _fields_ = [("dtype", ctypes.c_int32), ("npasses", ctypes.c_uint32),
                ("win", ctypes.c_uint32 * 4), ("overlap", ctypes.c_double * 4),
                ("variant", ctypes.c_int32), ("limit_search", ctypes.c_int32)]


class SynthDesc(ctypes.Structure):
    _fields_ = [
        ("w", ctypes.c_uint32), ("h", ctypes.c_uint32), ("seed", ctypes.c_uint64),
        ("density", ctypes.c_double), ("diameter", ctypes.c_double),
        ("i0_lo", ctypes.c_double), ("i0_hi", ctypes.c_double),
        ("noise_sigma", ctypes.c_double), ("field", ctypes.c_int),
        ("p0", ctypes.c_double), ("p1", ctypes.c_double),
        ("p2", ctypes.c_double), ("p3", ctypes.c_double),
        ("dtype", ctypes.c_int), ("nthreads", ctypes.c_int),
    ]


_lib = None
_synth = None


def library_path() -> Path:
    return LIB_PATH


def load() -> ctypes.CDLL:
    """Load the HIP engine library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise OSError(f"{LIB_PATH} is missing: build it with `make -C {PKG_ROOT}` "
                      "(or __graft_entry__.build())")
    lib = ctypes.CDLL(str(LIB_PATH))
    vp, sz, u32, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_int32
    lib.opiv_last_error.restype = ctypes.c_char_p
    lib.opiv_version.restype = ctypes.c_char_p
    lib.opiv_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    lib.opiv_ctx_destroy.argtypes = [vp]
    lib.opiv_ctx_synchronize.argtypes = [vp]
    lib.opiv_ctx_stream.argtypes = [vp]
    lib.opiv_ctx_stream.restype = vp
    lib.opiv_grid_offsets.argtypes = [u32, u32, ctypes.c_double, ctypes.POINTER(u32), ctypes.POINTER(u32)]
    lib.opiv_grid.argtypes = [u32, u32, u32, u32, u32, u32, ctypes.POINTER(u32), ctypes.POINTER(u32), vp]
    lib.opiv_plan_create.argtypes = [vp, ctypes.POINTER(PlanDesc), ctypes.POINTER(vp)]
    lib.opiv_plan_destroy.argtypes = [vp]
    lib.opiv_plan_window_count.argtypes = [vp, ctypes.POINTER(sz)]
    lib.opiv_plan_grid.argtypes = [vp, vp]
    lib.opiv_run_pair.argtypes = [vp, vp, vp, vp, sz]
    lib.opiv_run_pair_async.argtypes = [vp, vp, vp, vp, sz]
    lib.opiv_correlate_windows.argtypes = [vp, u32, sz, i32, vp, vp, vp]
    lib.opiv_auto_correlate_windows.argtypes = [vp, u32, sz, vp, vp]
    lib.opiv_fft_transform_windows.argtypes = [vp, u32, sz, i32, vp, vp]
    lib.opiv_batch_create.argtypes = [vp, ctypes.POINTER(BatchDesc), ctypes.POINTER(PairDesc), sz,
                                      ctypes.POINTER(vp)]
    lib.opiv_batch_destroy.argtypes = [vp]
    lib.opiv_batch_window_count.argtypes = [vp, ctypes.POINTER(sz)]
    lib.opiv_batch_offsets.argtypes = [vp, vp]
    lib.opiv_batch_set_pairs.argtypes = [vp, ctypes.POINTER(PairDesc), sz]
    lib.opiv_batch_shape.argtypes = [vp, u32, ctypes.POINTER(i32), ctypes.POINTER(i32)]
    lib.opiv_batch_form.argtypes = [vp, ctypes.POINTER(i32), ctypes.POINTER(i32)]
    lib.opiv_batch_run.argtypes = [vp, vp, sz]
    lib.opiv_vjp_pair.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    lib.opiv_vjp_pair_async.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    lib.opiv_vjp_pair_async_noalloc.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    lib.opiv_variance_pair.argtypes = [vp, vp, vp, vp, ctypes.c_double, vp]
    lib.opiv_average_subtract.argtypes = [vp, vp, sz, u32, u32, sz, i32, vp, vp]
    lib.opiv_image_probe.argtypes = [ctypes.c_char_p, ctypes.POINTER(ImageInfo)]
    lib.opiv_image_read.argtypes = [vp, ctypes.c_char_p, u32, vp, sz]
    for name in EXPORTS:
        if name not in ("opiv_last_error", "opiv_version", "opiv_ctx_stream"):
            getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != 0:
        msg = load().opiv_last_error().decode()
        raise OpivError(rc, msg)


def last_error() -> str:
    """opiv_last_error() of this thread."""
    return load().opiv_last_error().decode()


def image_info(path) -> tuple[int, int, np.dtype, int]:
    """(width, height, pixel dtype, images in the file) of a PNM/TIFF frame."""
    info = ImageInfo()
    _check(load().opiv_image_probe(str(path).encode(), ctypes.byref(info)))
    return info.width, info.height, np.dtype(np.uint16 if info.dtype == U16 else np.uint8), info.count


def read_image(path, index: int = 0) -> np.ndarray:
    """Image `index` of a PNM/TIFF file as uint8/uint16 (h, w), as stored."""
    w, h, dt, _ = image_info(path)
    out = np.empty((h, w), dt)
    _check(load().opiv_image_read(None, str(path).encode(), index, out.ctypes.data, 0))
    return out


def direct_chosen(win: int, dtype=np.float32) -> bool:
    """Whether OPIV_AUTO picks the direct MFMA path: for 8-bit frames at
    16 x 16, where the exact int8 MFMA kernel measures faster than the FFT
    kernel."""
    return win == 16 and np.dtype(dtype) == np.uint8


def version() -> str:
    return load().opiv_version().decode()


def grid_offsets(win_w: int, win_h: int, overlap: float) -> tuple[int, int]:
    """generate_cartesian_grid's step from an overlap fraction (grid.impl.h:10-22)."""
    ox, oy = ctypes.c_uint32(), ctypes.c_uint32()
    _check(load().opiv_grid_offsets(win_w, win_h, overlap, ctypes.byref(ox), ctypes.byref(oy)))
    return ox.value, oy.value


def grid(img_w: int, img_h: int, win_w: int, win_h: int, off_x: int, off_y: int) -> np.ndarray:
    """Bottom-left corners, shape (n, 2) int32, in the reference's y-outer order."""
    lib = load()
    nx, ny = ctypes.c_uint32(), ctypes.c_uint32()
    _check(lib.opiv_grid(img_w, img_h, win_w, win_h, off_x, off_y, ctypes.byref(nx), ctypes.byref(ny), None))
    xy = np.zeros((nx.value * ny.value, 2), dtype=np.int32)
    _check(lib.opiv_grid(img_w, img_h, win_w, win_h, off_x, off_y, ctypes.byref(nx), ctypes.byref(ny),
                         xy.ctypes.data))
    return xy


def _ptr(x) -> int:
    """Data pointer of a numpy array (host) or a torch tensor (host or device)."""
    if isinstance(x, np.ndarray):
        return x.ctypes.data
    if hasattr(x, "data_ptr"):
        return int(x.data_ptr())
    if isinstance(x, int):
        return x
    raise TypeError(f"unsupported buffer type {type(x)!r}")


class Context:
    """One device and one HIP stream (opiv_ctx)."""

    def __init__(self, device: int = 0):
        self._lib = load()
        h = ctypes.c_void_p()
        _check(self._lib.opiv_ctx_create(device, ctypes.byref(h)))
        self.handle = h
        self.device = device
        # plans and batches made on this context: closed before it, since
        # the C objects refer to the context
        self._children = weakref.WeakSet()

    def _adopt(self, child) -> None:
        self._children.add(child)

    def close(self) -> None:
        if self.handle:
            for child in list(self._children):
                child.close()
            self._lib.opiv_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        return self._lib.opiv_ctx_stream(self.handle) or 0

    def synchronize(self) -> None:
        _check(self._lib.opiv_ctx_synchronize(self.handle))

    def cross_correlate(self, a: np.ndarray, b: np.ndarray, variant: int = CIRCULAR_FFT_FP32) -> np.ndarray:
        """FFT::cross_correlate(_real) of stacked square windows (n, W, W) -> swapped planes."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        if a.ndim == 2:
            a, b = a[None], b[None]
        if a.shape != b.shape or a.shape[1] != a.shape[2]:
            raise OpivError(3, f"image size is different from expected: {a.shape}, {b.shape}")
        n, win = a.shape[0], a.shape[1]
        out = np.empty_like(a)
        _check(self._lib.opiv_correlate_windows(self.handle, win, n, variant, a.ctypes.data, b.ctypes.data,
                                                out.ctypes.data))
        return out

    def auto_correlate(self, a: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.ndim == 2:
            a = a[None]
        n, win = a.shape[0], a.shape[1]
        out = np.empty_like(a)
        _check(self._lib.opiv_auto_correlate_windows(self.handle, win, n, a.ctypes.data, out.ctypes.data))
        return out

    def average_subtract(self, frames, *, return_mean: bool = False):
        """examples/average_subtract/main.cpp:98-159 on a stack of n >= 2 host
        frames (n, h, w) of uint8/uint16/float32/float64: the uint16 images
        the reference writes (and the mean of frames 1..n-1)."""
        frames = np.ascontiguousarray(frames)
        n, h, w = frames.shape
        dt = DTYPE_CODES[frames.dtype]
        out = np.zeros((n, h, w), np.uint16)
        mean = np.zeros((h, w), np.float64)
        fp = (ctypes.c_void_p * n)(*[frames[i].ctypes.data for i in range(n)])
        op = (ctypes.c_void_p * n)(*[out[i].ctypes.data for i in range(n)])
        _check(self._lib.opiv_average_subtract(self.handle, fp, n, w, h, 0, dt, op, mean.ctypes.data))
        return (out, mean) if return_mean else out

    def average_subtract_device(self, frame_ptrs, out_ptrs, w: int, h: int, dtype, pitch: int = 0,
                                mean_ptr: int | None = None) -> None:
        """The same on device frames and outputs (lists of pointers)."""
        n = len(frame_ptrs)
        fp = (ctypes.c_void_p * n)(*frame_ptrs)
        op = (ctypes.c_void_p * n)(*out_ptrs)
        _check(self._lib.opiv_average_subtract(self.handle, fp, n, w, h, pitch, DTYPE_CODES[np.dtype(dtype)],
                                               op, mean_ptr))

    def read_image_device(self, path, dst_ptr: int, pitch: int = 0, index: int = 0) -> None:
        """Image `index` of a PNM/TIFF file into device memory (its own pixel type)."""
        _check(self._lib.opiv_image_read(self.handle, str(path).encode(), index, dst_ptr, pitch))

    def transform(self, z: np.ndarray, direction: int = FORWARD) -> np.ndarray:
        """FFT::transform (fft.h:86-119) of one or n stacked square windows
        (complex or real input); complex128 output."""
        z = np.ascontiguousarray(z, dtype=np.complex128)
        single = z.ndim == 2
        if single:
            z = z[None]
        n, win = z.shape[0], z.shape[1]
        out = np.empty_like(z)
        _check(self._lib.opiv_fft_transform_windows(self.handle, win, n, direction, z.ctypes.data, out.ctypes.data))
        return out[0] if single else out


class Plan:
    """A frame size + window + grid (opiv_plan); ``run`` processes one pair."""

    def __init__(self, ctx: Context, img_w: int, img_h: int, win: int, *, overlap: float = 0.5,
                 step: tuple[int, int] | None = None, dtype=np.float32, pitch: int = 0,
                 variant: int = CIRCULAR_FFT_FP32, limit_search: bool = False):
        self.ctx = ctx
        self._lib = ctx._lib
        self.dtype = np.dtype(dtype)
        d = PlanDesc()
        d.img_w, d.img_h, d.pitch = img_w, img_h, pitch
        d.dtype = DTYPE_CODES[self.dtype]
        d.win = win
        d.step_x, d.step_y = step if step else (0, 0)
        d.overlap = overlap
        d.variant = variant
        d.limit_search = 1 if limit_search else 0
        self.desc = d
        h = ctypes.c_void_p()
        _check(self._lib.opiv_plan_create(ctx.handle, ctypes.byref(d), ctypes.byref(h)))
        self.handle = h
        ctx._adopt(self)
        n = ctypes.c_size_t()
        _check(self._lib.opiv_plan_window_count(h, ctypes.byref(n)))
        self.count = n.value

    def close(self) -> None:
        if self.handle:
            self._lib.opiv_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def grid(self) -> np.ndarray:
        xy = np.zeros((self.count, 2), dtype=np.int32)
        _check(self._lib.opiv_plan_grid(self.handle, xy.ctypes.data))
        return xy

    def _host_frames(self, a, b):
        """Host frames as contiguous arrays of the plan's pixel type, each
        holding at least pitch * img_h pixels (the C ABI reads that many)."""
        a = np.ascontiguousarray(a, dtype=self.dtype)
        b = np.ascontiguousarray(b, dtype=self.dtype)
        need = (self.desc.pitch or self.desc.img_w) * self.desc.img_h
        if a.size != b.size or a.size < need:
            raise ValueError(f"frames of {a.size} and {b.size} pixels; the plan reads {need} "
                             f"(pitch {self.desc.pitch or self.desc.img_w} x height {self.desc.img_h})")
        return a, b

    def _host_results(self, out, what="out"):
        if not isinstance(out, np.ndarray) or out.dtype != RESULT_DTYPE or out.size != self.count \
                or not out.flags.c_contiguous:
            raise ValueError(f"{what} must be a contiguous array of {self.count} RESULT_DTYPE records")
        return out

    def run(self, a, b, out=None) -> np.ndarray:
        """Process one frame pair (host numpy arrays or device tensors); returns results."""
        if isinstance(a, np.ndarray) or isinstance(b, np.ndarray):
            a, b = self._host_frames(a, b)
        if out is None:
            out = np.zeros(self.count, dtype=RESULT_DTYPE)
        elif isinstance(out, np.ndarray):
            out = self._host_results(out)
        _check(self._lib.opiv_run_pair(self.handle, _ptr(a), _ptr(b), _ptr(out), self.count))
        return out

    def run_async(self, a_ptr: int, b_ptr: int, out_ptr: int) -> None:
        """Enqueue on the context stream; device pointers only."""
        _check(self._lib.opiv_run_pair_async(self.handle, a_ptr, b_ptr, out_ptr, self.count))

    def vjp(self, a, b, cot, fwd=None):
        """Frame gradients of sum_i cot[i] . (u_i, v_i): (grad_a, grad_b) float32 (h, pitch)."""
        a, b = self._host_frames(a, b)
        if fwd is not None:
            fwd = self._host_results(fwd, "fwd")
        cot = np.ascontiguousarray(cot, dtype=np.float64)
        if cot.size != 2 * self.count:
            raise ValueError(f"cot must hold 2 x {self.count} values, not {cot.size}")
        cot = cot.reshape(self.count, 2)
        pitch = self.desc.pitch or self.desc.img_w
        ga = np.zeros((self.desc.img_h, pitch), np.float32)
        gb = np.zeros((self.desc.img_h, pitch), np.float32)
        _check(self._lib.opiv_vjp_pair(self.handle, _ptr(a), _ptr(b), None if fwd is None else _ptr(fwd),
                                       cot.ctypes.data, ga.ctypes.data, gb.ctypes.data))
        return ga, gb

    def variance(self, a, b, sigma: float, fwd=None) -> np.ndarray:
        """(n, 2) variances of (u, v) under i.i.d. pixel noise sigma."""
        a, b = self._host_frames(a, b)
        if fwd is not None:
            fwd = self._host_results(fwd, "fwd")
        var = np.zeros((self.count, 2), np.float64)
        _check(self._lib.opiv_variance_pair(self.handle, _ptr(a), _ptr(b), None if fwd is None else _ptr(fwd),
                                            sigma, var.ctypes.data))
        return var


class Batch:
    """Many frame pairs x several independent window passes, one graph launch
    per run (opiv_batch).  ``pairs`` is a list of (a_ptr, b_ptr, width, height)
    device frames of one dtype; results of (pass p, pair i) start at
    ``offsets[p * npairs + i]``."""

    def __init__(self, ctx: Context, pairs, wins=(64, 32, 16), overlaps=None, *, dtype=np.float32,
                 variant: int = CIRCULAR_FFT_FP32, limit_search: bool = False):
        self.ctx = ctx
        self._lib = ctx._lib
        overlaps = overlaps or [0.5] * len(wins)
        d = BatchDesc()
        d.dtype = DTYPE_CODES[np.dtype(dtype)]
        d.npasses = len(wins)
        for i, (w, o) in enumerate(zip(wins, overlaps)):
            d.win[i] = w
            d.overlap[i] = o
        d.variant = variant
        d.limit_search = 1 if limit_search else 0
        self.npairs = len(pairs)
        arr = self._pairs(pairs)
        h = ctypes.c_void_p()
        _check(self._lib.opiv_batch_create(ctx.handle, ctypes.byref(d), arr, self.npairs, ctypes.byref(h)))
        self.handle = h
        ctx._adopt(self)
        n = ctypes.c_size_t()
        _check(self._lib.opiv_batch_window_count(h, ctypes.byref(n)))
        self.count = n.value
        off = np.zeros(len(wins) * self.npairs + 1, dtype=np.uint64)
        _check(self._lib.opiv_batch_offsets(h, off.ctypes.data))
        self.offsets = off.astype(np.int64)
        self.npasses = len(wins)

    @staticmethod
    def _pairs(pairs):
        arr = (PairDesc * len(pairs))()
        for i, (a, b, w, h) in enumerate(pairs):
            arr[i].a, arr[i].b, arr[i].img_w, arr[i].img_h, arr[i].pitch = _ptr(a), _ptr(b), w, h, 0
        return arr

    def set_pairs(self, pairs) -> None:
        _check(self._lib.opiv_batch_set_pairs(self.handle, self._pairs(pairs), len(pairs)))

    def shape(self, p: int) -> tuple[int, int]:
        t, b = ctypes.c_int32(), ctypes.c_int32()
        _check(self._lib.opiv_batch_shape(self.handle, p, ctypes.byref(t), ctypes.byref(b)))
        return t.value, b.value

    def form(self) -> tuple[bool, int]:
        """(True, blocks) when every pass runs in one persistent launch of
        ``blocks`` blocks, (False, 0) for one launch per pass."""
        single, blocks = ctypes.c_int32(), ctypes.c_int32()
        _check(self._lib.opiv_batch_form(self.handle, ctypes.byref(single), ctypes.byref(blocks)))
        return bool(single.value), blocks.value

    def run(self, out_ptr: int) -> None:
        """Enqueue all passes (one graph launch); out_ptr: device, ``count`` results."""
        _check(self._lib.opiv_batch_run(self.handle, out_ptr, self.count))

    def block(self, results: np.ndarray, p: int, i: int) -> np.ndarray:
        k = p * self.npairs + i
        return results[self.offsets[k]:self.offsets[k + 1]]

    def close(self) -> None:
        if self.handle:
            self._lib.opiv_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- device memory

_hip = None


def _hip_runtime():
    """The HIP runtime libopenpiv_amd.so is linked against (same soname, so
    the dynamic loader hands back the already loaded copy).  PyTorch ships a
    second copy of the runtime; loading both in one process leaves the later
    one without devices, so device buffers are taken from this one."""
    global _hip
    if _hip is None:
        load()
        h = ctypes.CDLL("libamdhip64.so.7")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        h.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
        h.hipFree.argtypes = [vp]
        h.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
        h.hipMemset.argtypes = [vp, ctypes.c_int, sz]
        h.hipDeviceSynchronize.argtypes = []
        for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipMemset, h.hipDeviceSynchronize):
            f.restype = ctypes.c_int
        _hip = h
    return _hip


class DeviceBuffer:
    """Device allocation (hipMalloc) with host copies; ``ptr`` is the device address.

    ``upload`` and ``zero`` run on the null stream and return only once the
    device has finished them: a context's stream is non-blocking, so work it
    is given afterwards is not otherwise ordered after them."""

    H2D, D2H = 1, 2

    def __init__(self, nbytes: int):
        self._h = _hip_runtime()
        p = ctypes.c_void_p()
        rc = self._h.hipMalloc(ctypes.byref(p), max(1, nbytes))
        if rc != 0:
            raise OpivError(5, f"hipMalloc({nbytes}) failed: {rc}")
        self.ptr = p.value
        self.nbytes = nbytes

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        buf = cls(a.nbytes)
        buf.upload(a)
        return buf

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError(f"{a.nbytes} bytes do not fit the {self.nbytes}-byte device buffer")
        if self._h.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, self.H2D) != 0:
            raise OpivError(5, "hipMemcpy H2D failed")
        if self._h.hipDeviceSynchronize() != 0:
            raise OpivError(5, "hipDeviceSynchronize failed")

    def download(self, out: np.ndarray) -> np.ndarray:
        if not out.flags.c_contiguous or out.nbytes > self.nbytes:
            raise ValueError(f"out must be a contiguous array of at most {self.nbytes} bytes")
        if self._h.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, self.D2H) != 0:
            raise OpivError(5, "hipMemcpy D2H failed")
        return out

    def zero(self) -> None:
        if self._h.hipMemset(self.ptr, 0, self.nbytes) != 0:
            raise OpivError(5, "hipMemset failed")
        if self._h.hipDeviceSynchronize() != 0:
            raise OpivError(5, "hipDeviceSynchronize failed")

    def free(self) -> None:
        if self.ptr:
            self._h.hipFree(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------- synthetic frames

FIELD_UNIFORM, FIELD_ROTATION, FIELD_LAMB_OSEEN = 0, 1, 2
_SYNTH_CODES = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1,
                np.dtype(np.float32): 2, np.dtype(np.float64): 3}


def _load_synth():
    global _synth
    if _synth is None:
        if not SYNTH_PATH.exists():
            raise OSError(f"{SYNTH_PATH} is missing: build it with `make -C {PKG_ROOT}`")
        _synth = ctypes.CDLL(str(SYNTH_PATH))
        _synth.opiv_synth_pair.argtypes = [ctypes.POINTER(SynthDesc), ctypes.c_void_p, ctypes.c_void_p]
        _synth.opiv_synth_pair.restype = ctypes.c_int
    return _synth


def synth_pair(w: int, h: int, *, seed: int = 1, density: float = 0.05, diameter: float = 2.5,
               i0=(100.0, 255.0), noise: float = 2.0, field: int = FIELD_UNIFORM,
               params=(0.0, 0.0, 0.0, 0.0), dtype=np.float64, threads: int = 0):
    """Seeded particle-image pair (frame A, frame B) as numpy arrays of shape (h, w)."""
    dt = np.dtype(dtype)
    d = SynthDesc()
    d.w, d.h, d.seed = w, h, seed
    d.density, d.diameter = density, diameter
    d.i0_lo, d.i0_hi = i0
    d.noise_sigma = noise
    d.field = field
    p = list(params) + [0.0] * (4 - len(params))
    d.p0, d.p1, d.p2, d.p3 = p[:4]
    d.dtype = _SYNTH_CODES[dt]
    d.nthreads = threads or min(16, os.cpu_count() or 1)
    a = np.empty((h, w), dtype=dt)
    b = np.empty((h, w), dtype=dt)
    if _load_synth().opiv_synth_pair(ctypes.byref(d), a.ctypes.data, b.ctypes.data) != 0:
        raise ValueError("opiv_synth_pair rejected the descriptor")
    return a, b


# ---------------------------------------------------------------- workloads

MULTIPASS_SIZES = [(512, 512), (1024, 1024), (2048, 2048), (4096, 4096), (1280, 800)]


def multipass_pairs(npairs: int = 24, seed: int = 24, dtype=np.float32):
    """The uneven multi-pair workload: npairs frame pairs whose sizes are
    drawn (seeded) from MULTIPASS_SIZES, each with a Lamb-Oseen vortex of core
    radius 40 px and peak 5 px at a seeded centre.  Yields (a, b) numpy
    frames one pair at a time."""
    rng = np.random.default_rng(seed)
    picks = rng.integers(0, len(MULTIPASS_SIZES), npairs)
    for i, k in enumerate(picks):
        w, h = MULTIPASS_SIZES[k]
        cx, cy = rng.uniform(0.25, 0.75) * w, rng.uniform(0.25, 0.75) * h
        yield synth_pair(w, h, seed=1000 + i, field=FIELD_LAMB_OSEEN, params=(cx, cy, 40.0, 5.0), dtype=dtype)
