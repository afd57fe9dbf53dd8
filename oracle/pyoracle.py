"""ctypes wrapper over oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

The CPU oracle restates footile 0.7.0's sequential fill/stroke algorithm in
plain C (oracle/footile_oracle.c).  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg import this module, as the checker; the product
package (footile_amd) never does.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

# Same 28-byte layout as ft_path_op / or_path_op: u32 tag + 6 x f32.
OP_DTYPE = np.dtype([("tag", "<u4"), ("v", "<f4", (6,))], align=False)
CLOSE, MOVE, LINE, QUAD, CUBIC, PEN_WIDTH = range(6)
MATTE8, RGBA8P, GRAYA8P = 0, 1, 2
NON_ZERO, EVEN_ODD = 0, 1
MITER, BEVEL, ROUND = 0, 1, 2
BPP = {MATTE8: 1, RGBA8P: 4, GRAYA8P: 2}

_lib = None


def build():
    """Compile liboracle.so with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        i32, u32, f32 = ctypes.c_int32, ctypes.c_uint32, ctypes.c_float
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        for name in ["or_fx_from_i32", "or_fx_to_i32", "or_fx_abs", "or_fx_floor",
                     "or_fx_ceil", "or_fx_round", "or_fx_trunc", "or_fx_fract"]:
            getattr(L, name).argtypes = [i32]
            getattr(L, name).restype = i32
        for name in ["or_fx_add", "or_fx_sub", "or_fx_mul", "or_fx_div", "or_fx_avg"]:
            getattr(L, name).argtypes = [i32, i32]
            getattr(L, name).restype = i32
        for name in ["or_fx_shl", "or_fx_shr"]:
            getattr(L, name).argtypes = [i32, u32]
            getattr(L, name).restype = i32
        L.or_fx_from_f32.argtypes = [f32]
        L.or_fx_from_f32.restype = i32
        L.or_fx_to_f32.argtypes = [i32]
        L.or_fx_to_f32.restype = f32
        L.or_fx_widdershins.argtypes = [i32, i32, i32, i32]
        L.or_fx_widdershins.restype = ctypes.c_int
        L.or_pixel_cov.argtypes = [i32]
        L.or_pixel_cov.restype = ctypes.c_int16
        L.or_ch8_mul.argtypes = [ctypes.c_uint8, ctypes.c_uint8]
        L.or_ch8_mul.restype = ctypes.c_uint8
        for name in ["or_accumulate_non_zero_scalar", "or_accumulate_non_zero_ssse3",
                     "or_accumulate_even_odd_scalar", "or_accumulate_even_odd_ssse3"]:
            getattr(L, name).argtypes = [vp, vp, sz]
            getattr(L, name).restype = None
        L.or_fig_new.restype = vp
        L.or_fig_free.argtypes = [vp]
        L.or_fig_add_point.argtypes = [vp, f32, f32]
        L.or_fig_close.argtypes = [vp]
        L.or_fig_points.argtypes = [vp, vp, sz]
        L.or_fig_points.restype = sz
        L.or_fig_fill.argtypes = [vp, ctypes.c_int, vp, u32, u32, ctypes.c_int, vp, vp, sz,
                                  ctypes.c_int]
        L.or_fig_fill.restype = ctypes.c_int
        L.or_plotter_new.argtypes = [u32, u32, ctypes.c_int]
        L.or_plotter_new.restype = vp
        L.or_plotter_free.argtypes = [vp]
        L.or_plotter_set_tolerance.argtypes = [vp, f32]
        L.or_plotter_set_transform.argtypes = [vp, vp]
        L.or_plotter_set_join.argtypes = [vp, ctypes.c_int, f32]
        L.or_plotter_set_ssse3.argtypes = [vp, ctypes.c_int]
        L.or_plotter_s_width.argtypes = [vp]
        L.or_plotter_s_width.restype = f32
        L.or_plotter_pixels.argtypes = [vp]
        L.or_plotter_pixels.restype = vp
        L.or_plotter_clear.argtypes = [vp]
        L.or_plotter_fill.argtypes = [vp, ctypes.c_int, vp, sz, vp]
        L.or_plotter_fill.restype = ctypes.c_int
        L.or_plotter_stroke.argtypes = [vp, vp, sz, vp]
        L.or_plotter_stroke.restype = ctypes.c_int
        L.or_plotter_flatten.argtypes = [vp, vp, sz, vp, sz, ctypes.POINTER(sz)]
        L.or_plotter_flatten.restype = ctypes.c_int
        L.or_plotter_stroke_ops.argtypes = [vp, vp, sz, vp, sz, ctypes.POINTER(sz)]
        L.or_plotter_stroke_ops.restype = ctypes.c_int
        L.or_plotter_last_counts.argtypes = [vp, ctypes.POINTER(sz), ctypes.POINTER(sz)]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _ops(ops):
    ops = np.ascontiguousarray(ops, dtype=OP_DTYPE)
    return ops, _ptr(ops), len(ops)


def _clr(clr):
    c = np.zeros(4, dtype=np.uint8)
    c[: len(clr)] = clr
    return c


class Fig:
    """Fig::{new, add_point, close, fill} (src/fig.rs:339-502)."""

    def __init__(self):
        self._f = lib().or_fig_new()

    def __del__(self):
        if getattr(self, "_f", None):
            lib().or_fig_free(self._f)
            self._f = None

    def add_point(self, x, y):
        lib().or_fig_add_point(self._f, float(x), float(y))

    def close(self):
        lib().or_fig_close(self._f)

    def points(self):
        n = lib().or_fig_points(self._f, None, 0)
        xy = np.zeros((max(n, 1), 2), dtype=np.int32)
        lib().or_fig_points(self._f, _ptr(xy), n)
        return xy[:n]

    def fill(self, rule, width, height, fmt=MATTE8, clr=(255, 255, 255, 255), sgn_len=None,
             pixels=None, ssse3=True):
        bpp = BPP[fmt]
        px = (np.zeros(width * height * bpp, dtype=np.uint8) if pixels is None
              else np.ascontiguousarray(pixels, dtype=np.uint8).reshape(-1).copy())
        n = sgn_len if sgn_len is not None else width
        sgn = np.zeros(n + 8, dtype=np.int16)
        c = _clr(clr)
        rc = lib().or_fig_fill(self._f, rule, _ptr(px), width, height, fmt, _ptr(c), _ptr(sgn), n,
                               1 if ssse3 else 0)
        if rc:
            raise RuntimeError(f"or_fig_fill -> {rc}")
        # Cells at or past `width` may stay non-zero (src/imgbuf.rs:54-61
        # zeroes only the first width cells); they never reach a pixel.
        assert not sgn[:width].any(), "signed-area row must be zero after a fill"
        return px


class Plotter:
    """Plotter<P> (src/plotter.rs:38-381) over the oracle."""

    def __init__(self, width, height, fmt=MATTE8):
        self.width, self.height, self.fmt = width, height, fmt
        self._p = lib().or_plotter_new(width, height, fmt)

    def __del__(self):
        if getattr(self, "_p", None):
            lib().or_plotter_free(self._p)
            self._p = None

    def set_tolerance(self, t):
        lib().or_plotter_set_tolerance(self._p, float(t))
        return self

    def set_transform(self, e):
        a = np.asarray(e, dtype=np.float32).reshape(6)
        lib().or_plotter_set_transform(self._p, _ptr(a))
        return self

    def set_join(self, kind, miter=4.0):
        lib().or_plotter_set_join(self._p, kind, float(miter))
        return self

    def set_ssse3(self, on):
        lib().or_plotter_set_ssse3(self._p, 1 if on else 0)
        return self

    @property
    def s_width(self):
        return lib().or_plotter_s_width(self._p)

    def pixels(self):
        n = self.width * self.height * BPP[self.fmt]
        buf = (ctypes.c_uint8 * n).from_address(lib().or_plotter_pixels(self._p))
        return np.frombuffer(buf, dtype=np.uint8).copy()

    def clear(self):
        lib().or_plotter_clear(self._p)

    def set_pixels(self, px):
        a = np.ascontiguousarray(px, dtype=np.uint8).reshape(-1)
        n = self.width * self.height * BPP[self.fmt]
        assert a.size == n
        ctypes.memmove(lib().or_plotter_pixels(self._p), a.ctypes.data, n)

    def fill(self, rule, ops, clr=(255, 255, 255, 255)):
        ops, p, n = _ops(ops)
        c = _clr(clr)
        rc = lib().or_plotter_fill(self._p, rule, p, n, _ptr(c))
        return rc

    def stroke(self, ops, clr=(255, 255, 255, 255)):
        ops, p, n = _ops(ops)
        c = _clr(clr)
        return lib().or_plotter_stroke(self._p, p, n, _ptr(c))

    def last_counts(self):
        a, b = ctypes.c_size_t(), ctypes.c_size_t()
        lib().or_plotter_last_counts(self._p, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def flatten(self, ops):
        """Points (x, y, w) handed to Fig/Stroke::add_point; close markers are
        rows with NaN x/y and w = 1.0 (joined) or 0.0."""
        ops, p, n = _ops(ops)
        cnt = ctypes.c_size_t()
        rc = lib().or_plotter_flatten(self._p, p, n, None, 0, ctypes.byref(cnt))
        if rc:
            return rc, None
        out = np.zeros((max(cnt.value, 1), 3), dtype=np.float32)
        rc = lib().or_plotter_flatten(self._p, p, n, _ptr(out), cnt.value, ctypes.byref(cnt))
        return rc, out[: cnt.value]

    def stroke_ops(self, ops):
        ops, p, n = _ops(ops)
        cnt = ctypes.c_size_t()
        rc = lib().or_plotter_stroke_ops(self._p, p, n, None, 0, ctypes.byref(cnt))
        if rc:
            return rc, None
        out = np.zeros(max(cnt.value, 1), dtype=OP_DTYPE)
        rc = lib().or_plotter_stroke_ops(self._p, p, n, _ptr(out), cnt.value, ctypes.byref(cnt))
        return rc, out[: cnt.value]
