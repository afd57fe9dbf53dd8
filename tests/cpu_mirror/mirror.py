"""ctypes wrapper of the host build of the pipeline stages (test infrastructure).

``run(...)`` mirrors a batch fill on host rasters.  Import with
``MIRROR_ASAN=1`` (and libasan preloaded) to load the AddressSanitizer build.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        build()
        name = "libft_mirror_asan.so" if os.environ.get("MIRROR_ASAN") else "libft_mirror.so"
        L = ctypes.CDLL(os.path.join(HERE, name))
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        L.ftm_run.argtypes = [vp, sz, vp, sz, ctypes.c_uint32, ctypes.c_uint32, vp, vp, vp, sz,
                              ctypes.POINTER(sz), ctypes.c_uint32, ctypes.c_float]
        L.ftm_run.restype = ctypes.c_int
        L.ftm_row_cells_check.argtypes = [ctypes.c_uint64, ctypes.c_int]
        L.ftm_row_cells_check.restype = ctypes.c_int
        L.ftm_split_check.argtypes = [ctypes.c_uint64, ctypes.c_int]
        L.ftm_split_check.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


MODE_FILL, MODE_STROKE = 0, 2
PATH_RAW, PATH_WHOLE = 2, 4


def run_batch(ops, descs, raw=False, mode=MODE_FILL, join=(0, 4.0)):
    """descs: list of FillDescC whose dst point at host memory.  Returns
    (status, offered, stored)."""
    from footile_amd._lib import FillDescC
    from footile_amd.path import ops_array
    a = ops_array(ops)
    arr = (FillDescC * max(len(descs), 1))(*descs)
    off = np.zeros(max(len(descs), 1), np.uint64)
    st = np.zeros(max(len(descs), 1), np.uint64)
    flags = (PATH_RAW if raw else 0) | (PATH_WHOLE if mode == MODE_STROKE else 0)
    rc = lib().ftm_run(_ptr(a), len(a), arr, len(descs), mode, flags, _ptr(off), _ptr(st), None, 0, None,
                       int(join[0]), float(join[1]))
    return rc, off[: len(descs)], st[: len(descs)]


def stroke(width, height, fmt, ops, clr=(255, 255, 255, 255), init=None, transform=(1, 0, 0, 0, 1, 0),
           tolerance=0.3, pen_width=1.0, join=(0, 4.0)):
    """One Plotter::stroke on a host raster through the mirrored stages."""
    return fill(width, height, fmt, 0, ops, clr, init, transform, tolerance, pen_width, mode=MODE_STROKE, join=join)


def fill(width, height, fmt, rule, ops, clr=(255, 255, 255, 255), init=None, transform=(1, 0, 0, 0, 1, 0),
         tolerance=0.3, pen_width=1.0, raw=False, mode=MODE_FILL, join=(0, 4.0)):
    """One Plotter::fill on a host raster through the mirrored stages."""
    from footile_amd._lib import FillDescC
    from footile_amd.path import ops_array
    from footile_amd.pixel import BPP
    a = ops_array(ops)
    bpp = BPP[fmt]
    px = (np.zeros(width * height * bpp, np.uint8) if init is None
          else np.ascontiguousarray(init, np.uint8).reshape(-1).copy())
    d = FillDescC()
    d.op_begin, d.op_count, d.rule = 0, len(a), rule
    d.dst, d.width, d.height, d.pitch, d.fmt = px.ctypes.data, width, height, width * bpp, fmt
    for i, v in enumerate(list(clr)[:4]):
        d.clr[i] = int(v) & 0xFF
    for i, v in enumerate(transform):
        d.transform[i] = float(v)
    d.tolerance, d.pen_width, d.fresh = float(tolerance), float(pen_width), 0
    rc, off, st = run_batch(a, [d], raw=raw, mode=mode, join=join)
    return rc, px, (int(off[0]), int(st[0]))


def row_cells_check(seed, iters):
    """Rows whose closed-form cells differ from the step-by-step walk."""
    return lib().ftm_row_cells_check(seed, iters)


def split_check(seed, iters):
    """Curves whose points, flattened by 2^sp lanes (sp = 0..6), differ from
    the walk from the root."""
    return lib().ftm_split_check(seed, iters)
