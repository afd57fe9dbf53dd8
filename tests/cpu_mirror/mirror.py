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
                              ctypes.POINTER(sz)]
        L.ftm_run.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def run_batch(ops, descs, raw=False):
    """descs: list of FillDescC whose dst point at host memory.  Returns
    (status, offered, stored)."""
    from footile_amd._lib import FillDescC
    from footile_amd.path import ops_array
    a = ops_array(ops)
    arr = (FillDescC * max(len(descs), 1))(*descs)
    off = np.zeros(max(len(descs), 1), np.uint64)
    st = np.zeros(max(len(descs), 1), np.uint64)
    rc = lib().ftm_run(_ptr(a), len(a), arr, len(descs), 0, 2 if raw else 0, _ptr(off), _ptr(st),
                       None, 0, None)
    return rc, off[: len(descs)], st[: len(descs)]


def fill(width, height, fmt, rule, ops, clr=(255, 255, 255, 255), init=None, transform=(1, 0, 0, 0, 1, 0),
         tolerance=0.3, pen_width=1.0, raw=False):
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
    rc, off, st = run_batch(a, [d], raw=raw)
    return rc, px, (int(off[0]), int(st[0]))
