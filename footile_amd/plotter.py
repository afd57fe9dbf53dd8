"""Plotter, Raster, JoinStyle and the batched fill (src/plotter.rs, src/stroker.rs).

``Plotter`` mirrors footile::Plotter<P> (src/plotter.rs:38-381): it owns its
raster -- held in GPU memory by libfootile_hip.so -- and keeps the tolerance,
transform, join style and stroke width across calls.  ``fill`` and ``stroke``
return the plotter; ``raster()`` copies the pixels back.
"""
import ctypes

import numpy as np

from ._lib import FillDescC, check, lib
from .path import ops_array
from .pixel import BPP, MATTE8

_vp = ctypes.c_void_p


def _ptr(a):
    return a.ctypes.data_as(_vp)


class JoinStyle:
    """src/stroker.rs:13-20; values are (kind, miter_limit)."""
    Bevel = (1, 0.0)
    Round = (2, 0.0)

    @staticmethod
    def Miter(limit=4.0):
        return (0, float(limit))


class Transform:
    """Affine transform elements e (x' = e0 x + e1 y + e2, y' = e3 x + e4 y + e5)."""

    def __init__(self, e=(1.0, 0.0, 0.0, 0.0, 1.0, 0.0)):
        self.e = tuple(float(v) for v in e)

    @classmethod
    def with_scale(cls, sx, sy):
        return cls((sx, 0.0, 0.0, 0.0, sy, 0.0))

    @classmethod
    def with_translate(cls, tx, ty):
        return cls((1.0, 0.0, tx, 0.0, 1.0, ty))


class Raster:
    """Host copy of a pix Raster<P>: ``pixels`` is uint8 (height, width, bpp)."""

    def __init__(self, width, height, fmt=MATTE8, pixels=None):
        self.width, self.height, self.fmt = int(width), int(height), fmt
        shape = (self.height, self.width, BPP[fmt])
        if pixels is None:
            self.pixels = np.zeros(shape, dtype=np.uint8)
        else:
            self.pixels = np.ascontiguousarray(pixels, dtype=np.uint8).reshape(shape)

    @classmethod
    def with_clear(cls, width, height, fmt=MATTE8):
        return cls(width, height, fmt)

    def as_u8_slice(self):
        return self.pixels.reshape(-1)


def _clr_bytes(clr):
    b = (ctypes.c_uint8 * 4)()
    vals = list(clr) if clr is not None else [255]
    for i, v in enumerate(vals[:4]):
        b[i] = int(v) & 0xFF
    return b


class Plotter:
    """footile::Plotter<P> on the GPU."""

    def __init__(self, raster):
        self._fmt = raster.fmt
        h = _vp()
        check(lib().ft_plotter_new_with_pixels(raster.width, raster.height, raster.fmt,
                                               _ptr(raster.pixels), ctypes.byref(h)),
              "ft_plotter_new_with_pixels")
        self._h = h
        self._w, self._hgt = raster.width, raster.height

    def __del__(self):
        self.close()

    def close(self):
        if getattr(self, "_h", None):
            lib().ft_plotter_free(self._h)
            self._h = None

    def width(self):
        return self._w

    def height(self):
        return self._hgt

    def set_tolerance(self, t):
        check(lib().ft_plotter_set_tolerance(self._h, float(t)), "set_tolerance")
        return self

    def set_transform(self, t):
        e = t.e if isinstance(t, Transform) else tuple(t)
        arr = np.asarray(e, dtype=np.float32)
        check(lib().ft_plotter_set_transform(self._h, _ptr(arr)), "set_transform")
        return self

    def set_join(self, js):
        kind, limit = js
        check(lib().ft_plotter_set_join(self._h, kind, float(limit)), "set_join")
        return self

    def fill(self, rule, ops, clr=(255,)):
        a = ops_array(ops)
        check(lib().ft_plotter_fill(self._h, int(rule), _ptr(a), len(a), _clr_bytes(clr)), "fill")
        return self

    def stroke(self, ops, clr=(255,)):
        a = ops_array(ops)
        check(lib().ft_plotter_stroke(self._h, _ptr(a), len(a), _clr_bytes(clr)), "stroke")
        return self

    def raster(self):
        r = Raster(self._w, self._hgt, self._fmt)
        check(lib().ft_plotter_read(self._h, _ptr(r.pixels), r.pixels.nbytes), "read")
        return r

    def write_raster(self, raster):
        """raster_mut(): replace the plotter's pixels."""
        px = np.ascontiguousarray(raster.pixels, dtype=np.uint8)
        check(lib().ft_plotter_write(self._h, _ptr(px), px.nbytes), "write")
        return self

    def into_raster(self):
        r = self.raster()
        self.close()
        return r

    @property
    def pen_width(self):
        w = ctypes.c_float()
        check(lib().ft_plotter_pen_width(self._h, ctypes.byref(w)), "pen_width")
        return w.value

    def last_counts(self):
        """(points offered to Fig::add_point, points stored) of the last fill."""
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().ft_plotter_last_counts(self._h, ctypes.byref(a), ctypes.byref(b)), "last_counts")
        return a.value, b.value

    def device_ptr(self):
        p, pitch = _vp(), ctypes.c_size_t()
        check(lib().ft_plotter_device_ptr(self._h, ctypes.byref(p), ctypes.byref(pitch)), "device_ptr")
        return p.value, pitch.value


def fig_fill(rule, points, closes, width, height, fmt=MATTE8, clr=(255, 255, 255, 255), pixels=None):
    """Fig-level entry (src/fig.rs:339-502): points as (x, y) f32 pairs, closes[i]
    true to call Fig::close after point i.  Returns the raster bytes."""
    xy = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 2)
    cl = np.ascontiguousarray(closes, dtype=np.uint8).reshape(-1)
    px = (np.zeros(width * height * BPP[fmt], dtype=np.uint8) if pixels is None
          else np.ascontiguousarray(pixels, dtype=np.uint8).reshape(-1).copy())
    check(lib().ft_fig_fill(int(rule), _ptr(xy), _ptr(cl), len(xy), fmt, width, height,
                            _ptr(px), _clr_bytes(clr)), "ft_fig_fill")
    return px


class DeviceBuffer:
    """A device allocation made by the library (no other runtime in-process)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = _vp()
        check(lib().ft_device_alloc(self.nbytes, ctypes.byref(p)), "ft_device_alloc")
        self.ptr = p.value

    def __del__(self):
        self.free()

    def free(self):
        if getattr(self, "ptr", None):
            lib().ft_device_free(_vp(self.ptr))
            self.ptr = None

    def upload(self, host):
        a = np.ascontiguousarray(host)
        check(lib().ft_memcpy_h2d(_vp(self.ptr), _ptr(a), a.nbytes), "h2d")

    def download(self, nbytes=None, dtype=np.uint8):
        n = self.nbytes if nbytes is None else nbytes
        a = np.empty(n, dtype=np.uint8)
        check(lib().ft_memcpy_d2h(_ptr(a), _vp(self.ptr), n), "d2h")
        return a.view(dtype)

    def zero(self):
        check(lib().ft_memset_d(_vp(self.ptr), 0, self.nbytes), "memset")


class FillDesc:
    """One fill of a batch (ft_fill_desc)."""

    def __init__(self, op_begin, op_count, dst, width, height, pitch, fmt=MATTE8, rule=0,
                 clr=(255, 255, 255, 255), transform=(1, 0, 0, 0, 1, 0), tolerance=0.3,
                 pen_width=1.0, fresh=False):
        self.c = FillDescC()
        self.c.op_begin, self.c.op_count, self.c.rule = op_begin, op_count, rule
        self.c.dst, self.c.width, self.c.height, self.c.pitch = dst, width, height, pitch
        self.c.fmt = fmt
        for i, v in enumerate(list(clr)[:4]):
            self.c.clr[i] = int(v) & 0xFF
        for i, v in enumerate(transform):
            self.c.transform[i] = float(v)
        self.c.tolerance, self.c.pen_width = float(tolerance), float(pen_width)
        self.c.fresh = 1 if fresh else 0


class _Tuning(ctypes.Structure):
    _fields_ = [("band_rows_min", ctypes.c_uint32), ("lds_cells", ctypes.c_uint32),
                ("dst_prefetch", ctypes.c_uint32), ("tile_groups", ctypes.c_uint32)]


class Batch:
    """Many independent fills from device-resident ops (ft_batch_*)."""

    def __init__(self, ops, descs):
        a = ops_array(ops)
        self._ops = a
        arr = (FillDescC * max(len(descs), 1))(*[d.c for d in descs])
        self.n_paths = len(descs)
        h = _vp()
        check(lib().ft_batch_new(_ptr(a), len(a), arr, len(descs), ctypes.byref(h)), "ft_batch_new")
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            lib().ft_batch_free(self._h)
            self._h = None

    def run(self):
        check(lib().ft_batch_run(self._h), "ft_batch_run")

    def launch(self):
        check(lib().ft_batch_launch(self._h), "ft_batch_launch")

    def sync(self):
        check(lib().ft_batch_sync(self._h), "ft_batch_sync")

    def set_timing(self, on=True):
        check(lib().ft_batch_set_timing(self._h, 1 if on else 0), "ft_batch_set_timing")

    def timing(self):
        """(total ms, raster-kernel ms, runs) summed since the last call."""
        t, r, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_uint64()
        check(lib().ft_batch_timing(self._h, ctypes.byref(t), ctypes.byref(r), ctypes.byref(n)),
              "ft_batch_timing")
        return t.value, r.value, n.value

    def set_tuning(self, band_rows_min, lds_cells, dst_prefetch=True, tile_groups=0):
        """Raster geometry (ft_batch_set_tuning); re-lays the batch out.
        tile_groups > 0 sends paths of at most that many LDS row groups to the
        tile scheduler."""
        t = _Tuning(int(band_rows_min), int(lds_cells), 1 if dst_prefetch else 0, int(tile_groups))
        check(lib().ft_batch_set_tuning(self._h, ctypes.byref(t)), "ft_batch_set_tuning")

    def tuning(self):
        """(band_rows_min, lds_cells, dst_prefetch, tile_groups)"""
        t = _Tuning()
        check(lib().ft_batch_get_tuning(self._h, ctypes.byref(t)), "ft_batch_get_tuning")
        return t.band_rows_min, t.lds_cells, bool(t.dst_prefetch), t.tile_groups

    def raster_stats(self):
        """(edge, row) pairs walked and nonzero cell deltas added by one run
        (ft_batch_raster_stats; runs the batch once)."""
        a, b = ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().ft_batch_raster_stats(self._h, ctypes.byref(a), ctypes.byref(b)), "ft_batch_raster_stats")
        return a.value, b.value

    def set_graph(self, on=True):
        """Replay launches from a captured HIP graph (ft_batch_set_graph)."""
        check(lib().ft_batch_set_graph(self._h, 1 if on else 0), "ft_batch_set_graph")

    def set_pipeline(self, depth=2):
        """Take `depth` workspaces in turn (True: 2, False: 1) so that later
        launches' front ends overlap an earlier launch's raster
        (ft_batch_set_pipeline); rasters keep launch order."""
        depth = 2 if depth is True else (1 if depth is False else int(depth))
        check(lib().ft_batch_set_pipeline(self._h, depth), "ft_batch_set_pipeline")

    def counts(self):
        """(offered, stored, edges) per path of the last run."""
        return self.results()[:3]

    def results(self):
        """(offered, stored, edges, top_row) per path of the last run."""
        n = max(self.n_paths, 1)
        off, st, ed = (np.zeros(n, dtype=np.uint64) for _ in range(3))
        top = np.zeros(n, dtype=np.int32)
        check(lib().ft_batch_counts(self._h, _ptr(off), _ptr(st), _ptr(ed), _ptr(top)), "ft_batch_counts")
        k = self.n_paths
        return off[:k], st[:k], ed[:k], top[:k]
