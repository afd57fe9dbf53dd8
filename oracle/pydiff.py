"""ctypes wrapper of the FP64 oracle of the differentiable fill (test
infrastructure only: tests/ and bench.py's checks use it, the product never does).
"""
import ctypes

import numpy as np

from oracle.pyoracle import lib as _oracle_lib

SEG = 16


class _Scene(ctypes.Structure):
    _fields_ = [("n_images", ctypes.c_uint32), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("n_paths", ctypes.c_uint32), ("path_image", ctypes.c_void_p), ("path_cubics", ctypes.c_void_p)]


_sigs_done = False


def lib():
    global _sigs_done
    L = _oracle_lib()
    if not _sigs_done:
        sp = ctypes.POINTER(_Scene)
        vp = ctypes.c_void_p
        L.fd_forward.argtypes = [sp, vp, vp, vp, vp]
        L.fd_forward.restype = ctypes.c_double
        L.fd_backward.argtypes = [sp, vp, vp, vp, vp, vp, vp]
        L.fd_backward.restype = ctypes.c_double
        L.fd_pixel_area.argtypes = [sp, vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]
        L.fd_pixel_area.restype = ctypes.c_double
        L.fd_polygon.argtypes = [sp, vp, ctypes.c_uint32, vp, ctypes.c_uint32]
        L.fd_polygon.restype = ctypes.c_uint32
        _sigs_done = True
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Scene:
    """Layout of a differentiable scene, for the oracle calls."""

    def __init__(self, n_images, width, height, path_image, path_cubics):
        self.n_images, self.width, self.height = int(n_images), int(width), int(height)
        self.path_image = np.ascontiguousarray(path_image, np.uint32)
        self.path_cubics = np.ascontiguousarray(path_cubics, np.uint32)
        self._s = _Scene(self.n_images, self.width, self.height, len(self.path_image),
                         self.path_image.ctypes.data, self.path_cubics.ctypes.data)

    @property
    def n_points(self):
        return int(np.sum(1 + 3 * self.path_cubics.astype(np.int64)))

    def forward(self, points, colours, target=None):
        """(loss, canvas[n_images, H, W, 4]) in FP64."""
        pts = np.ascontiguousarray(points, np.float64).reshape(-1)
        col = np.ascontiguousarray(colours, np.float64).reshape(-1)
        tgt = None if target is None else np.ascontiguousarray(target, np.float64).reshape(-1)
        canvas = np.zeros((self.n_images, self.height, self.width, 4), np.float64)
        loss = lib().fd_forward(ctypes.byref(self._s), _p(pts), _p(col), _p(tgt), _p(canvas))
        return loss, canvas

    def backward(self, points, colours, target=None, grad_canvas=None):
        """(loss, d points, d colours) in FP64; loss gradient when grad_canvas is None."""
        pts = np.ascontiguousarray(points, np.float64).reshape(-1)
        col = np.ascontiguousarray(colours, np.float64).reshape(-1)
        tgt = None if target is None else np.ascontiguousarray(target, np.float64).reshape(-1)
        gc = None if grad_canvas is None else np.ascontiguousarray(grad_canvas, np.float64).reshape(-1)
        gp = np.zeros(pts.shape, np.float64)
        gcol = np.zeros(col.shape, np.float64)
        loss = lib().fd_backward(ctypes.byref(self._s), _p(pts), _p(col), _p(tgt), _p(gc), _p(gp), _p(gcol))
        return loss, gp.reshape(-1, 2), gcol.reshape(-1, 4)

    def pixel_area(self, points, path, x, y):
        pts = np.ascontiguousarray(points, np.float64).reshape(-1)
        return lib().fd_pixel_area(ctypes.byref(self._s), _p(pts), path, x, y)

    def polygon(self, points, path):
        pts = np.ascontiguousarray(points, np.float64).reshape(-1)
        n = 1 + SEG * int(self.path_cubics[path])
        xy = np.zeros(2 * n, np.float64)
        lib().fd_polygon(ctypes.byref(self._s), _p(pts), path, _p(xy), n)
        return xy.reshape(-1, 2)
