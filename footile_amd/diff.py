"""Differentiable fill over the C ABI (ft_diff_*).

footile's fill is not differentiable; this is the optimisation-facing form of
the same rasteriser: a float32 analytic-coverage forward and a hand-written
adjoint (see include/footile_hip.h for the model).  Parameters live in device
memory (DeviceBuffer); `DiffFill` has the shape of a torch.autograd.Function
(static forward(ctx, ...) / backward(ctx, grad_output), ctx.save_for_backward)
so an autograd integration is a thin adapter over it.
"""
import ctypes

import numpy as np

from ._lib import check, lib
from .plotter import DeviceBuffer

_vp = ctypes.c_void_p


KEEP_COVERAGE = 1  # FT_DIFF_KEEP_COVERAGE


class _Layout(ctypes.Structure):
    _fields_ = [("n_images", ctypes.c_uint32), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("n_paths", ctypes.c_uint32), ("path_image", ctypes.c_void_p), ("path_cubics", ctypes.c_void_p),
                ("flags", ctypes.c_uint32)]


def device_array(a):
    """Upload a float32 array; returns a DeviceBuffer."""
    a = np.ascontiguousarray(a, np.float32)
    b = DeviceBuffer(max(a.nbytes, 4))
    b.upload(a)
    return b


class DiffRenderer:
    """A scene layout and its device workspace (ft_diff).

    By default the backward recomputes every path's coverage; keep_coverage
    (FT_DIFF_KEEP_COVERAGE) keeps it from the forward instead, at 4 B per path
    and pixel."""

    def __init__(self, n_images, width, height, path_image, path_cubics, keep_coverage=False):
        self.n_images, self.width, self.height = int(n_images), int(width), int(height)
        self.path_image = np.ascontiguousarray(path_image, np.uint32)
        self.path_cubics = np.ascontiguousarray(path_cubics, np.uint32)
        L = _Layout(self.n_images, self.width, self.height, len(self.path_image), self.path_image.ctypes.data,
                    self.path_cubics.ctypes.data, KEEP_COVERAGE if keep_coverage else 0)
        h = _vp()
        check(lib().ft_diff_new(ctypes.byref(L), ctypes.byref(h)), "ft_diff_new")
        self._h = h.value
        npts, npaths, nfl = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().ft_diff_info(_vp(self._h), ctypes.byref(npts), ctypes.byref(npaths), ctypes.byref(nfl)),
              "ft_diff_info")
        self.n_points, self.n_paths, self.canvas_floats = npts.value, npaths.value, nfl.value

    @classmethod
    def for_scene(cls, scene, keep_coverage=False):
        return cls(scene.n_images, scene.width, scene.height, scene.path_image, scene.path_cubics, keep_coverage)

    def __del__(self):
        self.close()

    def close(self):
        if getattr(self, "_h", None):
            lib().ft_diff_free(_vp(self._h))
            self._h = None

    @property
    def canvas_shape(self):
        return (self.n_images, self.height, self.width, 4)

    def new_canvas(self):
        return DeviceBuffer(4 * self.canvas_floats)

    def forward(self, points, colours, canvas):
        check(lib().ft_diff_forward(_vp(self._h), _vp(points.ptr), _vp(colours.ptr), _vp(canvas.ptr)),
              "ft_diff_forward")

    def loss(self, canvas, target, grad_canvas=None):
        out = ctypes.c_double()
        check(lib().ft_diff_loss(_vp(self._h), _vp(canvas.ptr), _vp(target.ptr),
                                 _vp(grad_canvas.ptr) if grad_canvas is not None else None, ctypes.byref(out)),
              "ft_diff_loss")
        return out.value

    def backward(self, points, colours, grad_canvas, grad_points, grad_colours):
        check(lib().ft_diff_backward(_vp(self._h), _vp(points.ptr), _vp(colours.ptr), _vp(grad_canvas.ptr),
                                     _vp(grad_points.ptr), _vp(grad_colours.ptr)), "ft_diff_backward")

    def step(self, points, colours, target, canvas, grad_canvas, grad_points, grad_colours):
        """forward + loss + backward; returns the loss."""
        out = ctypes.c_double()
        check(lib().ft_diff_step(_vp(self._h), _vp(points.ptr), _vp(colours.ptr), _vp(target.ptr), _vp(canvas.ptr),
                                 _vp(grad_canvas.ptr), _vp(grad_points.ptr), _vp(grad_colours.ptr),
                                 ctypes.byref(out)), "ft_diff_step")
        return out.value


class Context:
    """The ctx argument of DiffFill.forward / backward."""

    def __init__(self):
        self.saved_tensors = ()
        self.needs_input_grad = (False, True, True)

    def save_for_backward(self, *bufs):
        self.saved_tensors = bufs


class DiffFill:
    """canvas = DiffFill.apply(renderer, points, colours); grads = DiffFill.backward(ctx, grad_canvas)."""

    @staticmethod
    def forward(ctx, renderer, points, colours):
        canvas = renderer.new_canvas()
        renderer.forward(points, colours, canvas)
        ctx.renderer = renderer
        ctx.save_for_backward(points, colours)
        return canvas

    @staticmethod
    def backward(ctx, grad_canvas):
        points, colours = ctx.saved_tensors
        r = ctx.renderer
        gp = DeviceBuffer(8 * r.n_points)
        gc = DeviceBuffer(16 * r.n_paths)
        r.backward(points, colours, grad_canvas, gp, gc)
        return None, gp, gc

    @classmethod
    def apply(cls, *args):
        ctx = Context()
        return cls.forward(ctx, *args), ctx
