"""PathOp, FillRule and the Path2D builder (src/path.rs).

Ops are kept as a numpy structured array with the 28-byte layout of
``ft_path_op`` (u32 tag + 6 x f32), so they pass to the library unchanged.
"""
import numpy as np

OP_DTYPE = np.dtype([("tag", "<u4"), ("v", "<f4", (6,))], align=False)
CLOSE, MOVE, LINE, QUAD, CUBIC, PEN_WIDTH = range(6)


class FillRule:
    """src/path.rs:8-14"""
    NonZero = 0
    EvenOdd = 1


class PathOp:
    """Constructors of src/path.rs:18-31 (absolute coordinates)."""

    @staticmethod
    def Close():
        return (CLOSE, ())

    @staticmethod
    def Move(x, y):
        return (MOVE, (x, y))

    @staticmethod
    def Line(x, y):
        return (LINE, (x, y))

    @staticmethod
    def Quad(bx, by, cx, cy):
        return (QUAD, (bx, by, cx, cy))

    @staticmethod
    def Cubic(bx, by, cx, cy, dx, dy):
        return (CUBIC, (bx, by, cx, cy, dx, dy))

    @staticmethod
    def PenWidth(w):
        return (PEN_WIDTH, (w,))


def ops_array(ops):
    """Convert a sequence of (tag, values) to the ft_path_op array layout."""
    if isinstance(ops, np.ndarray) and ops.dtype == OP_DTYPE:
        return ops
    a = np.zeros(len(ops), dtype=OP_DTYPE)
    for i, (tag, vals) in enumerate(ops):
        a[i]["tag"] = tag
        a[i]["v"][: len(vals)] = vals
    return a


class Path2D:
    """Builder for PathOps (src/path.rs:44-171).

    Coordinates are relative to the pen by default (path.rs:57); ``close``
    moves the pen back to the origin (path.rs:89-93).  Values are rounded to
    f32 before the pen arithmetic, as the reference stores f32.
    """

    def __init__(self):
        self._ops = []
        self._absolute = False
        self._pen = (np.float32(0.0), np.float32(0.0))

    def absolute(self):
        self._absolute = True
        return self

    def relative(self):
        self._absolute = False
        return self

    def _pt(self, x, y):  # path.rs:80-86
        x, y = np.float32(x), np.float32(y)
        if self._absolute:
            return (x, y)
        return (np.float32(self._pen[0] + x), np.float32(self._pen[1] + y))

    def close(self):
        self._ops.append((CLOSE, ()))
        self._pen = (np.float32(0.0), np.float32(0.0))
        return self

    def move_to(self, x, y):
        pb = self._pt(x, y)
        self._ops.append((MOVE, pb))
        self._pen = pb
        return self

    def line_to(self, x, y):
        pb = self._pt(x, y)
        self._ops.append((LINE, pb))
        self._pen = pb
        return self

    def quad_to(self, bx, by, cx, cy):
        pb = self._pt(bx, by)
        pc = self._pt(cx, cy)
        self._ops.append((QUAD, pb + pc))
        self._pen = pc
        return self

    def cubic_to(self, bx, by, cx, cy, dx, dy):
        pb = self._pt(bx, by)
        pc = self._pt(cx, cy)
        pd = self._pt(dx, dy)
        self._ops.append((CUBIC, pb + pc + pd))
        self._pen = pd
        return self

    def pen_width(self, width):
        self._ops.append((PEN_WIDTH, (np.float32(width),)))
        return self

    def finish(self):
        """The ops as an ft_path_op array."""
        return ops_array(self._ops)
