"""footile_amd -- MI355X (gfx950) fill/stroke rasteriser with footile's API.

Python mirror of footile 0.7.0's public surface (src/lib.rs:39-41):
``Path2D``, ``PathOp``, ``FillRule``, ``JoinStyle`` and ``Plotter``, plus the
pix raster formats the plotter accepts (``Matte8``, ``Rgba8p``, ``Graya8p``).
Every fill and stroke runs in libfootile_hip.so on the GPU; there is no CPU
fallback: without the built library (``python -m footile_amd.build``) the
first plotter call raises.
"""
from ._lib import FootileError, lib, library_path  # noqa: F401
from .path import (CLOSE, CUBIC, LINE, MOVE, OP_DTYPE, PEN_WIDTH, QUAD, FillRule, Path2D,  # noqa: F401
                   PathOp, ops_array)
from .plotter import Batch, FillDesc, JoinStyle, Plotter, Raster, fig_fill  # noqa: F401
from .pixel import GRAYA8P, MATTE8, RGBA8P, Graya8p, Matte8, Rgba8p  # noqa: F401

__all__ = [
    "FillRule", "Path2D", "PathOp", "Plotter", "JoinStyle", "Raster", "Batch", "FillDesc",
    "Matte8", "Rgba8p", "Graya8p", "MATTE8", "RGBA8P", "GRAYA8P", "fig_fill", "ops_array",
    "FootileError",
]
