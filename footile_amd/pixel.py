"""Pixel formats of pix 0.14 that footile's Plotter accepts (src/plotter.rs:89-92).

A pixel value is a small tuple of channel bytes; the raster stores them
packed, row-major, channel order as pix lays them out.
"""
MATTE8, RGBA8P, GRAYA8P = 0, 1, 2
BPP = {MATTE8: 1, RGBA8P: 4, GRAYA8P: 2}


class _Pixel(tuple):
    fmt = None

    def bytes4(self):
        b = list(self) + [0] * (4 - len(self))
        return bytes(int(v) & 0xFF for v in b[:4])


class Matte8(_Pixel):
    """pix::matte::Matte8 -- one alpha channel."""
    fmt = MATTE8

    def __new__(cls, a=255):
        return super().__new__(cls, (a,))


class Rgba8p(_Pixel):
    """pix::rgb::Rgba8p -- premultiplied red, green, blue, alpha."""
    fmt = RGBA8P

    def __new__(cls, r, g, b, a):
        return super().__new__(cls, (r, g, b, a))


class Graya8p(_Pixel):
    """pix::gray::Graya8p -- premultiplied value, alpha."""
    fmt = GRAYA8P

    def __new__(cls, v, a):
        return super().__new__(cls, (v, a))
