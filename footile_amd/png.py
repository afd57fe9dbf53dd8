"""PNG output of a raster, in the role of the reference's example writer
(examples/png/mod.rs:12-27): Matte8 as 8-bit grayscale, Graya8p as gray +
alpha and Rgba8p as RGBA.

Premultiplied channels are divided by alpha (rounded to nearest) and stored
as they are, without an sRGB transfer: the reference converts to SGray8 /
SRgba8 through pix before encoding, a conversion this writer does not
reproduce.  Rows are stored unfiltered in one zlib stream.
"""
import struct
import zlib

import numpy as np

from .pixel import GRAYA8P, MATTE8, RGBA8P

_COLOUR_TYPE = {MATTE8: 0, GRAYA8P: 4, RGBA8P: 6}


def _chunk(kind, data):
    body = kind + data
    return struct.pack(">I", len(data)) + body + struct.pack(">I", zlib.crc32(body) & 0xFFFFFFFF)


def straight_alpha(pixels, fmt):
    """uint8 (h, w, bpp) premultiplied pixels -> straight alpha."""
    px = np.asarray(pixels, dtype=np.uint8)
    if fmt == MATTE8:
        return px.copy()
    out = px.astype(np.uint32)
    a = out[..., -1:]
    col = np.where(a > 0, (out[..., :-1] * 255 + a // 2) // np.maximum(a, 1), 0)
    out[..., :-1] = np.minimum(col, 255)
    return out.astype(np.uint8)


def encode(pixels, fmt):
    """PNG bytes of a (height, width, bpp) raster of format fmt."""
    px = straight_alpha(pixels, fmt)
    h, w = px.shape[0], px.shape[1]
    rows = px.reshape(h, -1)
    raw = b"".join(b"\x00" + rows[y].tobytes() for y in range(h))
    ihdr = struct.pack(">IIBBBBB", w, h, 8, _COLOUR_TYPE[fmt], 0, 0, 0)
    return (b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", ihdr) + _chunk(b"IDAT", zlib.compress(raw, 6))
            + _chunk(b"IEND", b""))


def write(raster, filename):
    """Write a Raster (footile_amd.Raster) to `filename` as PNG."""
    with open(filename, "wb") as f:
        f.write(encode(raster.pixels, raster.fmt))


def decode(data):
    """(pixels uint8 (h, w, channels), colour type) of a PNG this module wrote
    (8-bit, one IDAT, filter 0 rows)."""
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    at, idat = 8, b""
    w = h = ct = None
    while at < len(data):
        n, = struct.unpack(">I", data[at:at + 4])
        kind, body = data[at + 4:at + 8], data[at + 8:at + 8 + n]
        if zlib.crc32(kind + body) & 0xFFFFFFFF != struct.unpack(">I", data[at + 8 + n:at + 12 + n])[0]:
            raise ValueError("bad CRC")
        if kind == b"IHDR":
            w, h, depth, ct = struct.unpack(">IIBB", body[:10])
            assert depth == 8
        elif kind == b"IDAT":
            idat += body
        at += 12 + n
    ch = {0: 1, 4: 2, 6: 4}[ct]
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + w * ch)
    assert np.all(raw[:, 0] == 0)
    return raw[:, 1:].reshape(h, w, ch).copy(), ct
