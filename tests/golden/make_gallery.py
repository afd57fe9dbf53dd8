"""Writes tests/golden/gallery.json: the paths of footile's example programs
(examples/*.rs) as ops, with the calls each program makes and the SHA-256 of
the raster the oracle renders for it.

    python tests/golden/make_gallery.py

The paths are rebuilt with footile_amd.Path2D from the numbers in the
example sources (cited per entry); heptagram's vertices use the C library's
cosf / sinf, the functions Rust's f32::cos / f32::sin call on Linux.  Only
the ops and the expected hashes are stored: the test renders both sides.
"""
import ctypes
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from footile_amd.path import Path2D  # noqa: E402
from oracle import pyoracle as po  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gallery.json")
FMT = {"matte8": po.MATTE8, "rgba8p": po.RGBA8P, "graya8p": po.GRAYA8P}
JOIN = {"miter": 0, "bevel": 1, "round": 2}


def fish():  # examples/fishy.rs:9-18 (also fishy2.rs, fishyp.rs)
    return (Path2D().relative().pen_width(3.0).move_to(112.0, 24.0).line_to(-32.0, 24.0)
            .cubic_to(-96.0, -48.0, -96.0, 80.0, 0.0, 32.0).line_to(32.0, 24.0).line_to(-16.0, -40.0)
            .close().finish())


def eye():  # examples/fishy.rs:19-25
    return (Path2D().relative().pen_width(2.0).move_to(24.0, 48.0).line_to(8.0, 8.0).move_to(0.0, -8.0)
            .line_to(-8.0, 8.0).finish())


def heptagram():  # examples/heptagram.rs:12-24
    m = ctypes.CDLL("libm.so.6")
    m.cosf.argtypes = m.sinf.argtypes = [ctypes.c_float]
    m.cosf.restype = m.sinf.restype = ctypes.c_float
    pi = np.float32(np.pi)
    pb = Path2D().move_to(m.cosf(0.0), m.sinf(0.0))
    for n in range(1, 7):
        th = np.float32(np.float32(np.float32(pi * np.float32(4.0)) * np.float32(n)) / np.float32(7.0))
        pb = pb.line_to(m.cosf(float(th)), m.sinf(float(th)))
    return pb.close().finish()


def letter():  # examples/letter.rs:9-31
    q = [(62.433567, 64.895096, 47.88811, 81.79021), (33.342655, 98.57342, 33.342655, 127.88811),
         (33.342655, 156.86713, 48.44755, 174.54544), (63.664333, 192.11188, 89.51049, 192.11188),
         (122.62937, 192.11188, 139.3007, 159.32866), None,
         (147.02098, 188.41957, 129.34265, 199.04895), (111.77622, 209.67831, 88.503494, 209.67831),
         (64.671326, 209.67831, 47.21678, 199.83215), (29.874126, 189.87411, 20.6993, 171.52448),
         (11.636364, 153.06293, 11.636364, 127.88811), (11.636364, 90.18181, 32.0, 68.81119),
         (52.363636, 47.44055, 88.39161, 47.44055), (113.56643, 47.44055, 130.46153, 57.286713),
         (147.35664, 67.13286, 155.3007, 86.4895), "line2",
         (129.56644, 79.44055, 117.37063, 72.16783), (105.28671, 64.895096, 88.61539, 64.895096)]
    pb = Path2D().absolute().move_to(88.61539, 64.895096)
    for v in q:
        if v is None:
            pb = pb.line_to(156.75525, 168.05594)
        elif v == "line2":
            pb = pb.line_to(135.04895, 93.20279)
        else:
            pb = pb.quad_to(*v)
    return pb.finish()


def examples():
    """(name, source, width, height, fmt, init, transform, join, paths, calls)"""
    rel = lambda: Path2D().relative()  # noqa: E731
    E = []
    E.append(("cubic", "examples/cubic.rs:9-17", 64, 64, "matte8", None, None, None,
              {"p": rel().pen_width(2.0).move_to(8.0, 16.0).cubic_to(64.0, -16.0, 64.0, 48.0, 0.0, 32.0).finish()},
              [("stroke", "p", [255])]))
    E.append(("curve", "examples/curve.rs:9-18", 128, 128, "matte8", None, None, None,
              {"p": rel().pen_width(0.0).move_to(64.0, 48.0).pen_width(18.0)
               .cubic_to(-64.0, -48.0, -64.0, 80.0, 0.0, 32.0).finish()},
              [("stroke", "p", [255])]))
    E.append(("drop", "examples/drop.rs:9-20", 100, 100, "graya8p", None, None, None,
              {"p": rel().pen_width(3.0).move_to(50.0, 34.0).cubic_to(4.0, 16.0, 16.0, 28.0, 0.0, 32.0)
               .cubic_to(-16.0, -4.0, -4.0, -16.0, 0.0, -32.0).close().finish()},
              [("fill", "p", [128, 255], 0), ("stroke", "p", [255, 255])]))
    E.append(("figure", "examples/figure.rs:9-23", 64, 64, "matte8", None, None, None,
              {"p": rel().move_to(4.0, 4.0).line_to(28.0, 12.0).line_to(28.0, -12.0).line_to(-12.0, 28.0)
               .line_to(12.0, 28.0).line_to(-28.0, -4.0).line_to(-28.0, 4.0).line_to(12.0, -28.0).close()
               .finish()},
              [("fill", "p", [255], 0)]))
    fish_calls = [("fill", "fish", [127, 96, 96, 255], 0), ("stroke", "fish", [255, 208, 208, 255]),
                  ("stroke", "eye", [0, 0, 0, 255])]
    E.append(("fishy", "examples/fishy.rs:9-31", 128, 128, "rgba8p", None, None, None,
              {"fish": fish(), "eye": eye()}, fish_calls))
    E.append(("fishy2", "examples/fishy2.rs:27-32 (with_pixels of zeros)", 128, 128, "rgba8p", [0, 0, 0, 0],
              None, None, {"fish": fish(), "eye": eye()}, fish_calls))
    E.append(("heptagram", "examples/heptagram.rs:12-24", 100, 100, "matte8", None,
              [50.0, 0.0, 25.0, 0.0, 50.0, 25.0], None, {"p": heptagram()}, [("fill", "p", [255], 1)]))
    E.append(("letter", "examples/letter.rs:9-33", 165, 256, "matte8", None, None, None, {"p": letter()},
              [("fill", "p", [255], 0)]))
    E.append(("over", "examples/over.rs:9-19", 64, 64, "matte8", None, None, None,
              {"p": rel().pen_width(8.0).move_to(32.0, 16.0).line_to(16.0, 16.0).line_to(-16.0, 16.0)
               .line_to(-16.0, -16.0).line_to(16.0, -16.0).line_to(0.0, 32.0).finish()},
              [("stroke", "p", [255])]))
    E.append(("quad", "examples/quad.rs:9-17", 64, 64, "matte8", None, None, None,
              {"p": rel().pen_width(2.0).move_to(0.0, 16.0).quad_to(100.0, 16.0, 0.0, 32.0).finish()},
              [("stroke", "p", [255])]))
    E.append(("round", "examples/round.rs:9-18", 100, 100, "matte8", None, None, ("round", 4.0),
              {"p": rel().pen_width(40.0).move_to(10.0, 60.0).line_to(50.0, 0.0).line_to(0.0, -50.0).finish()},
              [("stroke", "p", [255])]))
    E.append(("stroke", "examples/stroke.rs:9-18", 64, 64, "matte8", None, None, None,
              {"p": rel().pen_width(5.0).move_to(16.0, 48.0).line_to(32.0, 0.0).line_to(-16.0, -32.0).close()
               .finish()},
              [("stroke", "p", [255])]))
    E.append(("stroke2", "examples/stroke2.rs:9-22 (Raster::with_color)", 64, 64, "rgba8p", [64, 128, 64, 255],
              None, None,
              {"p": rel().pen_width(6.0).move_to(16.0, 15.0).line_to(32.0, 1.0).line_to(-32.0, 1.0)
               .line_to(32.0, 15.0).line_to(-32.0, 15.0).line_to(32.0, 1.0).line_to(-32.0, 1.0).finish()},
              [("stroke", "p", [255, 255, 0, 255])]))
    teeth = rel().move_to(0.0, 8.0)
    for _ in range(4):
        teeth = teeth.line_to(8.0, 8.0).line_to(8.0, -8.0)
    teeth = teeth.move_to(-64.0, 32.0)
    for _ in range(4):
        teeth = teeth.line_to(8.0, 8.0).line_to(8.0, -8.0)
    E.append(("teeth", "examples/teeth.rs:8-30", 64, 64, "matte8", None, None, None, {"p": teeth.finish()},
              [("stroke", "p", [255])]))
    return E


def render_oracle(ex):
    """The oracle's raster bytes for one gallery entry (a dict as stored)."""
    fmt = FMT[ex["fmt"]]
    bpp = {po.MATTE8: 1, po.RGBA8P: 4, po.GRAYA8P: 2}[fmt]
    o = po.Plotter(ex["width"], ex["height"], fmt)
    if ex["init"] is not None:
        o.set_pixels(np.tile(np.asarray(ex["init"], np.uint8), ex["width"] * ex["height"]))
    if ex["transform"] is not None:
        o.set_transform(ex["transform"])
    if ex["join"] is not None:
        o.set_join(JOIN[ex["join"][0]], ex["join"][1])
    for c in ex["calls"]:
        ops = ops_of(ex["paths"][c["path"]])
        clr = tuple(c["clr"]) + (0,) * (4 - len(c["clr"]))
        rc = o.fill(c["rule"], ops, clr) if c["op"] == "fill" else o.stroke(ops, clr)
        assert rc == 0, (ex["name"], c)
    px = o.pixels()
    assert px.size == ex["width"] * ex["height"] * bpp
    return px


def ops_of(rows):
    from footile_amd.path import ops_array
    return ops_array([(int(r[0]), tuple(r[1:])) for r in rows])


def main():
    out = []
    for name, src, w, h, fmt, init, tr, join, paths, calls in examples():
        ex = {"name": name, "source": src, "width": w, "height": h, "fmt": fmt, "init": init, "transform": tr,
              "join": list(join) if join else None,
              "paths": {k: [[int(t)] + [float(x) for x in v] for t, v in zip(a["tag"], a["v"])]
                        for k, a in paths.items()},
              "calls": [{"op": c[0], "path": c[1], "clr": c[2], "rule": c[3] if len(c) > 3 else 0} for c in calls]}
        px = render_oracle(ex)
        ex["covered"] = int((px.reshape(h * w, -1)[:, -1] > 0).sum())
        ex["sha256"] = hashlib.sha256(px.tobytes()).hexdigest()
        out.append(ex)
        print(f"{name:10s} {w}x{h} {fmt:8s} covered {ex['covered']:6d}  {ex['sha256'][:16]}")
    with open(OUT, "w") as f:
        json.dump({"generator": "tests/golden/make_gallery.py", "examples": out}, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
