"""Run the device stage code on the host (tests/cpu_mirror) against the oracle.

Run under AddressSanitizer by tests/test_cpu_mirror.py:
    LD_PRELOAD=<libasan.so> MIRROR_ASAN=1 python tests/cpu_mirror/check_cases.py
Exits non-zero on the first raster or count that differs from the oracle;
AddressSanitizer aborts the process on any out-of-range access.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from footile_amd import workloads  # noqa: E402
from footile_amd.path import CLOSE, CUBIC, LINE, MOVE, PEN_WIDTH, QUAD, ops_array  # noqa: E402
from footile_amd.pixel import BPP, GRAYA8P, MATTE8, RGBA8P  # noqa: E402
from oracle import pyoracle as po  # noqa: E402
from tests.cpu_mirror import mirror  # noqa: E402

FAILS = []


def check(tag, w, h, fmt, rule, ops, clr=(255, 255, 255, 255), init=None, transform=None, tol=None):
    o = po.Plotter(w, h, fmt)
    if init is not None:
        o.set_pixels(init)
    if transform is not None:
        o.set_transform(transform)
    if tol is not None:
        o.set_tolerance(tol)
    rc_o = o.fill(rule, ops, clr)
    kw = {}
    if transform is not None:
        kw["transform"] = transform
    if tol is not None:
        kw["tolerance"] = tol
    rc, px, counts = mirror.fill(w, h, fmt, rule, ops, clr, init=init, **kw)
    if rc != rc_o:
        FAILS.append(f"{tag}: status {rc} vs oracle {rc_o}")
        return
    if rc:
        return
    want = o.pixels()
    if not np.array_equal(px, want):
        bad = np.flatnonzero(px != want)
        FAILS.append(f"{tag}: {bad.size} bytes differ, first {bad[:5]}")
    elif counts != o.last_counts():
        FAILS.append(f"{tag}: counts {counts} vs oracle {o.last_counts()}")


def check_stroke(tag, w, h, fmt, ops, clr=(255, 255, 255, 255), transform=None, tol=None, join=None):
    o = po.Plotter(w, h, fmt)
    if transform is not None:
        o.set_transform(transform)
    if tol is not None:
        o.set_tolerance(tol)
    if join is not None:
        o.set_join(*join)
    rc_o = o.stroke(ops, clr)
    kw = {}
    if transform is not None:
        kw["transform"] = transform
    if tol is not None:
        kw["tolerance"] = tol
    if join is not None:
        kw["join"] = join
    rc, px, counts = mirror.stroke(w, h, fmt, ops, clr, **kw)
    if rc != rc_o:
        FAILS.append(f"{tag}: status {rc} vs oracle {rc_o}")
        return
    if rc:
        return
    want = o.pixels()
    if not np.array_equal(px, want):
        bad = np.flatnonzero(px != want)
        FAILS.append(f"{tag}: {bad.size} bytes differ, first {bad[:5]}")
    elif counts != o.last_counts():
        FAILS.append(f"{tag}: counts {counts} vs oracle {o.last_counts()}")


def rect(x0, y0, x1, y1, ccw=False):
    pts = [(x0, y0), (x1, y0), (x1, y1), (x0, y1)]
    if ccw:
        pts = pts[::-1]
    return [(MOVE, pts[0])] + [(LINE, p) for p in pts[1:]] + [(CLOSE, ())]


def random_path(rng, w, h, n_sub=1, kinds=(LINE,), margin=10.0, close_prob=1.0):
    ops = []
    for _ in range(n_sub):
        n = int(rng.integers(2, 24))
        ops.append((MOVE, tuple(rng.uniform(-margin, [w + margin, h + margin]).astype(np.float32))))
        for _ in range(n):
            kind = kinds[int(rng.integers(0, len(kinds)))]
            m = {LINE: 1, QUAD: 2, CUBIC: 3}[kind]
            pts = rng.uniform(-margin, [w + margin, h + margin], size=(m, 2)).astype(np.float32)
            ops.append((kind, tuple(pts.reshape(-1))))
        if rng.uniform() < close_prob:
            ops.append((CLOSE, ()))
    return ops_array(ops)


def main(quick=False):
    kats = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_kats.json")))
    k = kats["plotter_overlapping"]
    tags = {"move": MOVE, "line": LINE, "cubic": CUBIC}
    check("overlapping", 16, 16, MATTE8, 0, ops_array([(tags[o[0]], tuple(o[1:])) for o in k["ops"]]))
    for i, ops in enumerate([rect(3, 4, 17, 9), rect(2.5, 0, 9, 8), rect(-5, 3, 40, 11),
                             rect(2, 2, 30, 30) + rect(10, 10, 20, 20, True), rect(4, -6, 12, 5),
                             rect(4, 50, 12, 60)]):
        for rule in (0, 1):
            check(f"rect{i}/{rule}", 32, 24, MATTE8, rule, ops_array(ops))
    rng = np.random.default_rng(1234)
    for t in range(20 if quick else 60):
        w, h = int(rng.integers(1, 300)), int(rng.integers(1, 120))
        check(f"poly{t}", w, h, MATTE8, t % 2, random_path(rng, w, h, n_sub=int(rng.integers(1, 4))))
    rng = np.random.default_rng(99)
    for t in range(15 if quick else 45):
        fmt = (MATTE8, RGBA8P, GRAYA8P)[t % 3]
        w, h = int(rng.integers(8, 200)), int(rng.integers(8, 100))
        init = rng.integers(0, 256, size=w * h * BPP[fmt]).astype(np.uint8) if t % 2 else None
        clr = tuple(int(v) for v in rng.integers(0, 256, size=4))
        ops = random_path(rng, w, h, n_sub=int(rng.integers(1, 3)), kinds=(LINE, QUAD, CUBIC))
        check(f"curve{t}", w, h, fmt, t % 2, ops, clr, init=init)
    check("plain_256", 256, 256, MATTE8, 0, workloads.plain_256(0))
    fish = ops_array([(PEN_WIDTH, (3.0,)), (MOVE, (112.0, 24.0)), (LINE, (80.0, 48.0)),
                      (CUBIC, (-16.0, 0.0, -16.0, 128.0, 80.0, 80.0)), (LINE, (112.0, 104.0)),
                      (LINE, (96.0, 64.0)), (CLOSE, ())])
    check("fish scale 2", 256, 256, MATTE8, 0, fish, transform=(2, 0, 0, 0, 2, 0))
    check("fish tol", 128, 128, RGBA8P, 1, fish, (127, 96, 96, 255), tol=0.05)
    # strokes: the stroker stages against the oracle's stroker
    check_stroke("plain_256 stroke", 256, 256, MATTE8, workloads.plain_256_stroke(0))
    check_stroke("fish stroke scale 2", 256, 256, MATTE8, fish, transform=(2, 0, 0, 0, 2, 0))
    check_stroke("fish stroke rgba", 128, 128, RGBA8P, fish, (255, 208, 208, 255), tol=0.05)
    rng = np.random.default_rng(5)
    for t in range(12 if quick else 36):
        ops = random_path(rng, 96, 96, n_sub=2, kinds=(LINE, QUAD, CUBIC), margin=-8.0, close_prob=0.5)
        ops = np.concatenate([ops_array([(PEN_WIDTH, (float(rng.uniform(1, 9)),))]), ops])
        join = [(0, 4.0), (1, 0.0), (2, 0.0), (0, 1.5)][t % 4]
        check_stroke(f"stroke{t} join{join}", 96, 96, MATTE8, ops, join=join)
    check_stroke("empty stroke", 32, 32, MATTE8, ops_array([(MOVE, (3.0, 3.0)), (CLOSE, ())]))
    # rotations, shears, reflections and translations
    rng = np.random.default_rng(91)
    for t in range(4 if quick else 8):
        a = rng.uniform(0, 2 * np.pi)
        sx, sy = rng.uniform(0.6, 1.4, size=2) * rng.choice([-1.0, 1.0], size=2)
        sh = rng.uniform(-0.5, 0.5)
        m = np.array([[np.cos(a), -np.sin(a)], [np.sin(a), np.cos(a)]]) @ np.array([[sx, sh], [0.0, sy]])
        off = np.array([64.0, 64.0]) - m @ np.array([48.0, 48.0]) + rng.uniform(-8, 8, size=2)
        e = tuple(float(v) for v in np.float32([m[0, 0], m[0, 1], off[0], m[1, 0], m[1, 1], off[1]]))
        ops = random_path(rng, 96, 96, n_sub=2, kinds=(LINE, QUAD, CUBIC), margin=-8.0, close_prob=0.5)
        ops = np.concatenate([ops_array([(PEN_WIDTH, (float(rng.uniform(1, 6)),))]), ops])
        fmt, clr = (MATTE8, (255, 255, 255, 255)) if t % 2 == 0 else (RGBA8P, (150, 90, 40, 200))
        check(f"affine{t}", 128, 128, fmt, t % 2, ops, clr, transform=e)
        check_stroke(f"affine stroke{t}", 128, 128, fmt, ops, clr, transform=e)
    # the 65,535-point cap and its close-pop interplay
    rng = np.random.default_rng(8)
    n = 66000
    pts = np.stack([rng.uniform(0, 250, n), rng.uniform(0, 250, n)], axis=1).astype(np.float32)
    pts[65534] = pts[0]
    base = [(MOVE, tuple(pts[0]))] + [(LINE, tuple(p)) for p in pts[1:]] + [(CLOSE, ())]
    tail = [(MOVE, (5.0, 5.0)), (LINE, (60.0, 9.0)), (LINE, (30.0, 70.0)), (CLOSE, ())]
    for i, ops in enumerate([base, base + tail, base[:65536] + [(CLOSE, ())] + tail,
                             base[:65535] + [(LINE, tuple(pts[0])), (CLOSE, ())] + tail]):
        check(f"cap{i}", 256, 256, MATTE8, 0, ops_array(ops))
    if FAILS:
        print("\n".join(FAILS))
        sys.exit(1)
    print("mirror ok", flush=True)


if __name__ == "__main__":
    main(quick="--quick" in sys.argv)
