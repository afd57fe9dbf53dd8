"""Seeded synthetic paths for the rasteriser's measured configurations.

All randomness comes from SplitMix64(seed); uniform f32 = (u >> 40) * 2^-24,
normals by Box-Muller in f64 then rounded to f32.  The same op arrays feed
the GPU library and the CPU oracle.

  plain_256      one closed path, move + shuffled 48 line_to / 16 cubic_to,
                 coords uniform in [8, 248), 256x256 Matte8 (fill NonZero; the
                 same path with pen_width 4 prepended is stroked).      seed 0
  random_walk    move_to(2048, 2048) + 10,000 cubic_to whose control points
                 are a random walk with N(0, 40) steps clamped to the canvas,
                 closed; 4096x4096 Rgba8p, NonZero and EvenOdd.          seed 1
  atlas          16,384 closed paths, lognormal(3.5, 1) segments clamped to
                 [4, 2000], 60% cubic / 30% quad / 10% line, each in its own
                 s x s tile (s = clamp(roundup8(12 sqrt(n)), 32, 256)),
                 shelf-packed into a 16384x16384 Matte8 atlas.          seed 2
"""
import numpy as np

from .path import CLOSE, CUBIC, LINE, MOVE, OP_DTYPE, PEN_WIDTH, QUAD

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    """Vectorised SplitMix64 stream."""

    def __init__(self, seed):
        self.state = np.uint64(seed)

    def next_u64(self, n):
        with np.errstate(over="ignore"):
            k = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + _GOLDEN * k
            self.state = self.state + _GOLDEN * np.uint64(n)
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            return z ^ (z >> np.uint64(31))

    def uniform(self, n):
        """f64 holding an exact f32-grid value in [0, 1)."""
        return (self.next_u64(n) >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)

    def normal(self, n):
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log(1.0 - u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        return z[:n]

    def below(self, bound):
        return int(self.next_u64(1)[0] % np.uint64(bound))


def _op(tag, *vals):
    a = np.zeros(1, dtype=OP_DTYPE)
    a[0]["tag"] = tag
    a[0]["v"][: len(vals)] = vals
    return a


def plain_256(seed=0, lines=48, cubics=16, lo=8.0, hi=248.0):
    """The plain CPU-runnable path (fill ops)."""
    rng = SplitMix64(seed)
    kinds = [LINE] * lines + [CUBIC] * cubics
    for i in range(len(kinds) - 1, 0, -1):  # Fisher-Yates
        j = rng.below(i + 1)
        kinds[i], kinds[j] = kinds[j], kinds[i]
    n = len(kinds) + 2
    ops = np.zeros(n, dtype=OP_DTYPE)
    span = hi - lo
    p = (lo + span * rng.uniform(2)).astype(np.float32)
    ops[0]["tag"] = MOVE
    ops[0]["v"][:2] = p
    for k, kind in enumerate(kinds, start=1):
        ops[k]["tag"] = kind
        m = 2 if kind == LINE else 6
        ops[k]["v"][:m] = (lo + span * rng.uniform(m)).astype(np.float32)
    ops[n - 1]["tag"] = CLOSE
    return ops


def plain_256_stroke(seed=0, width=4.0):
    """pen_width(width) followed by the plain path (stroke ops)."""
    return np.concatenate([_op(PEN_WIDTH, width), plain_256(seed)])


def random_walk(seed=1, n_cubics=10000, canvas=4096, sigma=40.0):
    """Self-intersecting random-walk path of cubics."""
    rng = SplitMix64(seed)
    steps = rng.normal(n_cubics * 6).reshape(n_cubics, 3, 2) * sigma
    hi = np.float64(np.float32(canvas - 0.01))
    ops = np.zeros(n_cubics + 2, dtype=OP_DTYPE)
    ops[0]["tag"] = MOVE
    ops[0]["v"][:2] = (canvas / 2, canvas / 2)
    pen = np.array([canvas / 2, canvas / 2], dtype=np.float64)
    pts = np.empty((n_cubics, 3, 2), dtype=np.float32)
    for i in range(n_cubics):
        for j in range(3):
            pen = np.clip(pen + steps[i, j], 0.0, hi)
            pen = pen.astype(np.float32).astype(np.float64)
            pts[i, j] = pen
    ops[1:-1]["tag"] = CUBIC
    ops[1:-1]["v"] = pts.reshape(n_cubics, 6)
    ops[-1]["tag"] = CLOSE
    return ops


def tile_side(n_seg):
    s = int(np.ceil(12.0 * np.sqrt(n_seg) / 8.0)) * 8
    return min(max(s, 32), 256)


def atlas(seed=2, n_paths=16384, atlas_side=16384, mu=3.5, sigma=1.0):
    """Batched atlas: returns (ops, tiles) where tiles is an int array of
    (op_begin, op_count, x, y, side) rows, shelf-packed into the atlas."""
    rng = SplitMix64(seed)
    z = rng.normal(n_paths)
    n_seg = np.clip(np.rint(np.exp(mu + sigma * z)), 4, 2000).astype(np.int64)
    sides = np.array([tile_side(n) for n in n_seg], dtype=np.int64)
    total_ops = int(n_seg.sum()) + 2 * n_paths
    ops = np.zeros(total_ops, dtype=OP_DTYPE)
    kind_u = rng.uniform(int(n_seg.sum()))
    coords = rng.uniform(int(n_seg.sum()) * 6 + 2 * n_paths)
    tiles = np.zeros((n_paths, 5), dtype=np.int64)
    o = 0
    ku = 0
    cu = 0
    for t in range(n_paths):
        s = float(sides[t])
        lo, span = 2.0, s - 4.0
        begin = o
        ops[o]["tag"] = MOVE
        ops[o]["v"][:2] = (lo + span * coords[cu:cu + 2]).astype(np.float32)
        cu += 2
        o += 1
        for _ in range(int(n_seg[t])):
            u = kind_u[ku]
            ku += 1
            kind = CUBIC if u < 0.6 else (QUAD if u < 0.9 else LINE)
            m = 6 if kind == CUBIC else (4 if kind == QUAD else 2)
            ops[o]["tag"] = kind
            ops[o]["v"][:m] = (lo + span * coords[cu:cu + m]).astype(np.float32)
            cu += 6
            o += 1
        ops[o]["tag"] = CLOSE
        o += 1
        tiles[t, 0] = begin
        tiles[t, 1] = o - begin
        tiles[t, 4] = sides[t]
    # shelf packing, tallest first
    order = np.argsort(-sides, kind="stable")
    x = y = shelf_h = 0
    for t in order:
        s = int(sides[t])
        if x + s > atlas_side:
            y += shelf_h
            x, shelf_h = 0, 0
        if y + s > atlas_side:
            raise ValueError("atlas overflow")
        tiles[t, 2], tiles[t, 3] = x, y
        x += s
        shelf_h = max(shelf_h, s)
    return ops, tiles
