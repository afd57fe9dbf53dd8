"""GPU parity: libfootile_hip.so (through its C ABI) against the CPU oracle.

Integer work (Fixed scan, accumulate, composite) must be bit-exact; the f32
flattening must emit the same points, so rasters are compared byte-for-byte.
"""
import numpy as np
import pytest

from footile_amd import workloads
from footile_amd.path import CLOSE, CUBIC, LINE, MOVE, PEN_WIDTH, QUAD, ops_array
from footile_amd.pixel import BPP, GRAYA8P, MATTE8, RGBA8P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ft():
    import footile_amd
    footile_amd.lib()
    return footile_amd


def gpu_plotter(ft, w, h, fmt, init=None, transform=None, tol=None, join=None):
    r = ft.Raster(w, h, fmt, init)
    p = ft.Plotter(r)
    if transform is not None:
        p.set_transform(transform)
    if tol is not None:
        p.set_tolerance(tol)
    if join is not None:
        p.set_join(join)
    return p


def oracle_plotter(oracle, w, h, fmt, init=None, transform=None, tol=None, join=None):
    p = oracle.Plotter(w, h, fmt)
    if init is not None:
        p.set_pixels(init)
    if transform is not None:
        p.set_transform(transform)
    if tol is not None:
        p.set_tolerance(tol)
    if join is not None:
        p.set_join(*join)
    return p


def compare_fill(ft, oracle, w, h, fmt, calls, init=None, transform=None, tol=None, join=None, tag=""):
    """Run the same sequence of ("fill", rule, ops, clr) / ("stroke", ops, clr)
    calls on both plotters and compare every byte after each call."""
    g = gpu_plotter(ft, w, h, fmt, init, transform, tol, join)
    o = oracle_plotter(oracle, w, h, fmt, init, transform, tol, join)
    for k, c in enumerate(calls):
        if c[0] == "fill":
            _, rule, ops, clr = c
            g.fill(rule, ops, clr)
            assert o.fill(rule, ops, clr) == 0
        else:
            _, ops, clr = c
            g.stroke(ops, clr)
            assert o.stroke(ops, clr) == 0
        got = g.raster().as_u8_slice()
        want = o.pixels()
        if not np.array_equal(got, want):
            bad = np.flatnonzero(got != want)
            raise AssertionError(f"{tag} call {k}: {bad.size} bytes differ, first at {bad[:8]} "
                                 f"got {got[bad[:8]]} want {want[bad[:8]]}")
    return g, o


def test_fig_known_answers(ft, kats):
    for k in kats["fig"]:
        fmt = MATTE8 if k["fmt"] == "matte8" else RGBA8P
        n = len(k["points"])
        closes = [0] * (n - 1) + [1]
        clr = list(k["clr"]) + [0] * (4 - len(k["clr"]))
        px = ft.fig_fill(0, k["points"], closes, k["w"], k["h"], fmt, clr)
        assert list(px) == k["expect"], k["name"]


def test_overlapping_completes(ft, oracle, kats):
    k = kats["plotter_overlapping"]
    tags = {"move": MOVE, "line": LINE, "cubic": CUBIC}
    ops = ops_array([(tags[o[0]], tuple(o[1:])) for o in k["ops"]])
    compare_fill(ft, oracle, k["w"], k["h"], MATTE8, [("fill", 0, ops, (255,))], tag="overlapping")


def _rect(x0, y0, x1, y1, ccw=False):
    pts = [(x0, y0), (x1, y0), (x1, y1), (x0, y1)]
    if ccw:
        pts = pts[::-1]
    return [(MOVE, pts[0])] + [(LINE, p) for p in pts[1:]] + [(CLOSE, ())]


def test_rectangles_holes_and_clipping(ft, oracle):
    cases = [_rect(3, 4, 17, 9), _rect(2.5, 0, 9, 8), _rect(-5, 3, 40, 11),
             _rect(2, 2, 30, 30) + _rect(10, 10, 20, 20, ccw=True),
             _rect(2, 2, 30, 30) + _rect(10, 10, 20, 20),
             _rect(4, -6, 12, 5),            # figure above the raster top (shifted rows)
             _rect(4, 50, 12, 60)]           # figure below the raster
    for i, ops in enumerate(cases):
        for rule in (0, 1):
            compare_fill(ft, oracle, 32, 24, MATTE8, [("fill", rule, ops_array(ops), (255,))],
                         tag=f"rect{i}/{rule}")


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


def test_random_polygons_matte8(ft, oracle):
    rng = np.random.default_rng(1234)
    for t in range(40):
        w = int(rng.integers(1, 300))
        h = int(rng.integers(1, 120))
        ops = random_path(rng, w, h, n_sub=int(rng.integers(1, 4)))
        compare_fill(ft, oracle, w, h, MATTE8, [("fill", t % 2, ops, (255,))], tag=f"poly{t} {w}x{h}")


def test_random_curves_all_formats(ft, oracle):
    rng = np.random.default_rng(99)
    for t in range(30):
        fmt = (MATTE8, RGBA8P, GRAYA8P)[t % 3]
        w = int(rng.integers(8, 200))
        h = int(rng.integers(8, 100))
        init = rng.integers(0, 256, size=w * h * BPP[fmt]).astype(np.uint8) if t % 2 else None
        clr = tuple(int(v) for v in rng.integers(0, 256, size=4))
        if fmt == RGBA8P:
            clr = (min(clr[0], clr[3]), min(clr[1], clr[3]), min(clr[2], clr[3]), clr[3])
        ops = random_path(rng, w, h, n_sub=int(rng.integers(1, 3)), kinds=(LINE, QUAD, CUBIC))
        compare_fill(ft, oracle, w, h, fmt, [("fill", t % 2, ops, clr)], init=init, tag=f"curve{t}")


def test_opaque_fills_over_existing_pixels(ft, oracle):
    """Opaque colours over existing pixels, with figures wide enough that whole
    wavefronts of pixels are fully covered (those store the colour's entry
    without the blend) or uncovered (left as they are over transparent
    pixels, blended over the others): Rgba8p and Graya8p, both rules."""
    rng = np.random.default_rng(7)
    w, h = 1400, 300
    rect = [(MOVE, (20.5, 10.0)), (LINE, (1380.0, 14.0)), (LINE, (1375.5, 280.0)), (LINE, (30.0, 290.0)),
            (CLOSE, ())]
    star = [(MOVE, (700.0, 5.0))]
    for k in range(1, 7):
        star.append((LINE, (700.0 + 690.0 * np.sin(k * 6 * np.pi / 7), 150.0 - 145.0 * np.cos(k * 6 * np.pi / 7))))
    star.append((CLOSE, ()))
    blob = [(MOVE, (100.0, 150.0)), (CUBIC, (100.0, -60.0, 1300.0, -60.0, 1300.0, 150.0)),
            (CUBIC, (1300.0, 360.0, 100.0, 360.0, 100.0, 150.0)), (CLOSE, ())]
    for fmt in (RGBA8P, GRAYA8P):
        for init_kind in ("random", "transparent_band"):
            init = rng.integers(0, 256, size=w * h * BPP[fmt]).astype(np.uint8)
            if init_kind == "transparent_band":
                init.reshape(h, -1)[:, : 600 * BPP[fmt]] = 0  # transparent left part
            clr = (200, 120, 40, 255) if fmt == RGBA8P else (90, 255, 0, 0)
            for rule in (0, 1):
                compare_fill(ft, oracle, w, h, fmt,
                             [("fill", rule, ops_array(rect), clr), ("fill", rule, ops_array(star), clr),
                              ("fill", rule, ops_array(blob), clr)],
                             init=init, tag=f"opaque fmt {fmt} {init_kind} rule {rule}")


def test_plain_256_fill_and_stroke(ft, oracle):
    compare_fill(ft, oracle, 256, 256, MATTE8, [("fill", 0, workloads.plain_256(0), (255,))],
                 tag="plain fill")
    compare_fill(ft, oracle, 256, 256, MATTE8, [("stroke", workloads.plain_256_stroke(0), (255,))],
                 tag="plain stroke")


def test_transform_tolerance_and_state(ft, oracle):
    fish = ft.Path2D().relative().pen_width(3.0).move_to(112.0, 24.0).line_to(-32.0, 24.0) \
        .cubic_to(-96.0, -48.0, -96.0, 80.0, 0.0, 32.0).line_to(32.0, 24.0) \
        .line_to(-16.0, -40.0).close().finish()
    compare_fill(ft, oracle, 256, 256, MATTE8,
                 [("fill", 0, fish, (255,)), ("stroke", fish, (255,)), ("fill", 1, fish, (255,))],
                 transform=(2, 0, 0, 0, 2, 0), tag="fishy scale 2")
    compare_fill(ft, oracle, 128, 128, RGBA8P,
                 [("fill", 0, fish, (127, 96, 96, 255)), ("stroke", fish, (255, 208, 208, 255))],
                 tol=0.05, tag="fishyp")


def test_general_affine_transforms(ft, oracle):
    """Rotations, shears, reflections and translations (src/plotter.rs:168-172)
    applied to curve paths: the device evaluates e0*x + e1*y + e2 in the same
    f32 order as the oracle, so fills and strokes stay bit-exact."""
    rng = np.random.default_rng(91)
    for t in range(8):
        a = rng.uniform(0, 2 * np.pi)
        sx, sy = rng.uniform(0.6, 1.4, size=2) * rng.choice([-1.0, 1.0], size=2)
        sh = rng.uniform(-0.5, 0.5)
        c, s = np.cos(a), np.sin(a)
        m = np.array([[c, -s], [s, c]]) @ np.array([[sx, sh], [0.0, sy]])
        off = np.array([64.0, 64.0]) - m @ np.array([48.0, 48.0]) + rng.uniform(-8, 8, size=2)
        e = tuple(float(v) for v in np.float32([m[0, 0], m[0, 1], off[0], m[1, 0], m[1, 1], off[1]]))
        ops = random_path(rng, 96, 96, n_sub=2, kinds=(LINE, QUAD, CUBIC), margin=-8.0, close_prob=0.5)
        ops = np.concatenate([ops_array([(PEN_WIDTH, (float(rng.uniform(1, 6)),))]), ops])
        if t % 2 == 0:
            fmt, clr = MATTE8, (255,)
        else:
            fmt, clr = RGBA8P, (150, 90, 40, 200)
        compare_fill(ft, oracle, 128, 128, fmt,
                     [("fill", t % 2, ops, clr), ("stroke", ops, clr), ("fill", 1 - t % 2, ops, clr)],
                     transform=e, tag=f"affine{t} {e}")


def test_stroke_joins(ft, oracle):
    rng = np.random.default_rng(5)
    for t in range(12):
        ops = random_path(rng, 96, 96, n_sub=2, kinds=(LINE, QUAD, CUBIC), margin=-8.0,
                          close_prob=0.5)
        ops = np.concatenate([ops_array([(PEN_WIDTH, (float(rng.uniform(1, 9)),))]), ops])
        join = [(0, 4.0), (1, 0.0), (2, 0.0), (0, 1.5)][t % 4]
        compare_fill(ft, oracle, 96, 96, MATTE8, [("stroke", ops, (255,))], join=join, tag=f"stroke{t}")


@pytest.mark.parametrize("w", [20000, 32768])
def test_maximum_width(ft, oracle, w):
    """Rasters wider than 16,384 px up to the widest the device path takes,
    FT_MAX_WIDTH = 32,768 px (one row of i32 cells is 128 KiB of LDS), in
    every format with both rules and a stroke; one pixel wider is FT_ENOTSUP,
    returned before any device work."""
    rng = np.random.default_rng(w)
    h = 6
    for t, fmt in enumerate((MATTE8, RGBA8P, GRAYA8P)):
        ops = random_path(rng, w, h, n_sub=2, kinds=(LINE, QUAD, CUBIC), margin=20.0, close_prob=0.7)
        clr = {MATTE8: (255,), RGBA8P: (200, 120, 60, 230), GRAYA8P: (90, 200)}[fmt]
        compare_fill(ft, oracle, w, h, fmt, [("fill", 0, ops, clr), ("stroke", ops, clr), ("fill", 1, ops, clr)],
                     tag=f"width {w} fmt {fmt}")
    with pytest.raises(ft.FootileError) as e:
        ft.Plotter(ft.Raster.with_clear(32768 + 1, 4))
    assert e.value.code == 5


def test_stroke_join_geometry(ft, oracle):
    """The geometric stroke cases that pin the oracle's stroker (closed
    squares, an acute V either side of its miter limit 1.8028, random round
    turns): the device stroker gives the same bytes."""
    square = [(MOVE, (10.0, 10.0)), (LINE, (40.0, 10.0)), (LINE, (40.0, 30.0)), (LINE, (10.0, 30.0)), (CLOSE, ())]
    vee = [(MOVE, (4.0, 30.0)), (LINE, (20.0, 6.0)), (LINE, (36.0, 30.0))]
    for shape, width, w, h in [(square, 10.0, 52, 42), (vee, 4.0, 40, 36)]:
        ops = ops_array([(PEN_WIDTH, (width,))] + shape)
        for join in [(0, 4.0), (0, 1.85), (0, 1.75), (0, 1.5), (0, 1.3), (1, 0.0), (2, 0.0)]:
            compare_fill(ft, oracle, w, h, MATTE8, [("stroke", ops, (255,))], join=join, tag=f"join {join}")
    rng = np.random.default_rng(3)
    for t in range(24):
        pts = [tuple(float(v) for v in rng.integers(6, 42, size=2)) for _ in range(3)]
        ops = ops_array([(PEN_WIDTH, (float(rng.integers(4, 14)),)), (MOVE, pts[0]), (LINE, pts[1]), (LINE, pts[2])])
        compare_fill(ft, oracle, 48, 48, MATTE8, [("stroke", ops, (255,))], join=(2, 0.0), tag=f"turn{t} {pts}")


def test_sequential_fills_keep_state(ft, oracle):
    rng = np.random.default_rng(77)
    for fmt in (MATTE8, RGBA8P):
        calls = [("fill", i % 2, random_path(rng, 64, 48, kinds=(LINE, CUBIC)), (40, 80, 120, 200))
                 for i in range(4)]
        compare_fill(ft, oracle, 64, 48, fmt, calls, tag=f"sequence fmt{fmt}")


def test_random_walk_4096(ft, oracle):
    ops = workloads.random_walk(1)
    clr = (200, 120, 40, 255)
    for rule in (0, 1):
        g, o = compare_fill(ft, oracle, 4096, 4096, RGBA8P, [("fill", rule, ops, clr)],
                            tag=f"random walk rule {rule}")
        assert g.last_counts() == o.last_counts()
        assert g.last_counts()[0] > 65535 and g.last_counts()[1] <= 65535


def test_point_cap_close_pop(ft, oracle):
    """The 65,535-point cap (src/vid.rs:17, src/fig.rs:429-431) and its
    interplay with the closing pop (src/fig.rs:373-383)."""
    rng = np.random.default_rng(8)
    n = 66000
    pts = np.stack([rng.uniform(0, 250, n), rng.uniform(0, 250, n)], axis=1).astype(np.float32)
    pts[65534] = pts[0]  # the point stored last under the cap equals the run's first
    base = [(MOVE, tuple(pts[0]))] + [(LINE, tuple(p)) for p in pts[1:]] + [(CLOSE, ())]
    tail = [(MOVE, (5.0, 5.0)), (LINE, (60.0, 9.0)), (LINE, (30.0, 70.0)), (CLOSE, ())]
    for ops in (base, base + tail, base[:65536] + [(CLOSE, ())] + tail,
                base[:65535] + [(LINE, tuple(pts[0])), (CLOSE, ())] + tail):
        a = ops_array(ops)
        g, o = compare_fill(ft, oracle, 256, 256, MATTE8, [("fill", 0, a, (255,))], tag="cap")
        assert g.last_counts() == o.last_counts()


def test_nan_curve_is_erange_and_leaves_raster(ft, oracle):
    ops = ops_array([(MOVE, (10.0, 20.0)), (CUBIC, (float("nan"), 24.0, 16.0, 28.0, 19.0, 32.0)),
                     (CLOSE, ())])
    init = np.arange(64 * 64, dtype=np.uint32).astype(np.uint8)
    g = gpu_plotter(ft, 64, 64, MATTE8, init)
    with pytest.raises(ft.FootileError) as e:
        g.fill(0, ops)
    assert e.value.code == 4
    assert np.array_equal(g.raster().as_u8_slice(), init)


def test_deterministic_repeat(ft):
    ops = workloads.random_walk(3, n_cubics=3000)
    out = []
    for _ in range(2):
        g = gpu_plotter(ft, 2048, 2048, MATTE8)
        g.fill(1, ops)
        out.append(g.raster().as_u8_slice().copy())
    assert np.array_equal(out[0], out[1])


def test_batch_matches_individual_plotters(ft, oracle):
    from footile_amd.plotter import Batch, DeviceBuffer, FillDesc
    ops, tiles = workloads.atlas(2, n_paths=160, atlas_side=2048)
    side = 2048
    atlas = DeviceBuffer(side * side)
    atlas.upload(np.full(side * side, 77, dtype=np.uint8))  # fresh tiles must not show it
    descs = [FillDesc(int(b), int(c), atlas.ptr + int(y) * side + int(x), int(s), int(s), side,
                      fresh=True) for b, c, x, y, s in tiles]
    batch = Batch(ops, descs)
    batch.run()
    img = atlas.download().reshape(side, side)
    offered, stored, edges = batch.counts()
    for t, (b, c, x, y, s) in enumerate(tiles):
        o = oracle.Plotter(int(s), int(s))
        assert o.fill(0, ops[b:b + c]) == 0
        want = o.pixels().reshape(s, s)
        assert np.array_equal(img[y:y + s, x:x + s], want), f"tile {t}"
        assert (offered[t], stored[t]) == o.last_counts()


def _atlas_batch(n_paths=96, side=1024):
    from footile_amd.plotter import Batch, DeviceBuffer, FillDesc
    ops, tiles = workloads.atlas(2, n_paths=n_paths, atlas_side=side)
    atlas = DeviceBuffer(side * side)
    atlas.upload(np.full(side * side, 77, dtype=np.uint8))
    descs = [FillDesc(int(b), int(c), atlas.ptr + int(y) * side + int(x), int(s), int(s), side,
                      fresh=True) for b, c, x, y, s in tiles]
    return Batch(ops, descs), atlas, descs, ops, tiles


def test_batch_tuning_keeps_rasters(ft, oracle):
    """Every raster geometry gives the same bytes (and the oracle's)."""
    batch, atlas, descs, ops, tiles = _atlas_batch()
    batch.run()
    want = atlas.download()
    for t, (b, c, x, y, s) in enumerate(tiles[:24]):
        o = oracle.Plotter(int(s), int(s))
        assert o.fill(0, ops[b:b + c]) == 0
        assert np.array_equal(want.reshape(1024, 1024)[y:y + s, x:x + s].reshape(-1), o.pixels()), f"tile {t}"
    for band_rows, cells, pf, tg in [(1, 1024, True, 0), (4, 2048, False, 0), (64, 16384, True, 0),
                                     (256, 8192, False, 0), (16, 4096, True, 0), (16, 4096, True, 1),
                                     (8, 2048, False, 2), (16, 8192, True, 8), (4, 1024, True, 64),
                                     (16, 32768, True, 4)]:
        try:
            batch.set_tuning(band_rows, cells, pf, tg)
        except ft.FootileError as e:
            assert e.code == 5 and cells > 16384  # more LDS per block than the device has
            continue
        assert batch.tuning() == (band_rows, cells, pf, tg)
        atlas.upload(np.full(1024 * 1024, 77, dtype=np.uint8))
        batch.run()
        assert np.array_equal(atlas.download(), want), (band_rows, cells, tg)
    with pytest.raises(ft.FootileError):
        batch.set_tuning(0, 4096)
    with pytest.raises(ft.FootileError):
        batch.set_tuning(16, 1000)


def test_batch_graph_replays_match_eager(ft, oracle):
    """Graph mode: several replays with eager work and synchronisation in
    between; every replay equals the eager result."""
    batch, atlas, descs, ops, tiles = _atlas_batch()
    batch.run()
    want = atlas.download()
    offered, stored, _, tops = batch.results()
    batch.set_graph(True)
    other = ft.Plotter(ft.Raster.with_clear(128, 96))
    for rep in range(4):
        atlas.upload(np.full(1024 * 1024, 77, dtype=np.uint8))  # pixels outside the tiles keep it
        other.fill(0, workloads.plain_256(rep))  # eager work on another stream between replays
        batch.launch()
        batch.sync()
        assert np.array_equal(atlas.download(), want), f"replay {rep}"
        o2, s2, _, t2 = batch.results()
        assert np.array_equal(o2, offered) and np.array_equal(s2, stored) and np.array_equal(t2, tops)
    # a re-layout drops the graph; the next launch captures again
    batch.set_tuning(8, 8192)
    atlas.upload(np.full(1024 * 1024, 77, dtype=np.uint8))
    batch.launch()
    batch.sync()
    assert np.array_equal(atlas.download(), want)
    batch.set_graph(False)


def test_batch_pipelined_launches_match_serial(ft, oracle):
    """Pipelined launches (two workspaces taking turns, one launch's front end
    overlapping the previous raster) give the bytes of the same launches run
    one at a time.  A translucent colour blends again at every launch, so a
    raster that ran out of order or beside the previous one, or a front end
    that overwrote buffers a raster still read, changes the result."""
    from footile_amd.plotter import Batch, DeviceBuffer, FillDesc
    W = H = 4096
    ops = workloads.random_walk(1)
    clr = (90, 50, 20, 96)
    A, B = DeviceBuffer(W * H * 4), DeviceBuffer(W * H * 4)
    both = np.concatenate([ops, ops])
    batch = Batch(both, [FillDesc(0, len(ops), A.ptr, W, H, W * 4, RGBA8P, rule=0, clr=clr),
                         FillDesc(len(ops), len(ops), B.ptr, W, H, W * 4, RGBA8P, rule=1, clr=clr)])
    batch.run()
    results = batch.results()

    def serial(n):
        A.zero()
        B.zero()
        for _ in range(n):
            batch.launch()
            batch.sync()
        return A.download(), B.download()

    want = {n: serial(n) for n in (1, 3, 4, 5)}
    assert not np.array_equal(want[4][0], want[5][0])  # every launch shows
    g = oracle.Plotter(W, H, RGBA8P)
    assert g.fill(0, ops, clr) == 0
    assert np.array_equal(want[1][0], g.pixels())

    def pipelined(n):
        A.zero()
        B.zero()
        for _ in range(n):
            batch.launch()
        batch.sync()
        return A.download(), B.download()

    for depth in (2, 3, 4):
        batch.set_pipeline(depth)
        for n in (5, 4, 1):
            got = pipelined(n)
            assert np.array_equal(got[0], want[n][0]) and np.array_equal(got[1], want[n][1]), (depth, n)
            r = batch.results()
            assert all(np.array_equal(x, y) for x, y in zip(r, results))
    with pytest.raises(ft.FootileError):
        batch.set_graph(True)
    with pytest.raises(ft.FootileError):
        batch.set_pipeline(5)
    # a re-layout reaches every workspace
    batch.set_tuning(8, 8192, True)
    got = pipelined(5)
    assert np.array_equal(got[0], want[5][0]) and np.array_equal(got[1], want[5][1])
    batch.set_pipeline(False)
    got = serial(3)
    assert np.array_equal(got[0], want[3][0]) and np.array_equal(got[1], want[3][1])

    # fresh atlas tiles, with the run that sizes the twin workspace in between
    batch, atlas, descs, ops, tiles = _atlas_batch()
    batch.set_pipeline(True)
    batch.run()
    want_atlas = atlas.download()
    for rep in range(3):
        atlas.upload(np.full(1024 * 1024, 77, dtype=np.uint8))
        for _ in range(3 + rep):
            batch.launch()
        batch.sync()
        assert np.array_equal(atlas.download(), want_atlas), f"atlas rep {rep}"
    for t, (b, c, x, y, s) in enumerate(tiles[:8]):
        o = oracle.Plotter(int(s), int(s))
        assert o.fill(0, ops[b:b + c]) == 0
        assert np.array_equal(want_atlas.reshape(1024, 1024)[y:y + s, x:x + s].reshape(-1), o.pixels()), t


def test_two_batches_launched_alternately(ft, oracle):
    """Two batches (each with its own workspace and stream) launched in turn
    without syncs, the second also through a Plotter between them: runs of
    different workspaces take turns, and every raster is what the batch gives
    alone."""
    from footile_amd.plotter import Batch, DeviceBuffer, FillDesc
    W = H = 1024
    ops = workloads.random_walk(1, n_cubics=2000, canvas=1024)
    clr = (90, 50, 20, 96)
    bufs = [DeviceBuffer(W * H * 4) for _ in range(2)]

    def make(buf, rule):
        return Batch(ops, [FillDesc(0, len(ops), buf.ptr, W, H, W * 4, RGBA8P, rule=rule, clr=clr)])

    b1, b2 = make(bufs[0], 0), make(bufs[1], 1)
    b1.run()
    b2.run()
    want = []
    for b, buf in ((b1, bufs[0]), (b2, bufs[1])):
        buf.zero()
        for _ in range(3):
            b.launch()
        b.sync()
        want.append(buf.download())
    for buf in bufs:
        buf.zero()
    p = ft.Plotter(ft.Raster.with_clear(640, 480))
    for k in range(3):
        b1.launch()
        p.fill(0, workloads.random_walk(2, n_cubics=300, canvas=480))
        b2.launch()
    b1.sync()
    b2.sync()
    assert np.array_equal(bufs[0].download(), want[0]) and np.array_equal(bufs[1].download(), want[1])


def test_look_back_kernels_finish_on_eight_compute_units(ft, oracle):
    """The scans and the one-pass flattening are sized for the whole device,
    and their blocks wait on each other (decoupled look-back).  Run on a
    stream that may use only 8 compute units, far fewer blocks than their
    grids are resident at once; blocks take look-back tickets on arrival, so
    none waits on a block that has not started and the run completes with
    the bytes of an ordinary run."""
    import ctypes
    from footile_amd._lib import FillDescC, lib
    from footile_amd.path import ops_array
    from footile_amd.plotter import Batch, DeviceBuffer, FillDesc
    hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so.7")
    stream = ctypes.c_void_p()
    mask = (ctypes.c_uint32 * 8)(0xFF, 0, 0, 0, 0, 0, 0, 0)
    assert hip.hipExtStreamCreateWithCUMask(ctypes.byref(stream), 8, mask) == 0
    try:
        W = H = 4096
        ops = workloads.random_walk(1)
        both = np.concatenate([ops, ops])

        def make():
            d = DeviceBuffer(2 * W * H * 4)
            d.zero()
            return d, [FillDesc(0, len(ops), d.ptr, W, H, W * 4, RGBA8P, rule=0, clr=(200, 120, 40, 255)),
                       FillDesc(len(ops), len(ops), d.ptr + W * H * 4, W, H, W * 4, RGBA8P, rule=1,
                                clr=(30, 60, 90, 128))]
        want_buf, descs = make()
        batch = Batch(both, descs)
        batch.run()
        want = want_buf.download()
        got_buf, descs = make()
        a = ops_array(both)
        arr = (FillDescC * 2)(*[d.c for d in descs])
        off = np.zeros(2, np.uint64)
        st = np.zeros(2, np.uint64)
        rc = lib().ft_batch_run_noalloc(a.ctypes.data_as(ctypes.c_void_p), len(a), arr, 2, stream,
                                        off.ctypes.data_as(ctypes.c_void_p), st.ctypes.data_as(ctypes.c_void_p))
        assert rc == 0, rc
        assert np.array_equal(got_buf.download(), want)
        w_off, w_st, _ = batch.counts()
        assert np.array_equal(off, np.asarray(w_off, np.uint64)) and np.array_equal(st, np.asarray(w_st, np.uint64))
    finally:
        hip.hipStreamDestroy(stream)


def test_batch_rmw_formats_with_and_without_prefetch(ft, oracle):
    """Rgba8p / Graya8p batches over existing content: the destination
    prefetch on and off, through the bands and through the tile scheduler,
    give the oracle's bytes."""
    from footile_amd.plotter import Batch, DeviceBuffer, FillDesc
    rng = np.random.default_rng(31)
    for fmt in (RGBA8P, GRAYA8P):
        bpp = BPP[fmt]
        side = 512
        init = rng.integers(0, 256, size=side * side * bpp).astype(np.uint8)
        dev = DeviceBuffer(side * side * bpp)
        ops_all, descs, tiles = [], [], []
        at = 0
        for t in range(12):  # disjoint tiles: fills of one batch run concurrently
            w, h = int(rng.integers(40, 168)), int(rng.integers(8, 126))
            x, y = (t % 3) * 170 + int(rng.integers(0, 170 - w)), (t // 3) * 128 + int(rng.integers(0, 128 - h))
            ops = random_path(rng, w, h, n_sub=2, kinds=(LINE, QUAD, CUBIC))
            clr = tuple(int(v) for v in rng.integers(0, 256, size=4))
            if fmt == RGBA8P:
                clr = (min(clr[0], clr[3]), min(clr[1], clr[3]), min(clr[2], clr[3]), clr[3])
            descs.append(FillDesc(at, len(ops), dev.ptr + (y * side + x) * bpp, w, h, side * bpp, fmt,
                                  rule=t % 2, clr=clr))
            tiles.append((x, y, w, h, ops, clr, t % 2))
            ops_all.append(ops)
            at += len(ops)
        batch = Batch(np.concatenate(ops_all), descs)
        for pf, tg in ((True, 0), (False, 0), (True, 4), (False, 4)):  # bands, then the tile scheduler
            dev.upload(init)
            batch.set_tuning(16, 4096, pf, tg)
            batch.run()
            got = dev.download().reshape(side, side * bpp)
            want = init.reshape(side, side * bpp).copy()
            for x, y, w, h, ops, clr, rule in tiles:
                o = oracle.Plotter(w, h, fmt)
                o.set_pixels(want[y:y + h, x * bpp:(x + w) * bpp].reshape(-1))
                assert o.fill(rule, ops, clr) == 0
                want[y:y + h, x * bpp:(x + w) * bpp] = o.pixels().reshape(h, w * bpp)
            assert np.array_equal(got, want), (fmt, pf, tg)


def test_batch_over_a_million_events(ft, oracle):
    """A batch past 2^20 events (its kept scan spans several tiles per block,
    so the kept flags take a pass of their own): sampled tiles and their
    point counts against the oracle."""
    from footile_amd.plotter import Batch, DeviceBuffer, FillDesc
    side = 4096
    ops, tiles = workloads.atlas(12, n_paths=1700, atlas_side=side)
    atlas = DeviceBuffer(side * side)
    atlas.upload(np.full(side * side, 77, dtype=np.uint8))
    descs = [FillDesc(int(b), int(c), atlas.ptr + int(y) * side + int(x), int(s), int(s), side,
                      fresh=True) for b, c, x, y, s in tiles]
    batch = Batch(ops, descs)
    batch.run()
    offered, stored, edges = batch.counts()
    assert int(offered.sum()) > (1 << 20)
    img = atlas.download().reshape(side, side)
    rng = np.random.default_rng(5)
    for t in sorted(rng.choice(len(tiles), size=40, replace=False)):
        b, c, x, y, s = tiles[t]
        o = oracle.Plotter(int(s), int(s))
        assert o.fill(0, ops[b:b + c]) == 0
        assert np.array_equal(img[y:y + s, x:x + s], o.pixels().reshape(s, s)), f"tile {t}"
        assert (offered[t], stored[t]) == o.last_counts()


def test_long_curves_fine_tolerance(ft, oracle):
    """Curves whose flattening lanes hold more than the eight points the count
    pass keeps per lane (the emit pass walks them again), mixed with short
    ones, at the finest tolerance (0.01) and the default."""
    rng = np.random.default_rng(21)
    ops = [(MOVE, (20.0, 20.0))]
    for k in range(12):
        pts = rng.uniform(0, 1000, size=6).astype(np.float32)
        ops.append((CUBIC if k % 3 else QUAD, tuple(pts if k % 3 else pts[:4])))
        ops.append((LINE, tuple(rng.uniform(0, 1000, size=2).astype(np.float32))))
    ops.append((CLOSE, ()))
    a = ops_array(ops)
    stored = {}
    for tol in (0.01, 0.3):
        g, o = compare_fill(ft, oracle, 1000, 1000, MATTE8, [("fill", 1, a, (255,))], tol=tol, tag=f"tol {tol}")
        assert g.last_counts() == o.last_counts()
        stored[tol] = o.last_counts()[1]
    assert stored[0.01] > 12 * 8 * 8  # far more leaves than 8 lanes x 8 kept points per curve


def test_degenerate_inputs(ft, oracle):
    """No ops, a lone move, a flat (horizontal-only) polygon, 1x1 and 1-wide
    rasters, figures wholly off the raster, and coordinates that saturate
    Fixed: every format and both rules, against the oracle."""
    cases = [
        ("no ops", 16, 16, []),
        ("lone move", 16, 16, [(MOVE, (3.0, 4.0)), (CLOSE, ())]),
        ("flat", 16, 16, [(MOVE, (1.0, 5.0)), (LINE, (12.0, 5.0)), (LINE, (7.0, 5.0)), (CLOSE, ())]),
        ("1x1", 1, 1, [(MOVE, (0.2, 0.1)), (LINE, (0.9, 0.6)), (LINE, (0.1, 0.95)), (CLOSE, ())]),
        ("1 wide", 1, 40, [(MOVE, (-3.0, -2.0)), (CUBIC, (5.0, 10.0, -4.0, 30.0, 2.0, 44.0)), (CLOSE, ())]),
        ("left of raster", 24, 24, [(MOVE, (-30.0, 2.0)), (LINE, (-5.0, 20.0)), (LINE, (-20.0, 22.0)), (CLOSE, ())]),
        ("right of raster", 24, 24, [(MOVE, (30.0, 2.0)), (LINE, (50.0, 20.0)), (LINE, (40.0, 22.0)), (CLOSE, ())]),
        ("below raster", 24, 24, [(MOVE, (2.0, 30.0)), (LINE, (20.0, 50.0)), (LINE, (4.0, 40.0)), (CLOSE, ())]),
        ("above raster", 24, 24, [(MOVE, (2.0, -30.0)), (LINE, (20.0, -5.0)), (LINE, (4.0, -12.0)), (CLOSE, ())]),
        ("saturating", 32, 32, [(MOVE, (-1.0e6, -1.0e6)), (LINE, (1.0e6, 3.0e4)), (LINE, (16.0, 1.0e7)),
                                (QUAD, (-5.0e5, 10.0, 8.0, 8.0)), (CLOSE, ())]),
    ]
    rng = np.random.default_rng(3)
    for name, w, h, ops in cases:
        a = ops_array(ops)
        for fmt in (MATTE8, RGBA8P, GRAYA8P):
            init = rng.integers(0, 256, size=w * h * BPP[fmt]).astype(np.uint8)
            clr = (90, 60, 30, 200) if fmt == RGBA8P else (200, 220, 0, 0)
            for rule in (0, 1):
                g, o = compare_fill(ft, oracle, w, h, fmt, [("fill", rule, a, clr)], init=init,
                                    tag=f"{name} fmt {fmt} rule {rule}")
                assert g.last_counts() == o.last_counts(), name


def _fig_cases():
    """Paths that exercise every Fig rule: the 65,535-point cap with and
    without its closing pop, duplicate and coincident points, several
    subfigures with closes, a close right after a move, a lone move, no ops,
    a flat figure, saturating coordinates, and plain random figures."""
    rng = np.random.default_rng(44)
    n = 66000
    pts = np.stack([rng.uniform(0, 250, n), rng.uniform(0, 250, n)], axis=1).astype(np.float32)
    pts[65534] = pts[0]
    base = [(MOVE, tuple(pts[0]))] + [(LINE, tuple(p)) for p in pts[1:]] + [(CLOSE, ())]
    tail = [(MOVE, (5.0, 5.0)), (LINE, (60.0, 9.0)), (LINE, (30.0, 70.0)), (CLOSE, ())]
    cases = [base, base + tail, base[:65536] + [(CLOSE, ())] + tail,
             base[:65535] + [(LINE, tuple(pts[0])), (CLOSE, ())] + tail,
             [], [(MOVE, (3.0, 4.0)), (CLOSE, ())],
             [(MOVE, (1.0, 5.0)), (LINE, (12.0, 5.0)), (LINE, (7.0, 5.0)), (CLOSE, ())],
             [(MOVE, (-1.0e6, -1.0e6)), (LINE, (1.0e6, 3.0e4)), (LINE, (16.0, 1.0e7)),
              (QUAD, (-5.0e5, 10.0, 8.0, 8.0)), (CLOSE, ())],
             [(MOVE, (10.0, 10.0)), (CLOSE, ()), (MOVE, (20.0, 20.0)), (LINE, (20.0, 20.0)), (LINE, (90.0, 30.0)),
              (LINE, (90.0, 30.0)), (LINE, (40.0, 100.0)), (LINE, (20.0, 20.0)), (CLOSE, ()), (CLOSE, ()),
              (MOVE, (150.0, 10.0)), (LINE, (200.0, 60.0)), (LINE, (120.0, 90.0))]]
    while len(cases) < 80:
        k = len(cases)
        ops = []
        for _ in range(int(rng.integers(1, 4))):  # subfigures, some left open
            ops.append((MOVE, tuple(rng.uniform(0, 256, 2).astype(np.float32))))
            for _ in range(int(rng.integers(1, 12))):
                r = rng.random()
                if r < 0.15 and ops[-1][0] == LINE:
                    ops.append(ops[-1])  # a duplicate point
                elif r < 0.5:
                    ops.append((LINE, tuple(rng.uniform(0, 256, 2).astype(np.float32))))
                elif r < 0.75:
                    ops.append((QUAD, tuple(rng.uniform(0, 256, 4).astype(np.float32))))
                else:
                    ops.append((CUBIC, tuple(rng.uniform(0, 256, 6).astype(np.float32))))
            if k % 3:
                ops.append((CLOSE, ()))
        cases.append(ops)
    return cases


def test_batch_of_many_paths_fig_rules(ft, oracle):
    """A batch of 80 paths holding the Fig's edge cases side by side (so the
    ingest scans run across path boundaries): every raster and every path's
    offered / stored counts against a Plotter of its own, both rules."""
    from footile_amd.plotter import Batch, DeviceBuffer, FillDesc
    cases = _fig_cases()
    arrays = [ops_array(c) for c in cases]
    ops = np.concatenate([a for a in arrays if len(a)])
    w = h = 256
    for rule in (0, 1):
        bufs, descs, at = [], [], 0
        for k, a in enumerate(arrays):
            b = DeviceBuffer(w * h)
            b.upload(np.full(w * h, k, dtype=np.uint8))
            bufs.append(b)
            descs.append(FillDesc(at, len(a), b.ptr, w, h, w, MATTE8, rule=rule, clr=(255,)))
            at += len(a)
        batch = Batch(ops, descs)
        batch.run()
        offered, stored, _ = batch.counts()
        for k, a in enumerate(arrays):
            o = oracle.Plotter(w, h)
            o.set_pixels(np.full(w * h, k, dtype=np.uint8))
            assert o.fill(rule, a, (255,)) == 0
            got = bufs[k].download()
            assert np.array_equal(got, o.pixels()), f"rule {rule} path {k}: {int((got != o.pixels()).sum())} bytes"
            assert (offered[k], stored[k]) == o.last_counts(), f"rule {rule} path {k}"


def _repeat_tail_path(rng, n, repeats, to_first, w=250.0):
    """One closed run of n random points whose last point repeats `repeats`
    times before the close; with to_first the repeated point is the run's
    first point, so the close pops it."""
    pts = rng.uniform(0, w, size=(n, 2)).astype(np.float32)
    last = tuple(pts[0]) if to_first else tuple(pts[-1])
    ops = [(MOVE, tuple(pts[0]))] + [(LINE, tuple(p)) for p in pts[1:]]
    ops += [(LINE, last)] * repeats + [(CLOSE, ())]
    return ops


def test_long_closing_walks_fall_back_to_the_ingest_pass(ft, oracle):
    """The kept flags find each run's last pushed point by walking back from
    its close over repeated points, at most 32 of them; a longer walk makes
    the run start again with the ingest pass and its running maximum.  Runs
    ending in 31, 32, 33, 40 and 50 repeats (of the last point, or of the
    first, which the close then pops), beside ordinary runs: through the
    Plotter (one scan pass for the kept flags) and in a batch of more than
    2^20 events (a kept pass of its own), every raster and count against the
    oracle."""
    from footile_amd.plotter import Batch, DeviceBuffer, FillDesc
    rng = np.random.default_rng(321)
    cases = []
    for repeats in (31, 32, 33, 40, 50):
        for to_first in (False, True):
            ops = _repeat_tail_path(rng, 2500, repeats, to_first)
            ops += _repeat_tail_path(rng, 60, 3, False)  # an ordinary run after it
            cases.append(ops_array(ops))
    for k, a in enumerate(cases):
        for rule in (0, 1):
            g, o = compare_fill(ft, oracle, 256, 256, MATTE8, [("fill", rule, a, (255,))], tag=f"walk {k} rule {rule}")
            assert g.last_counts() == o.last_counts(), k
    # a batch past 2^20 events
    paths = [ops_array(_repeat_tail_path(rng, 60000, (0, 40, 5, 33)[k % 4], k % 3 == 0, 64.0)) for k in range(18)]
    ops = np.concatenate(paths)
    w = h = 64
    bufs, descs, at = [], [], 0
    for k, a in enumerate(paths):
        b = DeviceBuffer(w * h)
        b.upload(np.full(w * h, k, dtype=np.uint8))
        bufs.append(b)
        descs.append(FillDesc(at, len(a), b.ptr, w, h, w, MATTE8, rule=k % 2, clr=(255,)))
        at += len(a)
    batch = Batch(ops, descs)
    batch.run()
    offered, stored, _ = batch.counts()
    for k, a in enumerate(paths):
        o = oracle.Plotter(w, h)
        o.set_pixels(np.full(w * h, k, dtype=np.uint8))
        assert o.fill(k % 2, a, (255,)) == 0
        got = bufs[k].download()
        assert np.array_equal(got, o.pixels()), f"path {k}: {int((got != o.pixels()).sum())} bytes"
        assert (offered[k], stored[k]) == o.last_counts(), f"path {k}"


def test_batch_band_pair_overflow_grows_and_reruns(ft, oracle):
    """Long near-vertical edges over one-row bands give far more (edge, band)
    pairs than the first run's list capacity: the run reports the overflow,
    the batch grows its lists and runs again, and the raster is the oracle's."""
    from footile_amd.plotter import Batch, DeviceBuffer, FillDesc
    w, h = 1024, 3000
    ops = []
    for k in range(6):
        x = 40.0 + 160.0 * k
        ops += [(MOVE, (x, 1.0)), (LINE, (x + 70.0, 2.0)), (LINE, (x + 90.0, 2999.0)), (LINE, (x + 5.0, 2990.0)),
                (CLOSE, ())]
    a = ops_array(ops)
    buf = DeviceBuffer(w * h)
    buf.upload(np.full(w * h, 9, dtype=np.uint8))
    batch = Batch(a, [FillDesc(0, len(a), buf.ptr, w, h, w, MATTE8, rule=0, fresh=True)])
    batch.set_tuning(1, 1024)
    batch.run()
    o = oracle.Plotter(w, h)
    assert o.fill(0, a) == 0
    assert np.array_equal(buf.download(), o.pixels())
    offered, stored, _ = batch.counts()
    assert (offered[0], stored[0]) == o.last_counts()


def test_pipelined_overflow_then_run_grows_every_workspace(ft, oracle):
    """Pipelined launches that outgrow the first capacities report FT_ENOMEM at
    the sync; ft_batch_run then grows the main workspace and its twins, and
    pipelined launches give the oracle's raster."""
    from footile_amd.plotter import Batch, DeviceBuffer, FillDesc
    w, h = 1024, 3000
    ops = []
    for k in range(6):
        x = 40.0 + 160.0 * k
        ops += [(MOVE, (x, 1.0)), (LINE, (x + 70.0, 2.0)), (LINE, (x + 90.0, 2999.0)), (LINE, (x + 5.0, 2990.0)),
                (CLOSE, ())]
    a = ops_array(ops)
    buf = DeviceBuffer(w * h)
    batch = Batch(a, [FillDesc(0, len(a), buf.ptr, w, h, w, MATTE8, rule=0, fresh=True)])
    batch.set_tuning(1, 1024)
    batch.set_pipeline(3)
    buf.upload(np.full(w * h, 9, dtype=np.uint8))
    for _ in range(3):
        batch.launch()
    with pytest.raises(ft.FootileError) as e:
        batch.sync()
    assert e.value.code == 2  # FT_ENOMEM: the band lists overflowed
    batch.run()
    o = oracle.Plotter(w, h)
    assert o.fill(0, a) == 0
    for n in (1, 4):
        buf.upload(np.full(w * h, 9, dtype=np.uint8))
        for _ in range(n):
            batch.launch()
        batch.sync()
        assert np.array_equal(buf.download(), o.pixels()), n
    offered, stored, _ = batch.counts()
    assert (offered[0], stored[0]) == o.last_counts()
    batch.set_pipeline(1)


def _oracle_tiles(oracle, ops, tiles, idx, threads=16):
    """The oracle's fresh s x s Matte8 tile and (offered, stored) counts for
    each tile index, on a thread pool (the oracle's calls release the GIL)."""
    from concurrent.futures import ThreadPoolExecutor

    def one(t):
        b, c, _, _, s = (int(v) for v in tiles[t])
        o = oracle.Plotter(s, s)
        assert o.fill(0, ops[b:b + c]) == 0
        return o.pixels().reshape(s, s), o.last_counts()

    with ThreadPoolExecutor(threads) as ex:
        return list(ex.map(one, idx))


def test_atlas_full_size_every_tile(ft, oracle):
    """The batched atlas at its full size: 16,384 paths (lognormal(3.5, 1)
    segments, seed 2) into fresh tiles of a 16384x16384 Matte8 atlas, with the
    autotuned raster geometry and the launch sequence replayed as a HIP graph.
    Every tile and its offered / stored point counts against the oracle, and
    the pixels between tiles untouched."""
    from footile_amd.plotter import Batch, DeviceBuffer, FillDesc
    from footile_amd.tune import autotune
    side = 16384
    ops, tiles = workloads.atlas(2, n_paths=16384, atlas_side=side)
    atlas = DeviceBuffer(side * side)
    atlas.upload(np.full(side * side, 77, dtype=np.uint8))
    descs = [FillDesc(int(b), int(c), atlas.ptr + int(y) * side + int(x), int(s), int(s), side,
                      fresh=True) for b, c, x, y, s in tiles]
    batch = Batch(ops, descs)
    autotune(batch, descs, len(ops), cache=None)
    atlas.upload(np.full(side * side, 77, dtype=np.uint8))
    batch.set_graph(True)
    for _ in range(2):  # capture, then a replay of the captured graph
        batch.launch()
        batch.sync()
    batch.set_graph(False)
    img = atlas.download().reshape(side, side)
    offered, stored, _ = batch.counts()
    covered = np.zeros((side, side), bool)
    for k0 in range(0, len(tiles), 2048):
        idx = list(range(k0, min(k0 + 2048, len(tiles))))
        for t, (want, counts) in zip(idx, _oracle_tiles(oracle, ops, tiles, idx)):
            b, c, x, y, s = (int(v) for v in tiles[t])
            if not np.array_equal(img[y:y + s, x:x + s], want):
                bad = np.argwhere(img[y:y + s, x:x + s] != want)
                raise AssertionError(f"tile {t} (side {s}): {len(bad)} pixels differ, first at {bad[:4].tolist()}")
            assert (int(offered[t]), int(stored[t])) == counts, f"tile {t}"
            covered[y:y + s, x:x + s] = True
    assert np.all(img[~covered] == 77)


def test_small_figure_path_and_its_limits(ft, oracle):
    """Plotter fills of up to 256 ops whose events fit 2,048 run as small
    figures (one block before the raster, one raster block per row group);
    past either limit (257 ops; fine curves with more events) the full
    pipeline runs, in the same plotter, between small fills.  Every case
    against the oracle, in every format, with rasters taller than one LDS row
    group."""
    rng = np.random.default_rng(77)
    cases = []
    for n_sub, kinds in ((20, (LINE,)), (40, (LINE, QUAD, CUBIC))):
        ops = random_path(rng, 200, 300, n_sub=n_sub, kinds=kinds)
        cases.append(ops[:256] if len(ops) > 256 else ops)
    cases.append(random_path(rng, 200, 300, n_sub=1, kinds=(QUAD, CUBIC)))  # curves within the event limit
    many = random_path(rng, 200, 300, n_sub=80, kinds=(LINE,))
    assert len(many) > 257
    cases.append(many[:257])
    fine = ops_array([(MOVE, (5.0, 5.0))] + [(CUBIC, (250.0 * (k % 2), 150.0, 5.0 + k, 290.0, 195.0, 5.0 + 2 * k))
                                            for k in range(40)] + [(CLOSE, ())])
    cases.append(fine)
    for fmt in (MATTE8, RGBA8P, GRAYA8P):
        clr = {MATTE8: (255,), RGBA8P: (90, 160, 30, 200), GRAYA8P: (70, 140)}[fmt]
        init = rng.integers(0, 256, size=200 * 300 * BPP[fmt]).astype(np.uint8) if fmt != MATTE8 else None
        calls = []
        for k, ops in enumerate(cases):
            calls.append(("fill", k % 2, ops, clr))
        compare_fill(ft, oracle, 200, 300, fmt, calls, init=init, tol=0.05, tag=f"small path fmt {fmt}")


def test_small_stroke_path_and_its_limits(ft, oracle):
    """Plotter strokes whose flattened events and outline fit 2,048 events
    (and 257 closes) run the stroker inside the small-figure block; bigger
    outlines (round joins on a long polyline, random curves) fall back to the
    full pipeline in the same plotter.  The oracle's outline point counts show
    which side of the limit each case is on; every stroke is compared byte
    for byte, with pen widths changed inside paths, degenerate sub-strokes and
    every join."""
    rng = np.random.default_rng(17)
    poly = ops_array([(PEN_WIDTH, (12.0,)), (MOVE, (100.0, 100.0))] +
                     [(LINE, tuple(rng.uniform(10, [190, 290]).astype(np.float32))) for _ in range(200)])
    widths = []
    for k in range(12):
        widths += [(PEN_WIDTH, (float(1 + 2 * k),)), (LINE, tuple(rng.uniform(10, [190, 290]).astype(np.float32)))]
    widths = ops_array([(MOVE, (20.0, 20.0))] + widths + [(CLOSE, ())])
    degenerate = ops_array([(MOVE, (5.0, 5.0)), (MOVE, (30.0, 40.0)), (LINE, (30.0, 40.0)), (LINE, (90.0, 40.0)),
                            (LINE, (90.0, 40.0)), (CLOSE, ()), (CLOSE, ()), (LINE, (150.0, 150.0)),
                            (MOVE, (60.0, 200.0)), (MOVE, (70.0, 210.0)), (LINE, (70.0, 210.0)), (CLOSE, ()),
                            (MOVE, (120.0, 250.0)), (QUAD, (180.0, 280.0, 120.0, 290.0)), (LINE, (120.0, 250.0))])
    lines = random_path(rng, 200, 300, n_sub=6, kinds=(LINE,), close_prob=0.5)
    many = random_path(rng, 200, 300, n_sub=60, kinds=(LINE,), close_prob=0.5)[:256]
    curves = random_path(rng, 200, 300, n_sub=5, kinds=(LINE, QUAD, CUBIC), close_prob=0.5)
    cases = [(poly, (0, 4.0), True), (poly, (1, 0.0), True), (poly, (2, 0.0), False), (widths, (0, 4.0), True),
             (widths, (2, 0.0), True), (degenerate, (0, 1.5), True), (lines, (0, 4.0), True),
             (many, (1, 0.0), True), (curves, (0, 4.0), False)]
    for fmt in (MATTE8, RGBA8P):
        clr = (255,) if fmt == MATTE8 else (30, 160, 90, 220)
        for k, (ops, join, small) in enumerate(cases):
            g, o = compare_fill(ft, oracle, 200, 300, fmt, [("stroke", ops, clr), ("fill", 1, ops, clr),
                                                           ("stroke", ops, clr)],
                                join=join, tag=f"small stroke {k} join {join} fmt {fmt}")
            assert g.last_counts() == o.last_counts()
            o2 = oracle_plotter(oracle, 200, 300, fmt, join=join)
            assert o2.stroke(ops, clr) == 0
            offered = o2.last_counts()[0]
            assert (offered < 1900) if small else (offered > 2100), (k, offered)
