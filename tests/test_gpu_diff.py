"""Differentiable fill on the GPU against the FP64 oracle (oracle/diff_oracle.c,
pinned by tests/test_diff_oracle.py).

Tolerances: the forward canvas is float32, within 2e-5 of FP64 per channel;
gradients are compared normwise, ||g - g64|| / ||g64|| <= 1e-4 (the bound the
differentiable path is specified to), per parameter group.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _scene(seed, n_images=2, w=40, h=32, paths=6, cubics=(1, 4)):
    from footile_amd.diffscene import DiffScene
    rng = np.random.default_rng(seed)
    image, cub = [], []
    for i in range(n_images):
        for _ in range(paths):
            image.append(i)
            cub.append(int(rng.integers(*cubics)))
    image = np.asarray(image, np.uint32)
    cub = np.asarray(cub, np.uint32)
    n_pts = int(np.sum(1 + 3 * cub.astype(np.int64)))
    pts = rng.uniform(-3.0, [w + 3.0, h + 3.0], size=(n_pts, 2)).astype(np.float32)
    a = rng.uniform(0.3, 0.9, size=len(image))
    col = np.stack([a * rng.uniform(0, 1, len(image)) for _ in range(3)] + [a], axis=1).astype(np.float32)
    return DiffScene(n_images, w, h, image, cub, pts, col)


def _oracle_scene(sc):
    from oracle import pydiff as pd
    return pd.Scene(sc.n_images, sc.width, sc.height, sc.path_image, sc.path_cubics)


def _gpu(sc, keep_coverage=False):
    from footile_amd.diff import DiffRenderer, device_array
    r = DiffRenderer.for_scene(sc, keep_coverage)
    return r, device_array(sc.points), device_array(sc.colours)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_forward_matches_fp64(oracle, seed):
    sc = _scene(seed)
    r, P, C = _gpu(sc)
    canvas = r.new_canvas()
    r.forward(P, C, canvas)
    got = canvas.download(dtype=np.float32).reshape(r.canvas_shape)
    _, want = _oracle_scene(sc).forward(sc.points, sc.colours)
    assert np.abs(got - want).max() < 2e-5


@pytest.mark.parametrize("seed", [4, 5])
def test_loss_gradient_matches_fp64(oracle, seed):
    from footile_amd.diff import device_array
    from footile_amd.plotter import DeviceBuffer
    sc = _scene(seed)
    r, P, C = _gpu(sc)
    target = np.random.default_rng(seed + 100).uniform(0, 1, size=r.canvas_shape).astype(np.float32)
    T = device_array(target)
    canvas, gcan = r.new_canvas(), r.new_canvas()
    gp, gc = DeviceBuffer(8 * r.n_points), DeviceBuffer(16 * r.n_paths)
    loss = r.step(P, C, T, canvas, gcan, gp, gc)
    want_loss, want_gp, want_gc = _oracle_scene(sc).backward(sc.points, sc.colours, target)
    assert abs(loss - want_loss) <= 1e-5 * abs(want_loss)
    got_gp = gp.download(dtype=np.float32).reshape(-1, 2)
    got_gc = gc.download(dtype=np.float32).reshape(-1, 4)
    assert _rel(got_gp, want_gp) < 1e-4, _rel(got_gp, want_gp)
    assert _rel(got_gc, want_gc) < 1e-4, _rel(got_gc, want_gc)


def test_vjp_matches_fp64(oracle):
    from footile_amd.diff import DiffFill, device_array
    sc = _scene(9, n_images=1, w=48, h=40, paths=8)
    r, P, C = _gpu(sc)
    G = np.random.default_rng(3).normal(size=r.canvas_shape).astype(np.float32)
    canvas, ctx = DiffFill.apply(r, P, C)
    _, gp, gc = DiffFill.backward(ctx, device_array(G))
    _, want_gp, want_gc = _oracle_scene(sc).backward(sc.points, sc.colours, grad_canvas=G)
    assert _rel(gp.download(dtype=np.float32).reshape(-1, 2), want_gp) < 1e-4
    assert _rel(gc.download(dtype=np.float32).reshape(-1, 4), want_gc) < 1e-4


def test_c4_shape_reduced(oracle):
    """The one-image scene law (8-cubic paths across the canvas) at 128x128, 32 paths."""
    from footile_amd import diffscene as ds
    from footile_amd.diff import device_array
    from footile_amd.plotter import DeviceBuffer
    sc = ds.scene_one_image(seed=3, width=128, height=128, n_paths=32)
    tp, tc = ds.perturbed(sc)
    _, target = _oracle_scene(sc).forward(tp, tc)
    target = target.astype(np.float32)
    r, P, C = _gpu(sc)
    T = device_array(target)
    canvas, gcan = r.new_canvas(), r.new_canvas()
    gp, gc = DeviceBuffer(8 * r.n_points), DeviceBuffer(16 * r.n_paths)
    r.step(P, C, T, canvas, gcan, gp, gc)
    _, want_gp, want_gc = _oracle_scene(sc).backward(sc.points, sc.colours, target)
    assert _rel(gp.download(dtype=np.float32).reshape(-1, 2), want_gp) < 1e-4
    assert _rel(gc.download(dtype=np.float32).reshape(-1, 4), want_gc) < 1e-4


def test_step_equals_forward_loss_backward(oracle):
    from footile_amd.diff import device_array
    from footile_amd.plotter import DeviceBuffer
    sc = _scene(12)
    r, P, C = _gpu(sc)
    T = device_array(np.full(r.canvas_shape, 0.25, np.float32))
    c1, g1 = r.new_canvas(), r.new_canvas()
    gp1, gc1 = DeviceBuffer(8 * r.n_points), DeviceBuffer(16 * r.n_paths)
    l1 = r.step(P, C, T, c1, g1, gp1, gc1)
    c2, g2 = r.new_canvas(), r.new_canvas()
    r.forward(P, C, c2)
    l2 = r.loss(c2, T, g2)
    gp2, gc2 = DeviceBuffer(8 * r.n_points), DeviceBuffer(16 * r.n_paths)
    r.backward(P, C, g2, gp2, gc2)
    assert l1 == l2
    assert np.array_equal(c1.download(), c2.download())
    a, b = gp1.download(dtype=np.float32), gp2.download(dtype=np.float32)
    assert np.allclose(a, b, rtol=1e-5, atol=1e-7 * np.abs(b).max())


def test_kept_coverage_matches_recompute(oracle):
    """The opt-in kept-coverage handle (FT_DIFF_KEEP_COVERAGE) and the default
    one, which recomputes coverage in the backward, give the same canvas and
    loss and gradients within float32 rounding, through a step and through a
    separate backward."""
    from footile_amd.diff import device_array
    from footile_amd.plotter import DeviceBuffer
    sc = _scene(12)
    out = []
    for keep in (False, True):
        r, P, C = _gpu(sc, keep)
        T = device_array(np.full(r.canvas_shape, 0.25, np.float32))
        c, g = r.new_canvas(), r.new_canvas()
        gp, gc = DeviceBuffer(8 * r.n_points), DeviceBuffer(16 * r.n_paths)
        loss = r.step(P, C, T, c, g, gp, gc)
        gp2, gc2 = DeviceBuffer(8 * r.n_points), DeviceBuffer(16 * r.n_paths)
        r.backward(P, C, g, gp2, gc2)  # replays the forward
        out.append((loss, c.download(), gp.download(dtype=np.float32), gc.download(dtype=np.float32),
                    gp2.download(dtype=np.float32), gc2.download(dtype=np.float32)))
    (l1, c1, *g1), (l2, c2, *g2) = out
    assert l1 == l2 and np.array_equal(c1, c2)
    for x in g1 + g2[2:]:
        y = g2[0] if x.size == g2[0].size else g2[1]
        assert np.allclose(x, y, rtol=1e-5, atol=1e-7 * np.abs(y).max())


@pytest.mark.parametrize("keep", [False, True])
def test_many_paths_long_cubics_vs_fp64(oracle, keep):
    """Layouts past the backward's fast cases against FP64: 21 paths per image
    (three checkpoint segments of 8, the last one partial) and paths of up to
    16 cubics, whose 257 vertices are past the 256 a block gathers in LDS (so
    their gradients go straight to global memory).  Recomputed and kept
    coverage, through ft_diff_step and through separate forward, loss and
    backward calls."""
    from footile_amd.diff import device_array
    from footile_amd.plotter import DeviceBuffer
    sc = _scene(21, n_images=2, w=64, h=48, paths=21, cubics=(12, 17))
    sc.path_cubics[::5] = 16  # several paths of exactly 257 vertices
    n_pts = int(np.sum(1 + 3 * sc.path_cubics.astype(np.int64)))
    rng = np.random.default_rng(22)
    sc.points = rng.uniform(-3.0, [64 + 3.0, 48 + 3.0], size=(n_pts, 2)).astype(np.float32)
    assert sc.n_paths == 42 and int(sc.path_cubics.max()) == 16
    target = rng.uniform(0, 1, size=(2, 48, 64, 4)).astype(np.float32)
    o = _oracle_scene(sc)
    want_loss, want_gp, want_gc = o.backward(sc.points, sc.colours, target)
    _, want_canvas = o.forward(sc.points, sc.colours)
    r, P, C = _gpu(sc, keep)
    T = device_array(target)
    canvas, gcan = r.new_canvas(), r.new_canvas()
    gp, gc = DeviceBuffer(8 * r.n_points), DeviceBuffer(16 * r.n_paths)
    loss = r.step(P, C, T, canvas, gcan, gp, gc)
    assert abs(loss - want_loss) <= 1e-5 * abs(want_loss)
    assert np.abs(canvas.download(dtype=np.float32).reshape(r.canvas_shape) - want_canvas).max() < 2e-5
    assert _rel(gp.download(dtype=np.float32).reshape(-1, 2), want_gp) < 1e-4
    assert _rel(gc.download(dtype=np.float32).reshape(-1, 4), want_gc) < 1e-4
    # the same through the separate entry points
    c2, g2 = r.new_canvas(), r.new_canvas()
    r.forward(P, C, c2)
    assert r.loss(c2, T, g2) == loss
    gp2, gc2 = DeviceBuffer(8 * r.n_points), DeviceBuffer(16 * r.n_paths)
    r.backward(P, C, g2, gp2, gc2)
    assert _rel(gp2.download(dtype=np.float32).reshape(-1, 2), want_gp) < 1e-4
    assert _rel(gc2.download(dtype=np.float32).reshape(-1, 4), want_gc) < 1e-4


def _elementwise(got, want):
    """Largest and mean |got - want| / max|want| over the elements."""
    d = np.abs(got - want) / max(float(np.abs(want).max()), 1e-30)
    return float(d.max()), float(d.mean())


def test_c4_full_size_vs_fp64(oracle):
    """The differentiable single-image scene at its full size: 512x512, 256
    paths of 8 cubics (seed 3), target rendered from perturbed parameters;
    canvas, loss and all gradients of one step against FP64, with coverage
    recomputed in the backward."""
    from footile_amd import diffscene as ds
    from footile_amd.diff import device_array
    from footile_amd.plotter import DeviceBuffer
    sc = ds.scene_one_image()
    r, P, C = _gpu(sc)
    tp, tc = ds.perturbed(sc)
    T = r.new_canvas()
    r.forward(device_array(tp), device_array(tc), T)
    target = T.download(dtype=np.float32).reshape(r.canvas_shape)
    canvas, gcan = r.new_canvas(), r.new_canvas()
    gp, gc = DeviceBuffer(8 * r.n_points), DeviceBuffer(16 * r.n_paths)
    loss = r.step(P, C, T, canvas, gcan, gp, gc)
    o = _oracle_scene(sc)
    want_loss, want_gp, want_gc = o.backward(sc.points, sc.colours, target)
    _, want_canvas = o.forward(sc.points, sc.colours)
    # 256 float32 composites deep: within 1e-4 per channel (measured 3.3e-5)
    assert np.abs(canvas.download(dtype=np.float32).reshape(r.canvas_shape) - want_canvas).max() < 1e-4
    assert abs(loss - want_loss) <= 1e-5 * abs(want_loss)
    got_gp = gp.download(dtype=np.float32).reshape(-1, 2)
    got_gc = gc.download(dtype=np.float32).reshape(-1, 4)
    assert _rel(got_gp, want_gp) < 1e-4 and _rel(got_gc, want_gc) < 1e-4
    assert _elementwise(got_gp, want_gp)[0] < 1e-4 and _elementwise(got_gc, want_gc)[0] < 1e-4


def test_c5_full_layout_vs_fp64(oracle):
    """The batched differentiable scene at its full size: 64 images of 256x256
    with U[32, 512] paths of U[4, 16] cubics (seed 4), one ft_diff_step with
    coverage recomputed in the backward.  The canvas of every image against
    FP64, and the gradients of 8 images (each image's parameters only touch
    its own pixels; the batched loss is the mean over all images)."""
    from concurrent.futures import ThreadPoolExecutor
    from footile_amd import diffscene as ds
    from footile_amd.diff import device_array
    from footile_amd.plotter import DeviceBuffer
    from oracle import pydiff as pd
    sc = ds.scene_many_images()
    r, P, C = _gpu(sc)
    tp, tc = ds.perturbed(sc)
    T = r.new_canvas()
    r.forward(device_array(tp), device_array(tc), T)
    target = T.download(dtype=np.float32).reshape(r.canvas_shape)
    canvas, gcan = r.new_canvas(), r.new_canvas()
    gp, gc = DeviceBuffer(8 * r.n_points), DeviceBuffer(16 * r.n_paths)
    r.step(P, C, T, canvas, gcan, gp, gc)
    got_canvas = canvas.download(dtype=np.float32).reshape(r.canvas_shape)
    got_gp = gp.download(dtype=np.float32).reshape(-1, 2)
    got_gc = gc.download(dtype=np.float32).reshape(-1, 4)
    first = np.searchsorted(sc.path_image, np.arange(sc.n_images + 1))
    pt_first = np.concatenate([[0], np.cumsum(1 + 3 * sc.path_cubics.astype(np.int64))])

    def image(i):
        k0, k1 = int(first[i]), int(first[i + 1])
        o = pd.Scene(1, sc.width, sc.height, np.zeros(k1 - k0, np.uint32), sc.path_cubics[k0:k1])
        pts, col = sc.points[pt_first[k0]:pt_first[k1]], sc.colours[k0:k1]
        _, want_canvas = o.forward(pts, col)
        err = float(np.abs(got_canvas[i] - want_canvas[0]).max())
        if i % 8:
            return err, None
        _, want_gp, want_gc = o.backward(pts, col, target[i:i + 1])
        scale = 1.0 / sc.n_images
        return err, (_rel(got_gp[pt_first[k0]:pt_first[k1]], want_gp * scale),
                     _rel(got_gc[k0:k1], want_gc * scale))

    with ThreadPoolExecutor(16) as ex:
        res = list(ex.map(image, range(sc.n_images)))
    assert max(e for e, _ in res) < 1e-4  # up to 512 float32 composites deep
    grads = [g for _, g in res if g is not None]
    assert len(grads) == 8
    assert max(g[0] for g in grads) < 1e-4 and max(g[1] for g in grads) < 1e-4


def test_empty_image_and_bad_layout(oracle):
    from footile_amd._lib import FootileError
    from footile_amd.diff import DiffRenderer, device_array
    # image 1 has no paths: its canvas stays transparent
    r = DiffRenderer(2, 16, 8, np.array([0], np.uint32), np.array([2], np.uint32))
    pts = np.array([[1, 1], [10, 1], [12, 6], [3, 7], [2, 3], [9, 2], [1, 1]], np.float32)
    col = np.array([[0.2, 0.1, 0.3, 0.5]], np.float32)
    canvas = r.new_canvas()
    r.forward(device_array(pts), device_array(col), canvas)
    got = canvas.download(dtype=np.float32).reshape(r.canvas_shape)
    assert np.all(got[1] == 0.0) and got[0].max() > 0
    with pytest.raises(FootileError):
        DiffRenderer(1, 16, 8, np.array([1], np.uint32), np.array([2], np.uint32))  # image out of range
    with pytest.raises(FootileError):
        DiffRenderer(1, 16, 8, np.array([0], np.uint32), np.array([0], np.uint32))  # no cubics
