#!/usr/bin/env python3
"""Benchmark of the footile fill path on one MI355X.

Measured workload (one step = one launch of the device pipeline over the
device-resident ops of two fills):
    random walk: one self-intersecting path of 10,000 cubic_to (SplitMix64
    seed 1, N(0, 40 px) steps clamped to the canvas), filled NonZero into a
    4096x4096 Rgba8p raster A and EvenOdd into a 4096x4096 Rgba8p raster B,
    colour Rgba8p(200, 120, 40, 255).
value = filled Mpixels/s = sum over the two fills of W * (H - max(top_row, 0))
per step / wall time per step, with the launches pipelined (the batch's two
workspaces take turns, so that a launch's flattening-to-binning front end
overlaps the previous launch's raster; the rasters keep launch order).  The
first step's rasters are compared byte for byte with the CPU oracle before
timing, and three pipelined launches with three launched one at a time
("parity": true when both match).  The same steps launched one at a time are
reported as extra.ms_per_step_one_at_a_time.

Also reported under "extra": the batched atlas (16,384 tiles in a 16384^2
Matte8 atlas, one launch) and the plain 256x256 fill/stroke latency through
the Plotter API.

    python bench.py [--gpus 1] [--steps 100] [--warmup 3]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

TIMING_EVERY = 4
# Workspaces taking turns in pipelined launches: each has its own front-end
# stream beside the batch's raster stream, and three plus that one are the
# four hardware queues HIP gives a process by default (depth 4 measured slower).
PIPE_DEPTH = 3
ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E peak
POINTS_PER_OP = {0: 0, 1: 1, 2: 1, 3: 2, 4: 3, 5: 0}


def op_bytes(ops):
    """Algorithmic op-input bytes: 8 B (x, y f32) per control or end point."""
    return 8 * int(sum(POINTS_PER_OP[int(t)] for t in ops["tag"]))


def cpu_random_walk(ops, budget_s):
    """The oracle (C restatement of the reference, 1 core) on the same fills:
    Mpix/s over whole steps (NonZero + EvenOdd fill of fresh rasters)."""
    from oracle import pyoracle as po
    t0 = time.perf_counter()
    outs = None
    times = []
    while True:
        ts = time.perf_counter()
        a = po.Plotter(4096, 4096, po.RGBA8P)  # raster allocation inside the sample, as fishyb.rs:35
        b = po.Plotter(4096, 4096, po.RGBA8P)
        assert a.fill(po.NON_ZERO, ops, (200, 120, 40, 255)) == 0
        assert b.fill(po.EVEN_ODD, ops, (200, 120, 40, 255)) == 0
        times.append(time.perf_counter() - ts)
        if outs is None:
            outs = (a.pixels(), b.pixels(), a.last_counts())
        if time.perf_counter() - t0 >= budget_s:
            break
    return outs, np.asarray(times)


def cpu_atlas(ops, tiles, budget_s, threads):
    """The oracle on a fixed sample of atlas tiles, one Plotter (fresh s x s
    Matte8) per tile: paths/s on one thread, and on `threads` threads with a
    static partition of the sample (the oracle's calls release the GIL)."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle import pyoracle as po
    sample = np.random.default_rng(1).permutation(len(tiles))

    def run(part, deadline):
        done = 0
        for t in part:
            b, c, _, _, s = (int(v) for v in tiles[t])
            o = po.Plotter(s, s)
            assert o.fill(po.NON_ZERO, ops[b:b + c]) == 0
            done += 1
            if time.perf_counter() >= deadline:
                break
        return done

    t0 = time.perf_counter()
    one = run(sample, t0 + budget_s)
    dt1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        parts = [sample[k::threads] for k in range(threads)]
        many = sum(ex.map(lambda p: run(p, t0 + budget_s), parts))
    dtn = time.perf_counter() - t0
    return {"paths_per_s_1_core": one / dt1, "paths_per_s_all_cores": many / dtn, "cores": threads,
            "kind": "port",
            "sample": f"random tiles of the same atlas: {one} in {dt1:.1f} s on 1 thread, {many} in {dtn:.1f} s "
                      f"on {threads} threads (static partition)"}


def bench_atlas(ft, steps, warmup, tuning=None, cpu_seconds=4.0):
    from footile_amd import workloads
    from footile_amd.plotter import Batch, DeviceBuffer, FillDesc
    side = 16384
    ops, tiles = workloads.atlas(2, n_paths=16384, atlas_side=side)
    atlas = DeviceBuffer(side * side)
    atlas.zero()
    descs = [FillDesc(int(b), int(c), atlas.ptr + int(y) * side + int(x), int(s), int(s), side,
                      fresh=True) for b, c, x, y, s in tiles]
    batch = Batch(ops, descs)
    if tuning:
        v = [int(x) for x in tuning.split(",")]
        b_rows, cells, tg = v[0], v[1], (v[2] if len(v) > 2 else 0)
        batch.set_tuning(b_rows, cells, True, tg)
        tuned = {"band_rows_min": b_rows, "lds_cells": cells, "tile_groups": tg, "tried": 0}
    else:
        from footile_amd.tune import autotune
        tuned = autotune(batch, descs, len(ops), cache=None)
    batch.run()
    # raster-kernel time from eager launches (kernels are not timed inside a graph)
    batch.set_timing(True)
    batch.timing()
    for _ in range(max(2, steps // 2)):
        batch.launch()
    batch.sync()
    eager_total_ms, raster_ms, runs = batch.timing()
    # one replay of the captured launch graph per launch
    batch.set_graph(True)
    for _ in range(warmup + 1):
        batch.launch()
        batch.sync()
    batch.timing()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.launch()
    batch.sync()
    dt_graph = (time.perf_counter() - t0) / steps
    total_ms, _, graph_runs = batch.timing()
    batch.set_graph(False)
    # the step: pipelined eager launches (later launches' front ends overlap
    # an earlier launch's raster)
    batch.set_pipeline(PIPE_DEPTH)
    for _ in range(warmup + 1):
        batch.launch()
    batch.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.launch()
    batch.sync()
    dt = (time.perf_counter() - t0) / steps
    batch.set_pipeline(1)
    tile_px = int((tiles[:, 4] ** 2).sum())
    bytes_per = tile_px + op_bytes(ops)
    offered, stored, edges = batch.counts()
    pairs, cell_adds = batch.raster_stats()  # one more (eager) run that counts the raster's work
    # spot-check tiles against the oracle (fresh s x s Matte8 per tile)
    from oracle import pyoracle as po
    img = atlas.download().reshape(side, side)
    rng = np.random.default_rng(0)
    checked = 0
    for t in rng.choice(len(tiles), size=64, replace=False):
        b, c, x, y, s = (int(v) for v in tiles[t])
        o = po.Plotter(s, s)
        assert o.fill(po.NON_ZERO, ops[b:b + c]) == 0
        if not np.array_equal(img[y:y + s, x:x + s].reshape(-1), o.pixels()):
            raise SystemExit(f"atlas tile {t} differs from the oracle")
        checked += 1
    raster_avg = raster_ms / max(runs, 1)
    return {
        "workload": "atlas: 16,384 paths, lognormal(3.5,1) segments, 60/30/10 cubic/quad/line, "
                    "tiles 32..256 in a 16384^2 Matte8 atlas, one launch, autotuned raster geometry",
        "paths_per_s": len(tiles) / dt,
        "mpix_per_s": tile_px / dt / 1e6,
        "ms_per_step": dt * 1e3,
        "launch": f"pipelined: {PIPE_DEPTH} workspaces take turns",
        "graph_ms_per_step": dt_graph * 1e3,
        "graph_device_ms_per_step": total_ms / max(graph_runs, 1),
        "eager_device_ms_per_step": eager_total_ms / max(runs, 1),
        "tuning": {"band_rows_min": tuned["band_rows_min"], "lds_cells": tuned["lds_cells"],
                   "tile_groups": tuned.get("tile_groups", 0), "candidates": tuned["tried"],
                   "tune_s": round(tuned.get("tune_s", 0.0), 3)},
        "raster_kernel_ms": raster_avg,
        "raster_kernel_gbs": bytes_per / (raster_avg * 1e-3) / 1e9 if raster_avg > 0 else None,
        "edges_per_pixel": float(edges.sum()) / tile_px,
        "edge_row_pairs_per_pixel": pairs / tile_px,
        "lds_atomics_per_pixel": cell_adds / tile_px,
        "lds_atomics_per_s": cell_adds / (raster_avg * 1e-3) if raster_avg > 0 else None,
        "points_stored": int(stored.sum()),
        "tiles_checked_vs_oracle": checked,
        "cpu_baseline": cpu_atlas(ops, tiles, cpu_seconds, threads=min(16, len(os.sched_getaffinity(0)))),
        "host_affinity_cores": len(os.sched_getaffinity(0)),
    }


def _median_mad(samples):
    a = np.asarray(samples)
    med = float(np.median(a))
    return med, float(np.median(np.abs(a - med)))


def bench_plain(ft, reps=50):
    """The plain 256x256 fill and stroke through the Plotter API, per call: median and MAD of `reps` calls
    (GPU: host ops, upload, launch, one host wait; CPU: the oracle on one
    core, raster allocation inside each sample as fishyb.rs:35 does)."""
    from footile_amd import workloads
    from oracle import pyoracle as po
    fill_ops = workloads.plain_256(0)
    stroke_ops = workloads.plain_256_stroke(0)
    out = {}
    for name, call in [("fill", lambda p: p.fill(0, fill_ops)), ("stroke", lambda p: p.stroke(stroke_ops))]:
        p = ft.Plotter(ft.Raster.with_clear(256, 256))
        call(p)
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            call(p)
            ts.append(time.perf_counter() - t0)
        med, mad = _median_mad(ts)
        out[f"gpu_{name}_us"], out[f"gpu_{name}_mad_us"] = med * 1e6, mad * 1e6
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            o = po.Plotter(256, 256)
            if name == "fill":
                o.fill(0, fill_ops)
            else:
                o.stroke(stroke_ops)
            ts.append(time.perf_counter() - t0)
        med, mad = _median_mad(ts)
        out[f"cpu_{name}_us"], out[f"cpu_{name}_mad_us"] = med * 1e6, mad * 1e6
    out["samples"] = reps
    out["workload"] = "plain 256x256 Matte8, 48 line + 16 cubic (seed 0); fill NonZero, stroke w=4 " \
                      "(Plotter API: host ops, upload + launch + sync per call); median and MAD per call"
    return out


def bench_plotter_random_walk(ft, ops, reps=10):
    """The random-walk NonZero fill through the drop-in entry (ft_plotter_fill:
    host ops uploaded per call, one host wait): ms per fill."""
    from footile_amd.pixel import RGBA8P
    p = ft.Plotter(ft.Raster.with_clear(4096, 4096, RGBA8P))
    p.fill(0, ops, (200, 120, 40, 255))
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        p.fill(0, ops, (200, 120, 40, 255))
        ts.append(time.perf_counter() - t0)
    med, mad = _median_mad(ts)
    return {"ms_per_fill": med * 1e3, "mad_ms": mad * 1e3, "mpix_per_s": 4096 * 4096 / med / 1e6,
            "workload": "random walk 4096x4096 Rgba8p NonZero through ft_plotter_fill (280 KB of ops "
                        "uploaded per call)", "samples": reps}


def _grad_rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _diff_buffers(r, scene, target_params=None):
    """Device parameters, a target rendered on the GPU from perturbed
    parameters, and output buffers for one differentiable scene."""
    from footile_amd.diff import device_array
    from footile_amd.diffscene import perturbed
    from footile_amd.plotter import DeviceBuffer
    P, C = device_array(scene.points), device_array(scene.colours)
    tp, tc = perturbed(scene) if target_params is None else target_params
    T = r.new_canvas()
    r.forward(device_array(tp), device_array(tc), T)
    bufs = dict(P=P, C=C, T=T, canvas=r.new_canvas(), gcan=r.new_canvas(), gp=DeviceBuffer(8 * r.n_points),
                gc=DeviceBuffer(16 * r.n_paths))
    return bufs


def _diff_time(r, b, steps, warmup):
    for _ in range(warmup):
        r.step(b["P"], b["C"], b["T"], b["canvas"], b["gcan"], b["gp"], b["gc"])
    t0 = time.perf_counter()
    for _ in range(steps):
        loss = r.step(b["P"], b["C"], b["T"], b["canvas"], b["gcan"], b["gp"], b["gc"])
    return (time.perf_counter() - t0) / steps, loss


def _grad_errors(got, want):
    """Normwise ||g - g64|| / ||g64||, and the largest and mean elementwise
    |g - g64| relative to max |g64| (an element's own |g64| can be ~0)."""
    d = np.abs(got - want) / max(float(np.abs(want).max()), 1e-30)
    return {"normwise": _grad_rel(got, want), "max_elementwise": float(d.max()),
            "mean_elementwise": float(d.mean())}


def bench_diff(steps, warmup, check=True):
    """Differentiable workloads: one 512^2 image (256 paths x 8 cubics) and
    64 images of 256^2 (U[32,512] paths x U[4,16] cubics), fwd + loss +
    bwd per step (one ft_diff_step call, coverage recomputed in the backward);
    gradients against the FP64 oracle.  The opt-in kept-coverage handle is
    timed beside each as its own row."""
    from footile_amd import diffscene as ds
    from footile_amd.diff import DiffRenderer
    out = {}
    # one 512x512 image
    sc = ds.scene_one_image()
    r = DiffRenderer.for_scene(sc)
    b = _diff_buffers(r, sc)
    dt, loss = _diff_time(r, b, steps, warmup)
    one = {"workload": "512x512 fp32 RGBA, 256 closed paths x 8 cubics (seed 3), L2 to a target rendered "
                      "from params + N(0,2px) / colours + N(0,0.05); one ft_diff_step, coverage recomputed in "
                      "the backward",
          "grad_steps_per_s": 1.0 / dt, "ms_per_step": dt * 1e3, "loss": loss}
    rk = DiffRenderer.for_scene(sc, keep_coverage=True)
    bk = _diff_buffers(rk, sc)
    dtk, _ = _diff_time(rk, bk, steps, warmup)
    one["kept_coverage_variant"] = {"grad_steps_per_s": 1.0 / dtk, "ms_per_step": dtk * 1e3,
                                   "extra_bytes": 4 * sc.n_paths * sc.width * sc.height}
    del rk, bk
    if check:
        from oracle import pydiff as pd
        target = b["T"].download(dtype=np.float32).reshape(r.canvas_shape)
        t0 = time.perf_counter()
        want_loss, want_gp, want_gc = pd.Scene(sc.n_images, sc.width, sc.height, sc.path_image,
                                               sc.path_cubics).backward(sc.points, sc.colours, target)
        one["fp64_oracle_s"] = time.perf_counter() - t0
        gp = b["gp"].download(dtype=np.float32).reshape(-1, 2)
        gc = b["gc"].download(dtype=np.float32).reshape(-1, 4)
        one["grad_err_points"] = _grad_errors(gp, want_gp)
        one["grad_err_colours"] = _grad_errors(gc, want_gc)
        one["loss_rel_err"] = abs(loss - want_loss) / abs(want_loss)
    out["image_512"] = one
    # 64 images of 256x256: the batched step vs one step per image
    sc = ds.scene_many_images()
    r = DiffRenderer.for_scene(sc)
    b = _diff_buffers(r, sc)
    dt, loss = _diff_time(r, b, max(2, steps // 2), 1)
    many = {"workload": "64 images 256x256 fp32 RGBA, U[32,512] paths x U[4,16] cubics (seed 4); all images "
                      "in one ft_diff_step (one host wait), coverage recomputed in the backward",
          "grad_steps_per_s": 1.0 / dt, "ms_per_step": dt * 1e3, "paths": int(sc.n_paths)}
    rk = DiffRenderer.for_scene(sc, keep_coverage=True)
    bk = _diff_buffers(rk, sc)
    dtk, _ = _diff_time(rk, bk, max(2, steps // 2), 1)
    many["kept_coverage_variant"] = {"grad_steps_per_s": 1.0 / dtk, "ms_per_step": dtk * 1e3}
    del rk, bk
    # per-image launches of the same scenes
    per_image = 0.0
    tp, tc = ds.perturbed(sc)
    first = np.searchsorted(sc.path_image, np.arange(sc.n_images + 1))
    pt_first = np.concatenate([[0], np.cumsum(1 + 3 * sc.path_cubics.astype(np.int64))])
    subs = []
    for i in range(sc.n_images):
        k0, k1 = int(first[i]), int(first[i + 1])
        sub = ds.DiffScene(1, sc.width, sc.height, np.zeros(k1 - k0, np.uint32), sc.path_cubics[k0:k1],
                           sc.points[pt_first[k0]:pt_first[k1]], sc.colours[k0:k1])
        rs = DiffRenderer.for_scene(sub)
        bs = _diff_buffers(rs, sub, (tp[pt_first[k0]:pt_first[k1]], tc[k0:k1]))
        subs.append((rs, bs, sub, k0, k1))
    for rs, bs, *_ in subs:
        rs.step(bs["P"], bs["C"], bs["T"], bs["canvas"], bs["gcan"], bs["gp"], bs["gc"])
    reps = max(1, steps // 4)
    t0 = time.perf_counter()
    for _ in range(reps):
        for rs, bs, *_ in subs:
            rs.step(bs["P"], bs["C"], bs["T"], bs["canvas"], bs["gcan"], bs["gp"], bs["gc"])
    per_image = (time.perf_counter() - t0) / reps
    many["per_image_step_ms"] = per_image * 1e3
    many["speedup_vs_per_image_step"] = per_image / dt
    if check:
        from concurrent.futures import ThreadPoolExecutor
        from oracle import pydiff as pd
        gp_all = b["gp"].download(dtype=np.float32).reshape(-1, 2)
        gc_all = b["gc"].download(dtype=np.float32).reshape(-1, 4)
        target = b["T"].download(dtype=np.float32).reshape(r.canvas_shape)
        scale = 1.0 / sc.n_images  # the batched loss is the mean over every image's channels
        checked = list(range(0, sc.n_images, 8))  # 8 images against FP64 (own params each)

        def grads(i):
            rs, bs, sub, k0, k1 = subs[i]
            _, want_gp, want_gc = pd.Scene(1, sub.width, sub.height, sub.path_image, sub.path_cubics).backward(
                sub.points, sub.colours, target[i:i + 1])
            return (gp_all[pt_first[k0]:pt_first[k1]], want_gp * scale, gc_all[k0:k1], want_gc * scale)

        with ThreadPoolExecutor(min(8, len(os.sched_getaffinity(0)))) as ex:
            res = list(ex.map(grads, checked))
        many["grad_err_points"] = _grad_errors(np.concatenate([g[0] for g in res]), np.concatenate([g[1] for g in res]))
        many["grad_err_colours"] = _grad_errors(np.concatenate([g[2] for g in res]),
                                              np.concatenate([g[3] for g in res]))
        many["grad_checked_images"] = checked
    out["images_64x256"] = many
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-extra", action="store_true", help="skip the atlas, plain and differentiable measurements")
    ap.add_argument("--no-diff-check", action="store_true", help="skip the FP64 gradient checks")
    ap.add_argument("--tuning", default=None,
                    help="band_rows_min,lds_cells,dst_prefetch for the random-walk batch instead of autotuning "
                         "(profiling runs)")
    ap.add_argument("--atlas-tuning", default=None,
                    help="band_rows_min,lds_cells[,tile_groups] for the atlas instead of autotuning "
                         "(profiling runs)")
    ap.add_argument("--only", choices=["atlas", "diff", "plain"], default=None,
                    help="run only the atlas, the differentiable or the plain 256x256 Plotter measurement and "
                         "print its JSON (profiling runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {args.gpus} is not supported "
              "(the rasteriser is measured on one GPU's share of a sharded job)",
              file=sys.stderr)
        sys.exit(2)

    import footile_amd as ft
    from footile_amd import workloads
    from footile_amd.pixel import RGBA8P
    from footile_amd.plotter import Batch, DeviceBuffer, FillDesc

    if args.only == "atlas":
        print(json.dumps(bench_atlas(ft, args.steps, args.warmup, args.atlas_tuning, min(4.0, args.cpu_seconds))),
              flush=True)
        return
    if args.only == "plain":
        print(json.dumps(bench_plain(ft)), flush=True)
        return
    if args.only == "diff":
        print(json.dumps(bench_diff(args.steps, args.warmup, check=not args.no_diff_check)), flush=True)
        return

    W = H = 4096
    clr = (200, 120, 40, 255)
    ops = workloads.random_walk(1)
    print(f"[bench] random walk: {len(ops)} ops", file=sys.stderr, flush=True)
    A, B = DeviceBuffer(W * H * 4), DeviceBuffer(W * H * 4)
    A.zero()
    B.zero()
    both = np.concatenate([ops, ops])
    descs = [FillDesc(0, len(ops), A.ptr, W, H, W * 4, RGBA8P, rule=0, clr=clr),
             FillDesc(len(ops), len(ops), B.ptr, W, H, W * 4, RGBA8P, rule=1, clr=clr)]
    batch = Batch(both, descs)
    if args.tuning:
        b_rows, cells, pf = (int(v) for v in args.tuning.split(","))
        batch.set_tuning(b_rows, cells, bool(pf))
        tuned = {"band_rows_min": b_rows, "lds_cells": cells, "dst_prefetch": bool(pf), "tried": 0}
    else:
        from footile_amd.tune import autotune
        tuned = autotune(batch, descs, len(both), cache=None)
    print(f"[bench] tuned: {tuned}", file=sys.stderr, flush=True)
    A.zero()
    B.zero()
    batch.run()  # first step on zeroed rasters with the tuned geometry: the parity sample
    img_a, img_b = A.download(), B.download()
    offered, stored, edges, tops = batch.results()
    px_per_step = int(sum(W * (H - max(int(t), 0)) for t in tops if t < H))
    bytes_per_step = px_per_step * 8 + op_bytes(both)

    print("[bench] CPU oracle baseline + parity", file=sys.stderr, flush=True)
    (cpu_a, cpu_b, cpu_counts), cpu_times = cpu_random_walk(ops, args.cpu_seconds)
    parity = bool(np.array_equal(img_a, cpu_a) and np.array_equal(img_b, cpu_b)
                  and (int(offered[0]), int(stored[0])) == tuple(cpu_counts))
    if not parity:
        print("[bench] PARITY FAILURE: GPU rasters differ from the CPU oracle", file=sys.stderr)

    print(f"[bench] warmup {args.warmup}, timed {args.steps} one launch at a time", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        batch.launch()
        batch.sync()
    # Launches one at a time (each launch's front end waits for the previous
    # raster), with event timing of the raster kernel on every
    # TIMING_EVERY-th step: timed launches carry events in their first and
    # raster dispatches, which costs a few microseconds of queue time.
    batch.timing()
    t0 = time.perf_counter()
    for k in range(args.steps):
        batch.set_timing(k % TIMING_EVERY == 0)
        batch.launch()
    batch.sync()
    t1 = time.perf_counter()
    batch.set_timing(False)
    total_ms, raster_ms, runs = batch.timing()
    sec_per_step_serial = (t1 - t0) / args.steps
    # the roofline's kernel duration: the raster alone, as the profiler's
    # kernel trace of the same command sees it
    raster_avg_ms = raster_ms / max(runs, 1)
    achieved = bytes_per_step / (raster_avg_ms * 1e-3) / 1e9

    # The step: pipelined launches (two workspaces take turns; a launch's
    # front end overlaps the previous launch's raster, rasters keep launch
    # order).  First their bytes against the same launches one at a time.
    def three_launches(pipelined):
        A.zero()
        B.zero()
        for _ in range(3):
            batch.launch()
            if not pipelined:
                batch.sync()
        batch.sync()
        return A.download(), B.download()

    serial3 = three_launches(False)
    batch.set_pipeline(PIPE_DEPTH)
    piped3 = three_launches(True)
    pipe_ok = bool(np.array_equal(serial3[0], piped3[0]) and np.array_equal(serial3[1], piped3[1]))
    del serial3, piped3
    if not pipe_ok:
        print("[bench] PARITY FAILURE: pipelined launches differ from serial ones", file=sys.stderr)
    print(f"[bench] warmup {args.warmup}, timed {args.steps} pipelined", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        batch.launch()
    batch.sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        batch.launch()
    batch.sync()
    t1 = time.perf_counter()
    sec_per_step = (t1 - t0) / args.steps
    # then, untimed, the raster kernel's events every TIMING_EVERY-th
    # pipelined launch: the raster beside the front ends it overlaps, which
    # share its compute units (the events cost queue time, so they stay out
    # of the timed loop)
    batch.timing()
    for k in range(max(TIMING_EVERY * 8, args.steps // 2)):
        batch.set_timing(k % TIMING_EVERY == 0)
        batch.launch()
    batch.sync()
    batch.set_timing(False)
    total_p_ms, raster_p_ms, runs_p = batch.timing()
    batch.set_pipeline(False)

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_random_walk.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    extra = {
        "paths_per_s": 2 / sec_per_step,
        "launch": f"pipelined: {PIPE_DEPTH} workspaces take turns, later launches' front ends overlap an "
                  "earlier launch's raster",
        "pipelined_matches_serial": pipe_ok,
        "ms_per_step_one_at_a_time": sec_per_step_serial * 1e3,
        "device_ms_per_step": total_ms / max(runs, 1),
        "device_ms_per_launch_pipelined": total_p_ms / max(runs_p, 1),
        "raster_kernel_ms": raster_avg_ms,
        "raster_kernel_ms_pipelined": raster_p_ms / max(runs_p, 1),
        # the whole step against the same HBM roofline (bytes the raster must
        # move per step over the pipelined step time)
        "step_hbm_frac": bytes_per_step / sec_per_step / 1e9 / HBM_PEAK_GBS,
        "points_offered_per_fill": int(offered[0]),
        "points_stored_per_fill": int(stored[0]),
        "edges_per_fill": int(edges[0]),
        "top_rows": [int(t) for t in tops],
        "tuning": {"band_rows_min": tuned["band_rows_min"], "lds_cells": tuned["lds_cells"],
                   "dst_prefetch": tuned["dst_prefetch"], "candidates": tuned["tried"]},
    }
    if not args.no_extra:
        print("[bench] atlas", file=sys.stderr, flush=True)
        extra["atlas"] = bench_atlas(ft, min(25, max(3, args.steps // 4)), 1, args.atlas_tuning,
                                     min(4.0, args.cpu_seconds))
        print("[bench] plain 256", file=sys.stderr, flush=True)
        extra["plain_256"] = bench_plain(ft)
        print("[bench] random walk through the Plotter API", file=sys.stderr, flush=True)
        extra["random_walk_plotter_api"] = bench_plotter_random_walk(ft, ops)
        print("[bench] differentiable fills", file=sys.stderr, flush=True)
        extra["diff"] = bench_diff(min(10, max(4, args.steps // 2)), 2, check=not args.no_diff_check)

    cpu_step_px = 2 * W * (H - max(int(tops[0]), 0))
    cpu_med = float(np.median(cpu_times))
    cpu_mad = float(np.median(np.abs(cpu_times - cpu_med)))
    line = {
        "metric": "filled Mpixels/s + paths/s (bit-exact vs CPU) at 1 GPU; fwd+bwd grad-steps/s",
        "value": px_per_step / sec_per_step / 1e6,
        "unit": "Mpixels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec_per_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "i32 (16.16 Fixed) / u8",
        "data": "synthetic",
        "config": {"workload": "random_walk 4096x4096 Rgba8p, 10,000 cubics (seed 1), NonZero + EvenOdd",
                   "width": W, "height": H, "n_cubics": 10000, "fills_per_step": 2},
        "parity": parity and pipe_ok,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_raster", "algorithmic_bytes_per_launch": bytes_per_step,
                     "timing": "raster kernel alone (launches one at a time, every "
                               f"{TIMING_EVERY}th timed by HIP events on its stream)"},
        "cpu_baseline": {"value": cpu_step_px / cpu_med / 1e6, "unit": "Mpixels/s", "cores": 1, "kind": "port",
                         "sample": f"{len(cpu_times)} step(s) of the same two fills (oracle, C, 1 thread, "
                                   f"SSSE3 accumulate), {cpu_times.sum():.1f} s; median step "
                                   f"{cpu_med * 1e3:.1f} ms, MAD {cpu_mad * 1e3:.1f} ms"},
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
