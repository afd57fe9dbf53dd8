"""Autotuner of a batch's raster geometry (ft_batch_set_tuning).

Candidates are (band_rows_min, lds_cells, dst_prefetch, tile_groups): band
height for edge binning, the LDS cell budget that fixes the rows per raster
block, (for read-modify-write rasters) whether raster blocks load their
destination before the cell pass, and the largest path (in LDS row groups)
the tile scheduler takes instead of the bands.  Each
candidate is timed on the batch itself with device events (whole launch
sequence, median of a few launches after one warm-up run); the fastest wins.
Results are cached in a JSON file keyed by a signature of the batch layout
(path count, raster width/height histogram in powers of two, formats, op
count), so a workload is tuned once per machine.
"""
import hashlib
import json
import os
import time

import numpy as np

from ._lib import FootileError

BAND_ROWS = (4, 8, 16, 32, 64)
LDS_CELLS = (2048, 3072, 4096, 8192, 16384)
# tile scheduler candidates: paths of at most TILE_GROUPS LDS row groups of
# TILE_CELLS cells leave the bands
TILE_GROUPS = (1, 2, 4, 8)
TILE_CELLS = (4096, 8192, 16384, 32768)
TILE_BAND_ROWS = (8, 16)
DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "footile_amd", "tuning.json")


def signature(descs, n_ops):
    """Stable key of a batch layout."""
    w = np.array([d.c.width for d in descs], np.int64)
    h = np.array([d.c.height for d in descs], np.int64)
    fmt = sorted({int(d.c.fmt) for d in descs})
    def hist(v):
        b = np.floor(np.log2(np.maximum(v, 1))).astype(np.int64)
        return {int(k): int(c) for k, c in zip(*np.unique(b, return_counts=True))}
    key = json.dumps({"paths": len(descs), "w": hist(w), "h": hist(h), "fmt": fmt,
                      "ops_log2": int(np.log2(max(n_ops, 1)))}, sort_keys=True)
    return hashlib.sha1(key.encode()).hexdigest()[:16]


def _tile_candidates(descs, prefetch):
    """Tile scheduler settings that move at least a tenth of the paths out of
    the bands (a path of height h and cell row stride wp takes
    ceil(h / (lds_cells // wp)) groups)."""
    w = np.array([d.c.width for d in descs], np.int64)
    h = np.array([d.c.height for d in descs], np.int64)
    wp = (w + 15) // 16 * 16
    out = []
    for c in TILE_CELLS:
        groups = -(-h // np.maximum(1, c // np.maximum(wp, 1)))
        for tg in TILE_GROUPS:
            if np.count_nonzero(groups <= tg) * 10 >= max(len(descs), 1):
                out += [(b, c, p, tg) for b in TILE_BAND_ROWS for p in prefetch]
    return out


def _time(batch, reps):
    batch.set_timing(True)
    batch.timing()
    for _ in range(reps):
        batch.launch()
    batch.sync()
    total, _, runs = batch.timing()
    return total / max(runs, 1)


def autotune(batch, descs, n_ops, cache=DEFAULT_CACHE, reps=3, candidates=None, verbose=False):
    """Pick and apply the fastest (band_rows_min, lds_cells) for `batch`.
    Returns {"band_rows_min", "lds_cells", "ms", "cached", "tried"}."""
    key = signature(descs, n_ops)
    table = {}
    if cache and os.path.exists(cache):
        try:
            with open(cache) as f:
                table = json.load(f)
        except (OSError, ValueError):
            table = {}
    if key in table:
        e = table[key]
        batch.set_tuning(e["band_rows_min"], e["lds_cells"], e.get("dst_prefetch", True), e.get("tile_groups", 0))
        return dict(e, cached=True, tried=0)
    # the destination prefetch only matters for read-modify-write rasters
    rmw = any(int(d.c.fmt) != 0 and not int(d.c.fresh) for d in descs)
    prefetch = (True, False) if rmw else (True,)
    cands = candidates or ([(b, c, p, 0) for b in BAND_ROWS for c in LDS_CELLS for p in prefetch] +
                           _tile_candidates(descs, prefetch))
    best = None
    t0 = time.perf_counter()
    for b, c, p, tg in cands:
        try:
            batch.set_tuning(b, c, p, tg)
        except FootileError as e:
            if e.code != 5:  # FT_ENOTSUP
                raise
            continue  # more LDS per block than the device has
        batch.run()
        ms = _time(batch, reps)
        if verbose:
            print(f"[tune] band_rows_min={b} lds_cells={c} dst_prefetch={p} tile_groups={tg}: {ms:.3f} ms")
        if best is None or ms < best[4]:
            best = (b, c, p, tg, ms)
    batch.set_tuning(*best[:4])
    entry = {"band_rows_min": best[0], "lds_cells": best[1], "dst_prefetch": best[2], "tile_groups": best[3],
             "ms": best[4], "tune_s": time.perf_counter() - t0}
    if cache:
        try:
            os.makedirs(os.path.dirname(cache), exist_ok=True)
            table[key] = entry
            with open(cache, "w") as f:
                json.dump(table, f, indent=1, sort_keys=True)
        except OSError:
            pass
    return dict(entry, cached=False, tried=len(cands))
