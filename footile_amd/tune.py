"""Autotuner of a batch's raster geometry (ft_batch_set_tuning).

Candidates are (band_rows_min, lds_cells, dst_prefetch): band height for
edge binning, the LDS cell budget that fixes the rows per raster block, and
(for read-modify-write rasters) whether raster blocks load their destination
before the cell pass.  Each
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
        batch.set_tuning(e["band_rows_min"], e["lds_cells"], e.get("dst_prefetch", True))
        return dict(e, cached=True, tried=0)
    # the destination prefetch only matters for read-modify-write rasters
    rmw = any(int(d.c.fmt) != 0 and not int(d.c.fresh) for d in descs)
    prefetch = (True, False) if rmw else (True,)
    cands = candidates or [(b, c, p) for b in BAND_ROWS for c in LDS_CELLS for p in prefetch]
    best = None
    t0 = time.perf_counter()
    for b, c, p in cands:
        try:
            batch.set_tuning(b, c, p)
        except FootileError as e:
            if e.code != 5:  # FT_ENOTSUP
                raise
            continue  # more LDS per block than the device has
        batch.run()
        ms = _time(batch, reps)
        if verbose:
            print(f"[tune] band_rows_min={b} lds_cells={c} dst_prefetch={p}: {ms:.3f} ms")
        if best is None or ms < best[3]:
            best = (b, c, p, ms)
    batch.set_tuning(best[0], best[1], best[2])
    entry = {"band_rows_min": best[0], "lds_cells": best[1], "dst_prefetch": best[2], "ms": best[3],
             "tune_s": time.perf_counter() - t0}
    if cache:
        try:
            os.makedirs(os.path.dirname(cache), exist_ok=True)
            table[key] = entry
            with open(cache, "w") as f:
                json.dump(table, f, indent=1, sort_keys=True)
        except OSError:
            pass
    return dict(entry, cached=False, tried=len(cands))
