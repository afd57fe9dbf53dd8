"""ft_batch_run_noalloc (the batch fill for callers that cannot allocate)
against the existing ft_batch_new + ft_batch_run, through the C ABI: the same
rasters byte for byte and the same point counts, on two calls in a row (the
second one smaller, over the same per-device scratch), plus the sizes it
rejects."""
import ctypes

import numpy as np
import pytest

from footile_amd import workloads
from footile_amd._lib import FillDescC, lib
from footile_amd.path import ops_array
from footile_amd.pixel import RGBA8P

pytestmark = pytest.mark.gpu

ENOMEM = 2


def _noalloc(ops, descs):
    a = ops_array(ops)
    arr = (FillDescC * max(len(descs), 1))(*[d.c for d in descs])
    off = np.zeros(max(len(descs), 1), np.uint64)
    st = np.zeros(max(len(descs), 1), np.uint64)
    rc = lib().ft_batch_run_noalloc(a.ctypes.data_as(ctypes.c_void_p), len(a), arr, len(descs), None,
                                    off.ctypes.data_as(ctypes.c_void_p), st.ctypes.data_as(ctypes.c_void_p))
    return rc, off[:len(descs)], st[:len(descs)]


def _random_walk_case():
    """Two 1024x1024 Rgba8p fills (NonZero, EvenOdd) of a 2,000-cubic random
    walk over existing pixels."""
    from footile_amd.plotter import DeviceBuffer, FillDesc
    ops = workloads.random_walk(1, n_cubics=2000, canvas=1024)
    both = np.concatenate([ops, ops])
    init = np.random.default_rng(3).integers(0, 256, size=2 * 1024 * 1024 * 4).astype(np.uint8)
    bufs = []

    def make():
        d = DeviceBuffer(2 * 1024 * 1024 * 4)
        d.upload(init)
        bufs.append(d)
        return d, [FillDesc(0, len(ops), d.ptr, 1024, 1024, 4096, RGBA8P, rule=0, clr=(200, 120, 40, 255)),
                   FillDesc(len(ops), len(ops), d.ptr + 1024 * 4096, 1024, 1024, 4096, RGBA8P, rule=1,
                            clr=(30, 60, 90, 128))]
    return both, make


def _atlas_case(n_paths, side):
    from footile_amd.plotter import DeviceBuffer, FillDesc
    ops, tiles = workloads.atlas(2, n_paths=n_paths, atlas_side=side)

    def make():
        d = DeviceBuffer(side * side)
        d.upload(np.full(side * side, 77, dtype=np.uint8))
        return d, [FillDesc(int(b), int(c), d.ptr + int(y) * side + int(x), int(s), int(s), side, fresh=True)
                   for b, c, x, y, s in tiles]
    return ops, make


def _compare(ops, make):
    from footile_amd.plotter import Batch
    want_buf, descs = make()
    batch = Batch(ops, descs)
    batch.run()
    want = want_buf.download()
    w_off, w_st, _ = batch.counts()
    got_buf, descs = make()
    rc, off, st = _noalloc(ops, descs)
    assert rc == 0, rc
    assert np.array_equal(got_buf.download(), want)
    assert np.array_equal(off, np.asarray(w_off, np.uint64)) and np.array_equal(st, np.asarray(w_st, np.uint64))


def test_noalloc_matches_batch_run_twice_then_smaller():
    _compare(*_random_walk_case())
    _compare(*_atlas_case(160, 2048))
    _compare(*_atlas_case(24, 512))  # smaller again: the scratch keeps the last call's contents


def test_noalloc_rejects_sizes_past_its_scratch():
    from footile_amd.plotter import DeviceBuffer, FillDesc
    d = DeviceBuffer(64 * 64)
    too_many_ops = np.zeros(1 << 20, dtype=ops_array([]).dtype)
    rc, _, _ = _noalloc(too_many_ops, [FillDesc(0, 1, d.ptr, 64, 64, 64)])
    assert rc == ENOMEM
    # a run whose events outgrow the scratch (10,000 cubics flattened at the
    # finest tolerance: about 800k events): FT_ENOMEM after the run
    big = DeviceBuffer(4096 * 4096)
    ops = workloads.random_walk(1)
    rc, _, _ = _noalloc(ops, [FillDesc(0, len(ops), big.ptr, 4096, 4096, 4096, tolerance=0.01)])
    assert rc == ENOMEM
    # and the next call in range still works
    _compare(*_atlas_case(24, 512))
