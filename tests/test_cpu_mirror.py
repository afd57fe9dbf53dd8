"""The device stage code (footile_amd/csrc/ft_stages.h) run on the host against the oracle.

tests/cpu_mirror builds the same stage functions the gfx950 kernels call into a
host library and drives them through the same launch order, one element at a
time, with every buffer sized exactly.  The AddressSanitizer build turns any
out-of-range index in the stage logic into an abort, which is how an indexing
mistake is found without running it on a GPU.
"""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _libasan():
    if shutil.which("g++") is None:
        return None
    out = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    return out if out and os.path.isabs(out) and os.path.exists(out) else None


def test_mirror_matches_oracle_under_asan():
    asan = _libasan()
    if asan is None:
        pytest.skip("libasan not available")
    from tests.cpu_mirror import mirror
    mirror.build()
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0", MIRROR_ASAN="1")
    r = subprocess.run([sys.executable, os.path.join(HERE, "cpu_mirror", "check_cases.py"), "--quick"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert "mirror ok" in r.stdout


def test_mirror_random_walk_4096(oracle):
    """The benchmark workload at full size, both fill rules, bit-exact."""
    from footile_amd import workloads
    from footile_amd.pixel import RGBA8P
    from tests.cpu_mirror import mirror
    ops = workloads.random_walk(1)
    for rule in (0, 1):
        o = oracle.Plotter(4096, 4096, RGBA8P)
        assert o.fill(rule, ops, (255, 208, 128, 255)) == 0
        rc, px, counts = mirror.fill(4096, 4096, RGBA8P, rule, ops, (255, 208, 128, 255))
        assert rc == 0
        assert counts == o.last_counts()
        assert np.array_equal(px, o.pixels())


def test_row_cells_closed_form_matches_pixel_walk():
    """The per-row set-up in closed form and the 32-bit cell walk
    (row_setup / row_walk, shared with the gfx950 raster) against the
    pixel-by-pixel walk of fig.rs:285-321 on random edges, rows and raster
    widths, including edges left of x = 0 and past the right edge."""
    from tests.cpu_mirror import mirror
    assert mirror.row_cells_check(12345, 200000) == 0


def test_curve_lanes_any_split_depth():
    """A curve flattened by 2^sp lanes (flatten_share, shared with the gfx950
    kernels; the pipeline splits 3 levels down, small figures 5) gives the
    same points in the same order as the whole subdivision tree walked from
    the root, for sp = 0..6, coarse and fine tolerances and large curves."""
    from tests.cpu_mirror import mirror
    assert mirror.split_check(5, 3000) == 0
