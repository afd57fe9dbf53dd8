"""footile's example programs (examples/*.rs) as a visual regression set.

tests/golden/gallery.json holds each example's paths (as ops), the Plotter
calls it makes (fill / stroke, colours, rule, join, transform, initial
pixels) and the SHA-256 of the raster the oracle renders for it
(tests/golden/make_gallery.py wrote it).  The CPU tests check that the oracle
still renders those bytes and that the PNG writer round-trips them; the GPU
test renders every example through the Plotter API and compares bytes.
"""
import hashlib
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gallery.json")
BPP = {"matte8": 1, "rgba8p": 4, "graya8p": 2}


def _examples():
    with open(GOLDEN) as f:
        return json.load(f)["examples"]


EXAMPLES = _examples()
IDS = [e["name"] for e in EXAMPLES]


def _ops(rows):
    from footile_amd.path import ops_array
    return ops_array([(int(r[0]), tuple(r[1:])) for r in rows])


def _clr(c):
    return tuple(c["clr"]) + (0,) * (4 - len(c["clr"]))


def _oracle_raster(oracle, ex):
    fmt = {"matte8": oracle.MATTE8, "rgba8p": oracle.RGBA8P, "graya8p": oracle.GRAYA8P}[ex["fmt"]]
    o = oracle.Plotter(ex["width"], ex["height"], fmt)
    if ex["init"] is not None:
        o.set_pixels(np.tile(np.asarray(ex["init"], np.uint8), ex["width"] * ex["height"]))
    if ex["transform"] is not None:
        o.set_transform(ex["transform"])
    if ex["join"] is not None:
        o.set_join({"miter": 0, "bevel": 1, "round": 2}[ex["join"][0]], ex["join"][1])
    for c in ex["calls"]:
        ops = _ops(ex["paths"][c["path"]])
        rc = o.fill(c["rule"], ops, _clr(c)) if c["op"] == "fill" else o.stroke(ops, _clr(c))
        assert rc == 0
    return o.pixels()


def test_gallery_covers_the_examples():
    names = set(IDS)
    for n in ["cubic", "curve", "drop", "figure", "fishy", "fishy2", "heptagram", "letter", "over", "quad",
              "round", "stroke", "stroke2", "teeth"]:
        assert n in names
    # with_pixels of zeros (fishy2.rs) is the same raster as with_clear (fishy.rs)
    by = {e["name"]: e for e in EXAMPLES}
    assert by["fishy"]["sha256"] == by["fishy2"]["sha256"]
    assert by["heptagram"]["calls"][0]["rule"] == 1  # EvenOdd
    assert by["drop"]["fmt"] == "graya8p" and by["round"]["join"][0] == "round"


@pytest.mark.parametrize("ex", EXAMPLES, ids=IDS)
def test_oracle_renders_the_recorded_bytes(oracle, ex):
    px = _oracle_raster(oracle, ex)
    assert hashlib.sha256(px.tobytes()).hexdigest() == ex["sha256"]
    assert int((px.reshape(-1, BPP[ex["fmt"]])[:, -1] > 0).sum()) == ex["covered"]


@pytest.mark.parametrize("ex", EXAMPLES, ids=IDS)
def test_png_round_trip(oracle, ex):
    from footile_amd import png
    from footile_amd.pixel import GRAYA8P, MATTE8, RGBA8P
    fmt = {"matte8": MATTE8, "rgba8p": RGBA8P, "graya8p": GRAYA8P}[ex["fmt"]]
    px = _oracle_raster(oracle, ex).reshape(ex["height"], ex["width"], BPP[ex["fmt"]])
    got, ct = png.decode(png.encode(px, fmt))
    assert ct == {MATTE8: 0, GRAYA8P: 4, RGBA8P: 6}[fmt]
    assert np.array_equal(got, png.straight_alpha(px, fmt))
    # opaque pixels keep their bytes; straight colour times alpha gives back the premultiplied byte
    a = px[..., -1].astype(np.int64)
    if fmt != MATTE8:
        back = (got[..., :-1].astype(np.int64) * a[..., None] + 127) // 255
        assert np.all(np.abs(back - px[..., :-1]) <= 1)


@pytest.mark.gpu
@pytest.mark.parametrize("ex", EXAMPLES, ids=IDS)
def test_gallery_on_the_gpu(oracle, ex):
    import footile_amd as ft
    from footile_amd.pixel import GRAYA8P, MATTE8, RGBA8P
    fmt = {"matte8": MATTE8, "rgba8p": RGBA8P, "graya8p": GRAYA8P}[ex["fmt"]]
    init = None
    if ex["init"] is not None:
        init = np.tile(np.asarray(ex["init"], np.uint8), ex["width"] * ex["height"])
    p = ft.Plotter(ft.Raster(ex["width"], ex["height"], fmt, init))
    if ex["transform"] is not None:
        p.set_transform(ex["transform"])
    if ex["join"] is not None:
        p.set_join(({"miter": 0, "bevel": 1, "round": 2}[ex["join"][0]], ex["join"][1]))
    for c in ex["calls"]:
        ops = _ops(ex["paths"][c["path"]])
        if c["op"] == "fill":
            p.fill(c["rule"], ops, _clr(c))
        else:
            p.stroke(ops, _clr(c))
    got = p.raster().as_u8_slice()
    want = _oracle_raster(oracle, ex)
    assert np.array_equal(got, want), f"{ex['name']}: {int((got != want).sum())} bytes differ"
    assert hashlib.sha256(got.tobytes()).hexdigest() == ex["sha256"]
