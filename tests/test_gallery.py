"""The gallery renderer (spd_k_amd/gallery.py) on the CPU: PNG round trip
and CRC check, the colour ramp's ends, the polyline raster, and panels of
tiny oracle runs in the driver's output format (1-D, 2-D with trouble maps,
the field-cut layout, 3-D mid-plane, induction)."""
import os
import struct
import zlib

import numpy as np
import pytest

import oracle_py as op
from spd_k_amd import gallery


def test_png_round_trip_and_crc():
    rng = np.random.default_rng(0)
    img = rng.integers(0, 256, (7, 5, 3), dtype=np.uint8)
    data = gallery.png_bytes(img)
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    w, h = struct.unpack(">II", data[16:24])
    assert (w, h) == (5, 7)
    assert (gallery.read_png(data) == img).all()
    # the image rows as zlib holds them: filter byte 0, then RGB
    n, = struct.unpack(">I", data[33:37])
    raw = zlib.decompress(data[41:41 + n])
    assert raw[0] == 0 and raw[1:4] == bytes(img[0, 0])
    bad = bytearray(data)
    bad[45] ^= 1  # inside the IDAT body
    with pytest.raises(ValueError, match="CRC"):
        gallery.read_png(bytes(bad))


def test_colour_ramp_ends_and_non_finite():
    c = gallery.colourise(np.array([0.0, 1.0, 0.5, -3.0, 7.0, np.nan]), 0.0, 1.0)
    ramp = gallery.RAMPS["viridis"]
    assert tuple(c[0]) == ramp[0] and tuple(c[1]) == ramp[-1] and tuple(c[2]) == ramp[2]
    assert tuple(c[3]) == ramp[0] and tuple(c[4]) == ramp[-1]  # clipped
    assert tuple(c[5]) == (255, 0, 255)
    flat = gallery.colourise(np.ones(3), 1.0, 1.0)  # empty range: the middle colour
    assert (flat == flat[0]).all()


def test_chart_draws_the_diagonal():
    img = gallery.chart(np.array([0.0, 1.0]), [(np.array([0.0, 1.0]), (255, 0, 0), False)], w=60, h=60, pad=5)
    red = np.argwhere((img == (255, 0, 0)).all(axis=2))
    assert len(red) > 40
    # rows fall as columns rise (y up), inside the frame
    r, c = red[:, 0], red[:, 1]
    assert np.corrcoef(r, c)[0, 1] < -0.99
    assert r.min() >= 1 and c.max() <= 60


CASES = {c["id"]: c for c in gallery.CASES}


@pytest.mark.parametrize("cid,extra", [
    ("sod_1d", ["mesh/nx1=16", "time/tlim=0.02", "output/dt=0.01"]),
    ("kelvin_helmholtz_2d", ["mesh/nx1=6", "mesh/nx2=6", "time/tlim=0.01", "output/dt=0.005"]),
    ("sine_wave_2d", ["mesh/nx1=4", "mesh/nx2=4", "time/tlim=0.02", "output/dt=0.02"]),
    ("sedov_3d", ["mesh/nx1=4", "mesh/nx2=4", "mesh/nx3=4", "time/tlim=0.002", "output/dt=0.002"]),
    ("induction_loop_3d", ["mesh/nx1=4", "mesh/nx2=4", "mesh/nx3=4", "time/tlim=0.01", "output/dt=0.01"]),
])
def test_panel_of_an_oracle_run(tmp_path, cid, extra):
    case = dict(CASES[cid], overrides=CASES[cid]["overrides"] + extra)
    d = str(tmp_path / cid)
    op.run_oracle(["-i", os.path.join(op.INPUTS, case["input"])] + case["overrides"], d)
    img = gallery.render(case, d)
    assert img.dtype == np.uint8 and img.ndim == 3 and img.shape[2] == 3
    assert (gallery.read_png(gallery.png_bytes(img)) == img).all()
    colours = {tuple(int(v) for v in x) for x in np.unique(img.reshape(-1, 3), axis=0)}
    if cid == "sod_1d":  # profiles: both lines drawn
        assert set(gallery.LINE_COLOURS[:2]) <= colours
    else:  # maps: not blank
        assert len(colours) > 16
    if cid == "kelvin_helmholtz_2d":  # two maps per row: density and the trouble map (magma ramp)
        g = op.grid(d)
        t = gallery.troubles(d, g, op.output_indices(d)[-1])
        assert t is not None and t.shape == (6 * 4, 6 * 4) and 0.0 <= t.min() and t.max() <= 1.0


def test_index_lists_every_case(tmp_path):
    p = str(tmp_path / "README.md")
    gallery.write_index(p, gallery.CASES, ".")
    text = open(p).read()
    for c in gallery.CASES:
        assert f"({'.'}/{c['id']}.png)" in text
    assert len([c for c in gallery.CASES]) == 18
    assert all(gallery.fast_case(c) is None for c in gallery.CASES if c["section"] == "knobs")
