"""Committed fixtures (tests/golden/walk_256.npz, written by make_golden.py):
the CPU restatement must reproduce them bit for bit, and its walk must agree
with the FP64 envelope oracle stored beside them (visited cells exactly on
rays away from ties and images to 1e-4 with fp32 faces; fp16 faces agree
on average)."""
import os

import numpy as np
import pytest

from oracle import oracle as orc

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "walk_256.npz")


@pytest.fixture(scope="module")
def fx():
    return dict(np.load(FIX))


def _scene(f):
    return orc.Scene(f["points"], f["radii"], f["density"], f["normals"], f["texel_sites"], f["texel_rgb"],
                     f["texel_height"], f["adjacency"], f["offsets"], f["faces_f16"])


def test_restatement_reproduces_walk_fixture(fx):
    w = orc.walk_forward(_scene(fx), fx["cam_words"], seq_cap=64)
    assert np.array_equal(w["visits"], fx["walk_visits"])
    assert np.array_equal(w["seq_hash"], fx["walk_seq_hash"])
    assert np.array_equal(w["seq"], fx["walk_seq"])
    for k in ("color", "log_t", "depth"):
        assert np.array_equal(w[k], fx["walk_" + k]), k
    gc = (2.0 * (w["color"] - fx["target"])).astype(np.float32)
    g = orc.walk_backward(_scene(fx), fx["cam_words"], w["color"], gc, nthreads=1)
    for k in ("spheres", "nsigmas", "texel_sites", "texel_rgb", "texel_height"):
        assert np.array_equal(g[k], fx["grad_" + k]), k


def test_walk_fixture_agrees_with_fp64_owners(fx):
    faces32 = orc.walk_faces(fx["points"], fx["radii"], fx["adjacency"], fx["offsets"], f16=False)
    s32 = orc.Scene(fx["points"], fx["radii"], fx["density"], fx["normals"], fx["texel_sites"], fx["texel_rgb"],
                    fx["texel_height"], fx["adjacency"], fx["offsets"], faces32)
    w32 = orc.walk_forward(s32, fx["cam_words"], seq_cap=64)
    assert np.array_equal(w32["seq"], fx["walk32_seq"]) and np.array_equal(w32["color"], fx["walk32_color"])
    seq = fx["walk32_seq"].reshape(256, 64)
    vis = fx["walk32_visits"].reshape(256)
    ref = fx["fp64_seq"]
    margin = fx["fp64_margin"].reshape(256)
    checked = 0
    for k in range(256):
        if margin[k] <= 1e-4:
            continue  # a near-tie that fp16 faces may resolve either way
        n = min(vis[k], 64)
        r = ref[k][ref[k] >= 0][:n]
        assert seq[k][:len(r)].tolist() == r.tolist(), k
        checked += 1
    assert checked > 200
    assert np.abs(fx["walk32_color"][0] - fx["fp64_color"]).max() < 1e-4
    assert np.abs(np.exp(fx["walk32_log_t"][0]) - np.exp(fx["fp64_log_t"])).max() < 1e-4
    # fp16 face storage moves cell boundaries slightly: same picture on average
    assert np.abs(fx["walk_color"][0] - fx["fp64_color"]).mean() < 1e-3


def test_restatement_reproduces_raster_and_colour_fixtures(fx):
    cadj, coff = orc.cech(fx["points"], fx["radii"])
    assert np.array_equal(cadj, fx["cech_adjacency"]) and np.array_equal(coff, fx["cech_offsets"])
    spheres = np.concatenate([fx["points"], fx["radii"][:, None]], 1).astype(np.float32)
    rfaces = orc.raster_faces(spheres, cadj, coff, f16=True)
    rs = orc.RasterScene(fx["points"], fx["radii"], fx["density"], fx["normals"], fx["texel_sites"], fx["texel_rgb"],
                         fx["texel_height"], cadj, coff, rfaces)
    boff, prims = orc.raster_bin(spheres, fx["cam_words"][0])
    rf = orc.raster_forward(rs, fx["cam_words"][0], boff, prims)
    assert np.array_equal(rf["color"], fx["raster_color"]) and np.array_equal(rf["log_t"], fx["raster_log_t"])
    v, w = orc.sv_forward(fx["texel_sites"], fx["sv_eye"], fx["sv_forward"], float(fx["sv_cutoff"]), fx["sv_axes"],
                          fx["sv_values"], fx["sv_temps"])
    assert np.array_equal(v, fx["sv_value"]) and np.array_equal(w, fx["sv_wsum"])
    # raster and walk render the same scene: same picture up to fp16 rounding
    assert np.abs(fx["raster_color"] - fx["walk_color"][0]).mean() < 1e-3
