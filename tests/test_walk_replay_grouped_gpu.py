"""Checkpoint and replay through the cell-grouped backward
(pf_walk_backward_replay_grouped): no record from the forward; the walk is
replayed band by band into a band-sized record, whose segments are sorted
by cell and differentiated, the gradients adding up over the bands.

Bar: gradients within 1e-3 relative L2 of the oracle's and within 1e-5 of
the replay backward's (pf_walk_backward: the same segments, another order
of float additions), for bands that split the tile order anywhere (1, 3
and 7 tiles, the whole frame), three cameras of 37x29 (tiles past the
image edge), a cap below the longest rays (those replayed in the gradient
launch), and per-camera texel colours.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as orc
from powerfoam import raytrace as rt
from powerfoam import workloads

from test_walk_gpu import _cams, _cpu_scene, _gpu_scene, _rel

pytestmark = pytest.mark.gpu
DEV = "cuda"
KEYS = ("spheres", "nsigmas", "texel_sites", "texel_rgb", "texel_height")


def _setup(sc, gs, cams):
    words = _cams(sc, cams)
    cams_dev = torch.from_numpy(words).to(DEV)
    W, H = cams[0].width, cams[0].height
    fwd = rt.launch_walk_forward(gs, cams_dev, W, H)
    rng = np.random.default_rng(1)
    target = torch.from_numpy(rng.uniform(0, 1, (len(cams), H, W, 3)).astype(np.float32)).to(DEV)
    gc = 2.0 * (fwd["color"] - target)
    gl = -0.3 * torch.exp(fwd["log_t"])
    return words, cams_dev, W, H, fwd, gc, gl


@pytest.mark.parametrize("band", [1, 3, 7, 1 << 20])
@pytest.mark.parametrize("cap", [8, 128])
def test_banded_replay_matches_replay_and_cpu(scene_textured, band, cap):
    sc = scene_textured
    gs, faces = _gpu_scene(sc)
    cams = workloads.sphere_cameras(np.random.default_rng(5), 3, 37, 29, radius=2.5)
    words, cams_dev, W, H, fwd, gc, gl = _setup(sc, gs, cams)
    print(f"[case] band {band} cap {cap}", flush=True)
    bnd = rt.launch_walk_backward(gs, cams_dev, W, H, fwd["color"], fwd["log_t"], gc, gl, grouped=True,
                                  band_tiles=band, replay_cap=cap)
    rep = rt.launch_walk_backward(gs, cams_dev, W, H, fwd["color"], fwd["log_t"], gc, gl)
    torch.cuda.synchronize()
    assert int(bnd["status"].item()) == 0
    for key in KEYS:
        assert _rel(bnd[key].cpu().numpy(), rep[key].cpu().numpy()) < 1e-5, key
    if band == 3 and cap == 8:
        cs = _cpu_scene(sc, faces)
        c = orc.walk_forward(cs, words)
        cg = orc.walk_backward(cs, words, c["color"], gc.cpu().numpy(), gl.cpu().numpy())
        for key in KEYS:
            assert _rel(bnd[key].cpu().numpy(), cg[key]) < 1e-3, key


def test_banded_replay_per_camera_colours(scene_textured):
    """Three cameras with their own texel colours: buckets per (camera,
    cell), camera b's colour gradients at stride N*K*3, equal to the
    grouped backward over the forward's whole record."""
    sc = scene_textured
    rng = np.random.default_rng(4)
    rgb = rng.uniform(0, 1, (3,) + sc["texel_rgb"].shape).astype(np.float32)
    gs1, _ = _gpu_scene(sc)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    gsb = rt.WalkScene(t(sc["points"]), t(sc["radii"]), t(sc["density"]), t(sc["normals"]), t(sc["texel_sites"]),
                       t(rgb), t(sc["texel_height"]), t(sc["adjacency"]), t(sc["offsets"]), gs1.faces)
    cams = workloads.sphere_cameras(np.random.default_rng(3), 3, 24, 24, radius=2.5)
    cams_dev = t(_cams(sc, cams))
    fwd = rt.launch_walk_forward(gsb, cams_dev, 24, 24, record_cap=64)
    gc = 2.0 * (fwd["color"] - 0.4)
    bnd = rt.launch_walk_backward(gsb, cams_dev, 24, 24, fwd["color"], fwd["log_t"], gc, None, grouped=True,
                                  band_tiles=5, replay_cap=64)
    grp = rt.launch_walk_backward(gsb, cams_dev, 24, 24, fwd["color"], fwd["log_t"], gc, None, record=fwd["record"],
                                  grouped=True)
    torch.cuda.synchronize()
    assert tuple(bnd["texel_rgb"].shape) == (3,) + sc["texel_rgb"].shape
    for key in KEYS:
        assert _rel(bnd[key].cpu().numpy(), grp[key].cpu().numpy()) < 1e-5, key


def test_banded_replay_refuses_bad_sizes(scene_textured):
    sc = scene_textured
    gs, _ = _gpu_scene(sc)
    cams = [workloads.front_camera(16, 16, hfov_deg=40.0)]
    words, cams_dev, W, H, fwd, gc, gl = _setup(sc, gs, cams)
    for kw in (dict(band_tiles=0), dict(replay_cap=12)):
        with pytest.raises(ValueError):
            rt.launch_walk_backward(gs, cams_dev, W, H, fwd["color"], fwd["log_t"], gc, gl, grouped=True, **kw)
    from powerfoam import _lib
    n = _lib.load().pf_walk_replay_grouped_workspace_bytes(gs.num_cells, 1, W, H, 128, 4)
    small = torch.empty(n - 256, dtype=torch.uint8, device=DEV)
    with pytest.raises(ValueError):
        rt.launch_walk_backward(gs, cams_dev, W, H, fwd["color"], fwd["log_t"], gc, gl, grouped=True,
                                band_tiles=4, workspace=small)
