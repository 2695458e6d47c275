"""The gfx950 path against the committed fixtures alone: walk forward
(visit counts, sequence hashes and first 64 cells bit-exact; colour, log T,
depth within 1e-4), walk backward (1e-3 relative L2 per gradient array),
raster forward (1e-4) and the spherical-Voronoi pre-pass (bit-exact)."""
import os
import types

import numpy as np
import pytest
import torch

from powerfoam import raytrace as rt
from powerfoam.bvh import AABBTree
from powerfoam.color_fn import SphericalVoronoi
from powerfoam.rasterize import Rasterizer
from powerfoam.camera import TorchCamera

pytestmark = pytest.mark.gpu
DEV = "cuda"
FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "walk_256.npz")


@pytest.fixture(scope="module")
def fx():
    return dict(np.load(FIX))


def t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _walk_scene(f):
    faces = t(f["faces_f16"]).view(torch.float16)
    return rt.WalkScene(t(f["points"]), t(f["radii"]), t(f["density"]), t(f["normals"]), t(f["texel_sites"]),
                        t(f["texel_rgb"]), t(f["texel_height"]), t(f["adjacency"]), t(f["offsets"]), faces)


def _camera(f):
    w = f["cam_words"][0]
    fl = w.view(np.float32)
    return TorchCamera(eye=torch.from_numpy(fl[0:3].copy()), right=torch.from_numpy(fl[3:6].copy()),
                       up=torch.from_numpy(fl[6:9].copy()), width=int(w[10]), height=int(w[11]))


def test_walk_matches_fixture(fx):
    s = _walk_scene(fx)
    cams = t(fx["cam_words"])
    g = rt.launch_walk_forward(s, cams, 16, 16, want_sequences=True, seq_cap=64)
    torch.cuda.synchronize()
    assert np.array_equal(g["visits"].cpu().numpy(), fx["walk_visits"])
    assert np.array_equal(g["seq_hash"].cpu().numpy().view(np.uint64), fx["walk_seq_hash"])
    assert np.array_equal(g["seq"].cpu().numpy(), fx["walk_seq"])
    for k in ("color", "log_t", "depth"):
        assert np.abs(g[k].cpu().numpy() - fx["walk_" + k]).max() <= 1e-4, k
    gc = 2.0 * (g["color"] - t(fx["target"]))
    gr = rt.launch_walk_backward(s, cams, 16, 16, g["color"], g["log_t"], gc, None)
    torch.cuda.synchronize()
    assert int(gr["status"].item()) == 0
    for k in ("spheres", "nsigmas", "texel_sites", "texel_rgb", "texel_height"):
        a, b = gr[k].cpu().numpy(), fx["grad_" + k]
        assert np.linalg.norm(a - b) <= 1e-3 * max(np.linalg.norm(b), 1e-30), k


def test_raster_matches_fixture(fx):
    tree = AABBTree(DEV)
    tree.update(t(fx["points"]), t(fx["radii"]))
    adj, off = tree.build_cech_complex()
    assert np.array_equal(adj.cpu().numpy(), fx["cech_adjacency"])
    assert np.array_equal(off.cpu().numpy(), fx["cech_offsets"])
    rz = Rasterizer(types.SimpleNamespace(num_texel_sites=8), DEV)
    with torch.no_grad():
        out = rz.forward(_camera(fx), None, t(fx["points"]), t(fx["radii"]), t(fx["density"]), t(fx["normals"]),
                         t(fx["texel_sites"]), t(fx["texel_rgb"]), t(fx["texel_height"]), adj, off, None, False)
    assert np.abs(out[0].cpu().numpy() - fx["raster_color"]).max() <= 1e-4
    assert np.abs(out[1].cpu().numpy() - (1 - np.exp(fx["raster_log_t"]))).max() <= 1e-4


def test_colour_prepass_matches_fixture(fx):
    sv = SphericalVoronoi(types.SimpleNamespace(sv_dof=fx["sv_axes"].shape[0]), DEV)
    eye, fwd = torch.from_numpy(fx["sv_eye"]), torch.from_numpy(fx["sv_forward"])
    v = torch.empty(fx["sv_value"].shape, device=DEV)
    w = torch.empty(fx["sv_wsum"].shape, device=DEV)
    from powerfoam import _lib

    sites = t(fx["texel_sites"].reshape(-1, 3))
    ax, vals, temps = t(fx["sv_axes"]), t(fx["sv_values"]), t(fx["sv_temps"])  # held across the launch
    rc = _lib.load().pf_sv_forward(_lib.ptr(sites), sites.shape[0], eye.data_ptr(), fwd.data_ptr(),
                                   float(fx["sv_cutoff"]), _lib.ptr(ax), _lib.ptr(vals), _lib.ptr(temps), sv.sv_dof,
                                   _lib.PF_F32, _lib.ptr(v), _lib.ptr(w), _lib.stream_ptr())
    _lib.check(rc, "pf_sv_forward")
    torch.cuda.synchronize()
    assert np.array_equal(v.cpu().numpy(), fx["sv_value"]) and np.array_equal(w.cpu().numpy(), fx["sv_wsum"])
