"""Parity of the gfx950 rasterizer and Čech builder (through the C ABI)
with the CPU restatement.  Bar: tile lists and early-stop positions equal,
images within 1e-4 max-abs, per-cell sums and gradients within 1e-3
relative L2; the Čech CSR equal to brute force."""
import types

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from powerfoam import workloads
from powerfoam.bvh import AABBTree
from powerfoam.camera import cameras_to_device, pack_cameras
from powerfoam.rasterize import Rasterizer, VisOptions, bin_tiles, raster_faces

pytestmark = pytest.mark.gpu
DEV = "cuda"
TOL = 1e-4
ARGS = types.SimpleNamespace(render_objective="volume", num_texel_sites=8, is_pinhole=True,
                             disable_coop_prim_load=False, disable_coop_adj_load=False)


def t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.fixture(scope="module")
def cech4k(scene4k):
    return orc.cech(scene4k["points"], scene4k["radii"])


def test_cech_equals_brute_force(scene4k, cech4k):
    tree = AABBTree(DEV)
    tree.update(t(scene4k["points"]), t(scene4k["radii"]))
    adj, off = tree.build_cech_complex()
    want_adj, want_off = cech4k
    assert adj.dtype == torch.int32 and off.dtype == torch.int32
    assert np.array_equal(off.cpu().numpy(), want_off)
    assert np.array_equal(adj.cpu().numpy(), want_adj)


def test_cech_textured_and_isolated_spheres(scene_textured):
    sc = scene_textured
    pts = np.concatenate([sc["points"], [[5.0, 5.0, 5.0]]]).astype(np.float32)
    rad = np.concatenate([sc["radii"], [0.01]]).astype(np.float32)
    tree = AABBTree(DEV)
    tree.update(t(pts), t(rad))
    adj, off = tree.build_cech_complex()
    wa, wo = orc.cech(pts, rad)
    assert np.array_equal(off.cpu().numpy(), wo) and np.array_equal(adj.cpu().numpy(), wa)


def test_cech_heavy_tailed_radii_equals_brute_force():
    """60,000 cells with Pareto radii (most small, a few up to 200 times the
    median, as trained scenes have): the grid takes several levels, level 0
    sized by the mean radius, and the edge set still equals brute force.
    One grid sized by the largest radius would hold ~10^3 times more
    candidates per query."""
    import time

    rng = np.random.default_rng(11)
    sc = workloads.gmm_scene(60_000, seed=4)
    rad = (0.004 * (1.0 + rng.pareto(1.5, 60_000))).clip(max=0.8).astype(np.float32)
    tree = AABBTree(DEV)
    tree.update(t(sc["points"]), t(rad))
    assert tree.grid.levels >= 5
    t0 = time.perf_counter()
    adj, off = tree.build_cech_complex()
    torch.cuda.synchronize()
    took = time.perf_counter() - t0
    wa, wo = orc.cech(sc["points"], rad)
    assert np.array_equal(off.cpu().numpy(), wo)
    assert np.array_equal(adj.cpu().numpy(), wa)
    assert took < 5.0, took


def test_cech_60k_cells_equals_brute_force():
    """A clustered 60,000-cell scene (radii tripled, so rows are not empty):
    the grid and row offsets go through the multi-chunk scan (15 chunks of
    rows, more of grid cells)."""
    sc = workloads.gmm_scene(60_000, seed=3)
    rad = (3.0 * sc["radii"]).astype(np.float32)
    tree = AABBTree(DEV)
    tree.update(t(sc["points"]), t(rad))
    adj, off = tree.build_cech_complex()
    wa, wo = orc.cech(sc["points"], rad)
    assert wa.size > 60_000
    assert np.array_equal(off.cpu().numpy(), wo)
    assert np.array_equal(adj.cpu().numpy(), wa)


def test_raster_faces_bit_equal(scene4k, cech4k):
    sph = np.concatenate([scene4k["points"], scene4k["radii"][:, None]], 1).astype(np.float32)
    adj, off = cech4k
    got = raster_faces(t(sph), t(adj), t(off)).cpu().view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(got, orc.raster_faces(sph, adj, off, f16=True))


@pytest.mark.parametrize("W,H", [(64, 48), (37, 29)])
def test_binning_equal(scene4k, W, H):
    sph = np.concatenate([scene4k["points"], scene4k["radii"][:, None]], 1).astype(np.float32)
    cam = workloads.front_camera(W, H)
    off_g, prims_g, mask = bin_tiles(cameras_to_device([cam], None, DEV), t(sph), W, H)
    off_c, prims_c = orc.raster_bin(sph, pack_cameras([cam])[0])
    assert np.array_equal(off_g.cpu().numpy(), off_c)
    assert np.array_equal(prims_g.cpu().numpy(), prims_c)
    m = np.zeros(len(sph), bool)
    m[prims_c] = True
    assert np.array_equal(mask.cpu().numpy(), m)


def _rasterize_both(sc, adj, off, cam, depth_q=None, ray_gt=None, point_err=False, thr=1e-3):
    rz = Rasterizer(ARGS, DEV)
    rz.transmittance_threshold = thr
    p = {k: t(sc[k]).requires_grad_(True) for k in ("points", "radii", "density", "normals", "texel_sites",
                                                     "texel_rgb", "texel_height")}
    out = rz.forward(cam, None if depth_q is None else t(depth_q), p["points"], p["radii"], p["density"],
                     p["normals"], p["texel_sites"], p["texel_rgb"], p["texel_height"], t(adj), t(off),
                     None if ray_gt is None else t(ray_gt), point_err)
    sph = np.concatenate([sc["points"], sc["radii"][:, None]], 1).astype(np.float32)
    w = pack_cameras([cam])[0]
    S = orc.RasterScene(sc["points"], sc["radii"], sc["density"], sc["normals"], sc["texel_sites"],
                        sc["texel_rgb"], sc["texel_height"], adj, off, orc.raster_faces(sph, adj, off, True))
    o, pr = orc.raster_bin(sph, w)
    c = orc.raster_forward(S, w, o, pr, thr=thr, depth_q=depth_q, ray_gt=ray_gt, point_err=point_err)
    return rz, p, out, S, w, o, pr, c


def test_raster_forward_matches_cpu(scene_textured):
    sc = scene_textured
    adj, off = orc.cech(sc["points"], sc["radii"])
    cam = workloads.front_camera(72, 56)
    rng = np.random.default_rng(4)
    dq = np.tile(np.array([0.9, 0.5, 0.2], np.float32), (56, 72, 1))
    gt = rng.uniform(0, 1, (56, 72, 3)).astype(np.float32)
    _, _, out, _, _, _, _, c = _rasterize_both(sc, adj, off, cam, depth_q=dq, ray_gt=gt, point_err=True)
    color, opacity, ndist, normal, qd, err, contrib, perr, mask = out
    assert err is None and mask.dtype == torch.bool
    assert np.abs(color.detach().cpu().numpy() - c["color"]).max() <= TOL
    assert np.abs(opacity.detach().cpu().numpy() - (1 - np.exp(c["log_t"]))).max() <= TOL
    assert np.abs(ndist.detach().cpu().numpy() - c["ndist"]).max() <= TOL
    assert np.abs(normal.detach().cpu().numpy() - c["normal"]).max() <= TOL
    assert np.abs(qd.detach().cpu().numpy() - c["qdepth"]).max() <= TOL
    assert _rel(contrib.detach().cpu().numpy(), c["contrib"]) < 1e-5
    assert _rel(perr.detach().cpu().numpy(), c["perr"]) < 1e-5


def test_raster_early_stop_positions_equal(scene4k, cech4k):
    adj, off = cech4k
    cam = workloads.front_camera(64, 64)
    rz, p, out, S, w, o, pr, c = _rasterize_both(scene4k, adj, off, cam, thr=0.97)
    out[0].sum().backward()  # ensures the saved state is consistent
    assert np.abs(out[0].detach().cpu().numpy() - c["color"]).max() <= TOL


def test_raster_backward_matches_cpu(scene_textured):
    sc = scene_textured
    adj, off = orc.cech(sc["points"], sc["radii"])
    cam = workloads.front_camera(48, 48)
    rng = np.random.default_rng(9)
    dq = np.tile(np.array([0.7, 0.3], np.float32), (48, 48, 1))
    rz, p, out, S, w, o, pr, c = _rasterize_both(sc, adj, off, cam, depth_q=dq)
    color, opacity, ndist, normal, qd, err, contrib, perr, mask = out
    target = t(rng.uniform(0, 1, (48, 48, 3)).astype(np.float32))
    loss = ((color - target) ** 2).sum() + 0.3 * opacity.sum() + 0.2 * ndist.sum() + 0.1 * normal.sum() \
        + 0.05 * qd.sum() + 100.0 * contrib.sum()
    loss.backward()
    gc = (2 * (color - target)).detach().cpu().numpy()
    gl = (-0.3 * (1 - opacity)).detach().cpu().numpy()
    cg = orc.raster_backward(S, w, o, pr, c["early"], c["log_t"], gc, gl, g_ndist=np.full((48, 48), 0.2, np.float32),
                             g_normal=np.full((48, 48, 3), 0.1, np.float32), depth_q=dq,
                             g_qdepth=np.full(dq.shape, 0.05, np.float32),
                             g_contrib=np.full(len(sc["radii"]), 100.0, np.float32))
    assert _rel(p["points"].grad.cpu().numpy(), cg["spheres"][:, :3]) < 1e-3
    assert _rel(p["radii"].grad.cpu().numpy(), cg["spheres"][:, 3]) < 1e-3
    assert _rel(p["density"].grad.cpu().numpy(), cg["nsigmas"][:, 3]) < 1e-3
    assert _rel(p["normals"].grad.cpu().numpy(), cg["nsigmas"][:, :3]) < 1e-3
    assert _rel(p["texel_sites"].grad.cpu().numpy(), cg["texel_sites"]) < 1e-3
    assert _rel(p["texel_rgb"].grad.cpu().numpy(), cg["texel_rgb"]) < 1e-3
    assert _rel(p["texel_height"].grad.cpu().numpy(), cg["texel_height"]) < 1e-3


def test_raster_benchmark_matches_cpu(scene4k, cech4k):
    sc = scene4k
    adj, off = cech4k
    cam = workloads.front_camera(64, 40)
    rz = Rasterizer(ARGS, DEV, attr_dtype="half")
    wf = orc.walk_faces(sc["points"], sc["radii"], adj, off, f16=True)
    faces = t(wf.view(np.int16)).view(torch.float16)
    col = rz.benchmark(cam, t(sc["points"]), t(sc["radii"]), t(sc["density"]), t(sc["normals"]),
                       t(sc["texel_sites"]), t(sc["texel_rgb"]), t(sc["texel_height"]), t(adj), t(off), faces, 1e-2)
    h = lambda a: torch.from_numpy(a).half().float().numpy()
    sph = np.concatenate([sc["points"], sc["radii"][:, None]], 1).astype(np.float32)
    S = orc.RasterScene(sc["points"], sc["radii"], h(sc["density"]), h(sc["normals"]), h(sc["texel_sites"]),
                        h(sc["texel_rgb"]), h(sc["texel_height"]), adj, off, wf)
    w = pack_cameras([cam])[0]
    o, pr = orc.raster_bin(sph, w)
    c = orc.raster_plain(S, w, o, pr, thr=1e-2, vis=False)
    assert np.abs(col.cpu().numpy() - c["color"]).max() <= TOL


def test_raster_visualize_matches_cpu(scene4k, cech4k):
    sc = scene4k
    adj, off = cech4k
    cam = workloads.front_camera(48, 48)
    rz = Rasterizer(ARGS, DEV)
    opts = VisOptions(transmittance_threshold=1e-3, max_intersections=40, depth_quantile=0.5,
                      bkgd_color=(0.1, 0.2, 0.3))
    color, depth, normal, alpha, isect = rz.visualize(cam, t(sc["points"]), t(sc["radii"]), t(sc["density"]),
                                                      t(sc["normals"]), t(sc["texel_sites"]), t(sc["texel_rgb"]),
                                                      t(sc["texel_height"]), t(adj), t(off), opts)
    sph = np.concatenate([sc["points"], sc["radii"][:, None]], 1).astype(np.float32)
    S = orc.RasterScene(sc["points"], sc["radii"], sc["density"], sc["normals"], sc["texel_sites"],
                        sc["texel_rgb"], sc["texel_height"], adj, off, orc.raster_faces(sph, adj, off, True))
    w = pack_cameras([cam])[0]
    o, pr = orc.raster_bin(sph, w)
    c = orc.raster_plain(S, w, o, pr, thr=1e-3, vis=True, max_isect=40, depth_qv=0.5, bkgd=(0.1, 0.2, 0.3))
    for a, b in ((color, c["color"]), (depth, c["depth"]), (normal, c["normal"]), (alpha, c["alpha"])):
        assert np.abs(a.cpu().numpy() - b).max() <= TOL
    assert np.array_equal(isect.cpu().numpy(), c["isect"])


def test_rasterizer_error_conventions(scene4k, cech4k):
    sc = scene4k
    adj, off = cech4k
    cam = workloads.front_camera(16, 16)
    rz = Rasterizer(ARGS, DEV)
    args = [t(sc[k]) for k in ("points", "radii", "density", "normals", "texel_sites", "texel_rgb",
                               "texel_height")]
    with pytest.raises(ValueError):  # rasterize.py:1634-1637
        rz.forward(cam, None, *args, t(adj), t(off), None, True)
    with pytest.raises(ValueError):  # rasterize.py:1620-1621
        Rasterizer(ARGS, DEV, attr_dtype="half").forward(cam, None, *args, t(adj), t(off), None, False)
    with pytest.raises(ValueError):
        Rasterizer(ARGS, DEV, attr_dtype="double")
    surf = types.SimpleNamespace(**{**vars(ARGS), "render_objective": "surface"})
    gt = t(np.zeros((16, 16, 3), np.float32))
    p0 = args[0].clone().requires_grad_(True)
    out = Rasterizer(surf, DEV).forward(cam, None, p0, *args[1:], t(adj), t(off), gt, False)
    assert out[5] is not None
    with pytest.raises(NotImplementedError):  # rasterize.py:1232-1234
        out[0].sum().backward()


def test_mis_shaped_inputs_are_refused(scene_textured):
    """Per-pixel and per-cell inputs whose shapes do not match the camera or
    the scene raise ValueError instead of being read past their ends."""
    from powerfoam import raytrace as rt

    sc = scene_textured
    adj, off = orc.cech(sc["points"], sc["radii"])
    cam = workloads.front_camera(24, 16)
    with pytest.raises(ValueError, match="depth_quantiles"):
        _rasterize_both(sc, adj, off, cam, depth_q=np.full((16, 12, 2), 0.5, np.float32))
    with pytest.raises(ValueError, match="ray_gt"):
        _rasterize_both(sc, adj, off, cam, ray_gt=np.zeros((16, 23, 3), np.float32), point_err=True)
    short = dict(sc, texel_rgb=np.ascontiguousarray(sc["texel_rgb"][:-1]))
    with pytest.raises(ValueError, match="texel_rgb"):
        _rasterize_both(short, adj, off, cam)
    p = {k: t(sc[k]) for k in ("points", "radii", "density", "normals", "texel_sites", "texel_rgb", "texel_height")}
    with pytest.raises(ValueError, match="texel_height"):
        rt.WalkScene(p["points"], p["radii"], p["density"], p["normals"], p["texel_sites"], p["texel_rgb"],
                     p["texel_height"][:, :4], t(sc["adjacency"]), t(sc["offsets"]),
                     rt.walk_faces(p["points"], p["radii"], t(sc["adjacency"]).int(), t(sc["offsets"]).int()))
