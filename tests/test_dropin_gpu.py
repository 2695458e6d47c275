"""The literal drop-ins, called the way the reference's own callers call them.

* ``RayTracer.benchmark`` exactly as benchmark.py:263-372 drives it: a
  scene with half attributes (benchmark.py:267), the power adjacency of
  build_power_adjacency (:290-292), reference-layout (E, 4) fp16
  ``adjacency_diff`` built by the harness's own torch code (:295-302), an
  int start cell from the harness's argmin (:320-324) and threshold 1e-2
  (:371).  Bar: the image equals the CPU restatement's bit for bit.
* ``PowerfoamScene.forward`` (scene.py:314-353) of a scene read back from a
  ``save_pt`` checkpoint with ``load_pt`` (scene.py:772-821): its 9 outputs
  against the restatement's spherical-Voronoi pre-pass and rasterizer on
  the same activations (colour and side outputs within 1e-4, per-cell
  sums within 1e-5 relative).
"""
import types

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from powerfoam import workloads
from powerfoam.adjacency import build_power_adjacency
from powerfoam.camera import pack_cameras
from powerfoam.color_fn import camera_frame
from powerfoam.raytrace import RayTracer
from powerfoam.scene import PowerfoamScene

pytestmark = pytest.mark.gpu
DEV = "cuda"
ARGS = types.SimpleNamespace(num_texel_sites=8, sv_dof=3, render_objective="volume", is_pinhole=True,
                             disable_coop_prim_load=False, disable_coop_adj_load=False)


def _harness_benchmark(sc, cam):
    """benchmark.py:290-372 with render_type 'raytrace', on a half scene."""
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    points = t(sc["points"])                 # model.points stays f32
    radii = t(sc["radii"]).half()            # activations of a half model
    density = t(sc["density"]).half()
    normals = t(sc["normals"]).half()
    texel_sites = t(sc["texel_sites"])       # points + offsets: f32
    texel_rgb = t(sc["texel_rgb"]).half()
    texel_height = t(sc["texel_height"]).half()
    adjacency, adjacency_offsets = build_power_adjacency(points.cpu(), radii.float().cpu())
    adjacency, adjacency_offsets = adjacency.to(DEV), adjacency_offsets.to(DEV)
    num_adjs = adjacency_offsets.diff()
    pm = 0.5 * (points.norm(dim=-1) ** 2 - radii ** 2)
    self_points = points.repeat_interleave(num_adjs, dim=0)
    adjacency_diff = points[adjacency] - self_points
    pm_diff = pm[adjacency] - pm.repeat_interleave(num_adjs, dim=0)
    adjacency_diff = torch.cat([adjacency_diff, pm_diff[:, None]], dim=-1)
    adjacency_diff = adjacency_diff.to(torch.float16)
    camera_eye = cam.eye.to(DEV)
    dists = torch.linalg.norm(points - camera_eye[None, :], dim=-1)
    start_point_idx = int(torch.argmin(dists ** 2 - radii ** 2))
    raytracer = RayTracer(ARGS, DEV, attr_dtype="half")
    rgb = raytracer.benchmark(cam, start_point_idx, points, radii, density, normals, texel_sites, texel_rgb,
                              texel_height, adjacency, adjacency_offsets, adjacency_diff, 1e-2)
    torch.cuda.synchronize()
    inputs = dict(points=points, radii=radii, density=density, normals=normals, texel_sites=texel_sites,
                  texel_rgb=texel_rgb, texel_height=texel_height)
    return rgb, start_point_idx, inputs, adjacency, adjacency_offsets, adjacency_diff


@pytest.mark.parametrize("which", ["uniform_4k_128", "textured_512_96"])
def test_raytracer_benchmark_as_the_reference_harness_calls_it(scene_textured, which):
    if which == "uniform_4k_128":
        sc = workloads.uniform_scene(4096, seed=0)   # the 4,096-cell CPU-path configuration
        cam = workloads.front_camera(128, 128)
    else:
        sc, cam = scene_textured, workloads.front_camera(96, 80, hfov_deg=40.0)
    rgb, start, inp, adj, off, diff = _harness_benchmark(sc, cam)
    assert rgb.shape == (cam.height, cam.width, 3) and rgb.dtype == torch.float32
    # the restatement on the same values: attributes as the kernel reads them
    # (f16 storage widened exactly), the harness's own f16 faces
    h = lambda k: inp[k].half().float().cpu().numpy()
    faces = diff.cpu().view(torch.int16).numpy().view(np.uint16)
    cs = orc.Scene(inp["points"].cpu().numpy(), inp["radii"].float().cpu().numpy(), h("density"), h("normals"),
                   h("texel_sites"), h("texel_rgb"), h("texel_height"), adj.int().cpu().numpy(),
                   off.int().cpu().numpy(), faces)
    c = orc.walk_forward(cs, pack_cameras([cam], [start]), thr=1e-2)
    assert c["status"] == 0
    assert np.array_equal(rgb.cpu().numpy(), c["color"][0])
    assert float(rgb.max()) > 0.05  # the scene is visible


def test_scene_forward_from_checkpoint_matches_oracle(tmp_path):
    src = PowerfoamScene(ARGS).initialize_random(768, DEV, seed=11)
    with torch.no_grad():
        src.density.fill_(0.3)
        src.texel_height.uniform_(-0.3, 0.3)
    path = tmp_path / "model.pt"
    src.save_pt(path)
    s = PowerfoamScene(ARGS).load_pt(path, device=DEV)
    W = H = 48
    cam = workloads.front_camera(W, H, hfov_deg=50.0)
    rng = np.random.default_rng(2)
    dq = np.tile(np.array([0.8, 0.4], np.float32), (H, W, 1))
    gt = rng.uniform(0, 1, (H, W, 3)).astype(np.float32)
    with torch.no_grad():
        out = s.forward(cam, torch.from_numpy(dq).to(DEV), torch.from_numpy(gt).to(DEV), True)
        color, opacity, ndist, normal, qd, err, contrib, perr, mask = out
        sites, heights = s.world_texels()
        axis, vals, temps = s.get_att_sv()
        radii, density, normals = s.get_radii(), s.get_density(), s.get_normals()
    n = lambda x: x.detach().float().cpu().numpy()
    eye, fwd, cut = camera_frame(cam)
    rgb, _ = orc.sv_forward(n(sites).reshape(-1, 3), eye.numpy(), fwd.numpy(), cut, n(axis), n(vals), n(temps))
    pts, rad = n(s.points), n(radii)
    sph = np.concatenate([pts, rad[:, None]], 1).astype(np.float32)
    adj, off = s.adjacency.cpu().numpy(), s.adjacency_offsets.cpu().numpy()
    S = orc.RasterScene(pts, rad, n(density), n(normals), n(sites), rgb.reshape(-1, 8, 3), n(heights), adj, off,
                        orc.raster_faces(sph, adj, off, f16=True))
    w = pack_cameras([cam])[0]
    o, pr = orc.raster_bin(sph, w)
    c = orc.raster_forward(S, w, o, pr, thr=1e-3, depth_q=dq, ray_gt=gt, point_err=True)
    assert float(opacity.max()) > 0.05
    tol = 1e-4
    assert np.abs(n(color) - c["color"]).max() <= tol
    assert np.abs(n(opacity) - (1 - np.exp(c["log_t"]))).max() <= tol
    assert np.abs(n(ndist) - c["ndist"]).max() <= tol
    assert np.abs(n(normal) - c["normal"]).max() <= tol
    assert np.abs(n(qd) - c["qdepth"]).max() <= tol
    rel = lambda a, b: np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
    assert rel(n(contrib), c["contrib"]) < 1e-5 and rel(n(perr), c["perr"]) < 1e-5
    assert err is None
    visible = np.zeros(len(rad), bool)
    visible[pr] = True
    assert np.array_equal(mask.cpu().numpy(), visible)
