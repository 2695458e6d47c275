"""Scene glue on the GPU: forward equals the explicit composition of the
colour pre-pass and the rasterizer, every parameter receives a finite
gradient, a checkpoint written and read back renders bit-identically, the
walk render agrees with the raster render of the same scene, and the
launcher runs an entry point on torch's stream."""
import types

import numpy as np
import pytest
import torch

from powerfoam import workloads
from powerfoam.adjacency import build_power_adjacency
from powerfoam.scene import PowerfoamScene
from powerfoam.warp_interop import launch_kernel_from_torch

pytestmark = pytest.mark.gpu
DEV = "cuda"
ARGS = types.SimpleNamespace(num_texel_sites=8, sv_dof=3, render_objective="volume", is_pinhole=True)


@pytest.fixture(scope="module")
def scene():
    s = PowerfoamScene(ARGS).initialize_random(512, DEV, seed=3)
    with torch.no_grad():
        s.density.fill_(0.3)  # softplus(0.3, beta=100) ~ 0.3: visible cells
    return s


def test_forward_is_prepass_then_rasterizer(scene):
    cam = workloads.front_camera(64, 48)
    out = scene.forward(cam)
    sites, h = scene.world_texels()
    axis, rgb, temp = scene.get_att_sv()
    tex = scene.sv.forward(sites.reshape(-1, 3).detach(), cam, axis, rgb, temp).view(512, 8, 3)
    ref = scene.rasterizer.forward(cam, None, scene.points, scene.get_radii(), scene.get_density(),
                                   scene.get_normals(), sites, tex, h, scene.adjacency, scene.adjacency_offsets,
                                   None, False)
    assert len(out) == 9
    assert torch.equal(out[0], ref[0]) and torch.equal(out[1], ref[1])
    assert float(out[1].detach().max()) > 0.05  # something is visible


def test_gradients_reach_every_parameter(scene):
    cam = workloads.front_camera(64, 48)
    scene.zero_grad()
    color, opacity = scene.forward(cam)[:2]
    (color.square().sum() + opacity.sum()).backward()
    for name in ("points", "radii", "density", "quaternions", "texel_sites", "texel_sv_axis", "texel_sv_rgb",
                 "texel_height"):
        g = getattr(scene, name).grad
        assert g is not None and torch.isfinite(g).all(), name
        assert float(g.abs().sum()) > 0, name


def test_checkpoint_renders_identically(scene, tmp_path):
    cam = workloads.front_camera(40, 40)
    path = tmp_path / "model.pt"
    scene.save_pt(path)
    s2 = PowerfoamScene(ARGS).load_pt(path, device=DEV)
    with torch.no_grad():
        a = scene.forward(cam)[0]
        b = s2.forward(cam)[0]
    assert torch.equal(a, b)


def test_walk_render_agrees_with_raster(scene):
    cam = workloads.front_camera(48, 48)
    pts, rad = scene.points.detach().cpu(), scene.get_radii().detach().cpu()
    adj = build_power_adjacency(pts, rad)
    with torch.no_grad():
        walk = scene.render_walk(cam, adj)[0][0]
        rast = scene.forward(cam)[0]
    assert float((walk - rast).abs().mean()) < 5e-3


def test_launcher_runs_entry_point(scene):
    c = scene.points.detach().contiguous()
    r = scene.get_radii().detach().float().contiguous()
    ws = torch.empty(8, dtype=torch.int32, device=DEV)
    bounds = torch.empty(8, dtype=torch.float32, device=DEV)
    launch_kernel_from_torch(DEV, "pf_cech_bounds", c.shape[0], [c, r, c.shape[0], ws, bounds], 256)
    b = bounds.cpu().numpy()
    # bounds of the sphere centres recovered from their boxes, and the largest radius
    np.testing.assert_allclose(b[0:3], c.min(0).values.cpu().numpy(), rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(b[3:6], c.max(0).values.cpu().numpy(), rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(b[6], float(r.max()), rtol=1e-5)
    np.testing.assert_allclose(b[7], float(r.mean()), rtol=1e-4)  # mean radius
