"""Serving a trained scene coloured by spherical-Voronoi lobes.

A PowerfoamScene written with ``save_pt`` and read back with ``load_pt``
(scene.py:772-821) becomes a ResidentScene (``from_powerfoam``) on a GPU
power-diagram adjacency.  Requests from two cameras, forward and
forward+backward, are served in one mix:

* every served image equals ``PowerfoamScene.render_walk`` on the same
  adjacency bit for bit (colour and alpha), and the CPU restatement (its
  spherical-Voronoi pre-pass, then its walk) on every pixel;
* the forward+backward request's lobe gradients (the server's
  ``finish_gradients``) equal autograd through ``SphericalVoronoi.forward``
  and ``RayTracer.forward`` on the same loss within 1e-4 relative L2, and
  its scene gradients equal the plain per-request path's.
"""
import types

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from powerfoam import workloads
from powerfoam.camera import cameras_to_device
from powerfoam.color_fn import camera_frame
from powerfoam.power import build_power_adjacency_gpu
from powerfoam.raytrace import start_cells_device, walk_faces
from powerfoam.scene import PowerfoamScene
from powerfoam.serving import PlainRenderer, RenderRequest, RenderServer, ResidentScene

pytestmark = pytest.mark.gpu
DEV = "cuda"
ARGS = types.SimpleNamespace(num_texel_sites=8, sv_dof=3, render_objective="volume", is_pinhole=True,
                             disable_coop_prim_load=False, disable_coop_adj_load=False)
KEYS = ("spheres", "nsigmas", "texel_sites", "texel_rgb", "texel_height")


def _rel(a, b):
    return float(torch.linalg.norm((a - b).flatten()) / max(float(torch.linalg.norm(b.flatten())), 1e-30))


@pytest.fixture(scope="module")
def checkpoint_scene(tmp_path_factory):
    src = PowerfoamScene(ARGS).initialize_random(2048, DEV, seed=5)
    with torch.no_grad():
        src.density.fill_(0.4)
        src.texel_height.uniform_(-0.3, 0.3)
    path = tmp_path_factory.mktemp("ckpt") / "model.pt"
    src.save_pt(path)
    s = PowerfoamScene(ARGS).load_pt(path, device=DEV)
    with torch.no_grad():
        adj, off = build_power_adjacency_gpu(s.points.detach(), s.get_radii().detach().float())
    return s, (adj, off)


def _cams():
    from powerfoam.camera import look_at_camera

    return [workloads.front_camera(64, 48, hfov_deg=50.0),
            look_at_camera([2.2, -1.6, 1.0], [0.0, 0.0, 0.0], [0.0, 0.0, 1.0], 40, 56, 55.0)]


def test_served_lobe_scene_equals_render_walk_and_oracle(checkpoint_scene):
    s, padj = checkpoint_scene
    rs = ResidentScene.from_powerfoam(s, padj)
    cams = _cams()
    g = np.random.default_rng(3)
    targets = [torch.from_numpy(g.uniform(0, 1, (c.height, c.width, 3)).astype(np.float32)) for c in cams]
    reqs = [RenderRequest(scene="ckpt", camera=cams[0], mode="forward", rid=0),
            RenderRequest(scene="ckpt", camera=cams[1], mode="forward", rid=1),
            RenderRequest(scene="ckpt", camera=cams[0], mode="forward_backward", target=targets[0], rid=2)]
    srv = RenderServer(DEV, tile_capacity=32, quantum=8)
    srv.add_scene("ckpt", rs)
    for r in reqs:
        srv.submit(r)
    while srv.queue or srv.active:
        srv.step()
    srv.poll(block=True)
    torch.cuda.synchronize()
    assert not srv.rejected and int(srv.status.item()) == 0
    lobe_grads = srv.finish_gradients(reqs[2])

    # against PowerfoamScene.render_walk on the same checkpoint and adjacency
    for r in reqs:
        with torch.no_grad():
            color, alpha, _ = s.render_walk(r.camera, padj)
        torch.cuda.synchronize()
        assert torch.equal(r.color, color[0]), r.rid
        assert torch.equal(1.0 - torch.exp(r.log_t), alpha[0]), r.rid
        assert float(color.max()) > 0.05  # the scene is visible

    # against the restatement: its lobe pre-pass, then its walk, every pixel
    n = lambda t: t.detach().float().cpu().numpy()
    sites = rs.walk.texel_sites
    fh = walk_faces(rs.points, rs.radii, *padj).cpu().view(torch.int16).numpy().view(np.uint16)
    for r in reqs[:2]:
        eye, fwd, cut = camera_frame(r.camera)
        rgb, _ = orc.sv_forward(n(sites).reshape(-1, 3), eye.numpy(), fwd.numpy(), cut, *(n(t) for t in rs.sv))
        cam_dev = cameras_to_device([r.camera], None, DEV)
        start_cells_device(rs.points, rs.radii, cam_dev)
        cs = orc.Scene(n(rs.points), n(rs.radii), n(rs.walk.nsigmas[:, 3]), n(rs.walk.nsigmas[:, :3]), n(sites),
                       rgb.reshape(-1, 8, 3), n(rs.walk.texel_height), n(padj[0]).astype(np.int32),
                       n(padj[1]).astype(np.int32), fh)
        c = orc.walk_forward(cs, cam_dev.cpu().numpy())
        assert c["status"] == 0
        assert np.array_equal(n(r.color), c["color"][0]), r.rid

    # lobe gradients against autograd through the reference-shaped modules
    axis, vals, temps = (t.detach().clone().requires_grad_(True) for t in s.get_att_sv())
    with torch.no_grad():
        w_sites, heights = s.world_texels()
        radii, density, normals = s.get_radii(), s.get_density(), s.get_normals()
    rgb = s.sv.forward(w_sites.reshape(-1, 3).float(), cams[0], axis, vals, temps).view(-1, 8, 3)
    adj, off = padj
    faces = walk_faces(s.points.detach(), radii.float(), adj, off)
    color, _, _ = s.raytracer.forward(cams[0], None, s.points.detach(), radii, density, normals, w_sites, rgb, heights,
                                      adj, off, faces, 1e-3)
    loss = ((color[0] - targets[0].to(DEV)) ** 2).sum()
    loss.backward()
    for got, want, name in zip(lobe_grads, (axis.grad, vals.grad, temps.grad), ("axes", "values", "temps")):
        assert float(want.abs().max()) > 0, name
        assert _rel(got, want) <= 1e-4, name

    # scene gradients against the plain per-request path
    p = PlainRenderer({"ckpt": rs}).render(RenderRequest(scene="ckpt", camera=cams[0], mode="forward_backward",
                                                         target=targets[0], rid=9))
    torch.cuda.synchronize()
    assert torch.equal(p.color, reqs[2].color)
    for k in KEYS:
        assert _rel(reqs[2].grads[k], p.grads[k]) <= 1e-4, k
    for got, want in zip(lobe_grads, p.colour_grads):
        assert _rel(got, want) <= 1e-4


def test_lobe_scene_checks_its_arrays(checkpoint_scene):
    s, padj = checkpoint_scene
    rs = ResidentScene.from_powerfoam(s, padj)
    axes, vals, temps = rs.sv
    with pytest.raises(ValueError, match="lobes"):
        ResidentScene(rs.points, rs.radii, rs.walk.nsigmas[:, 3].contiguous(), rs.walk.nsigmas[:, :3].contiguous(),
                      rs.walk.texel_sites, rs.walk.texel_height, padj[0], padj[1], rs.walk.faces,
                      sv=(axes, vals, temps[:, :-8]))
