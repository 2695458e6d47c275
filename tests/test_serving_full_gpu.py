"""Serving at the benchmark configurations' sizes.

* One 3,000,000-cell resident scene (Gaussian-mixture sites, SH degree 3),
  one request per resolution of {256^2, 512^2, 1024^2, 1920x1080}, each
  forward and forward+backward, all eight served in one mix at the default
  tile capacity (16,384) and record cap (128): every result equals the
  same request served alone and the plain per-request path (colour bit for
  bit, gradients within 1e-4 relative L2), and 512 sampled rays of the
  1080p request equal the CPU restatement bit for bit.
* Eight resident scenes with cells log-uniform in [1e5, 2e6] under the
  64 GB admission budget, a stream of forward and forward+backward requests
  whose reservations exceed the budget: admission holds some back, none is
  rejected, and every result equals the plain path's.
* A scene added while steps are in flight: the steps still read the old
  scene table, which is released only after they finish.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as orc
from powerfoam import raytrace as rt
from powerfoam import shading, workloads
from powerfoam.camera import cameras_to_device, pack_cameras
from powerfoam.geometry import morton_sort
from powerfoam.power import build_power_adjacency_gpu
from powerfoam.serving import PlainRenderer, RenderRequest, RenderServer, ResidentScene

pytestmark = pytest.mark.gpu
DEV = "cuda"
RES = [(256, 256), (512, 512), (1024, 1024), (1920, 1080)]
KEYS = ("spheres", "nsigmas", "texel_sites", "texel_rgb", "texel_height")


def _served_scene(n, seed):
    """A resident scene as the serving benchmark builds it: Morton-ordered
    cells, the GPU power-diagram adjacency, (E, 4) fp16 walk faces."""
    sc = workloads.gmm_scene(n, seed=seed)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    pts = t(sc["points"])
    perm = morton_sort(pts).long()
    pts, rad = pts[perm].contiguous(), t(sc["radii"])[perm].contiguous()
    adj, off = build_power_adjacency_gpu(pts, rad)
    faces = rt.walk_faces(pts, rad, adj, off)
    sc_perm = {k: (v[perm.cpu().numpy()] if isinstance(v, np.ndarray) else v) for k, v in sc.items()}
    return ResidentScene(pts, rad, t(sc_perm["density"]), t(sc_perm["normals"]), t(sc_perm["texel_sites"]),
                         t(sc_perm["texel_height"]), adj, off, faces, sh=t(sc_perm["sh"])), sc_perm, faces


def _req(scene, cam, mode, rng, rid):
    tgt = None
    if mode == "forward_backward":
        tgt = torch.from_numpy(rng.uniform(0, 1, (cam.height, cam.width, 3)).astype(np.float32))
    return RenderRequest(scene=scene, camera=cam, mode=mode, target=tgt, rid=rid)


def _drain(server, reqs, watch=None):
    for r in reqs:
        server.submit(r)
    while server.queue or server.active:
        server.step()
        if watch is not None:
            watch(server)
    server.poll(block=True)
    torch.cuda.synchronize()


def _rel_dev(a, b):
    return float(torch.linalg.norm((a - b).flatten()) / max(float(torch.linalg.norm(b.flatten())), 1e-30))


@pytest.fixture(scope="module")
def scene3m():
    return _served_scene(3_000_000, seed=21)


def test_3m_cells_every_resolution_alone_mixed_plain_and_oracle(scene3m):
    scene, sc, faces = scene3m
    rng = np.random.default_rng(31)
    cams = [workloads.sphere_cameras(rng, 1, w, h)[0] for w, h in RES]

    def make():
        g = np.random.default_rng(7)
        out = []
        for k, cam in enumerate(cams):
            for mode in ("forward", "forward_backward"):
                out.append(_req("big", cam, mode, g, len(out)))
        return out

    mixed = make()
    srv = RenderServer(DEV)  # default 64 GB budget, 16,384-tile steps, record cap 128
    srv.add_scene("big", scene)
    _drain(srv, mixed)
    assert int(srv.status.item()) & ~rt_cycle() == 0 and not srv.rejected
    plain = PlainRenderer({"big": scene})
    for r in make():
        alone = RenderServer(DEV)
        alone.add_scene("big", scene)
        _drain(alone, [r])
        q = next(x for x in mixed if x.rid == r.rid)
        p = plain.render(RenderRequest(scene="big", camera=r.camera, mode=r.mode, target=r.target, rid=r.rid))
        torch.cuda.synchronize()
        assert torch.equal(q.color, r.color) and torch.equal(q.log_t, r.log_t), (r.width, r.height, r.mode)
        assert torch.equal(q.color, p.color), (r.width, r.height, r.mode)
        if r.mode == "forward_backward":
            for k in KEYS:
                assert _rel_dev(q.grads[k], r.grads[k]) <= 1e-4, k
                assert _rel_dev(q.grads[k], p.grads[k]) <= 1e-4, k
        del alone
    # 512 sampled rays of the 1080p forward request against the restatement
    q = next(x for x in mixed if x.mode == "forward" and x.width == 1920)
    cam_dev = cameras_to_device([q.camera], None, DEV)
    rt.start_cells_device(scene.points, scene.radii, cam_dev)
    rgb = shading.sh_forward(scene.sh, 3, scene.walk.texel_sites, cam_dev)[0]
    words = cam_dev.cpu().numpy()
    fh = faces.cpu().view(torch.int16).numpy().view(np.uint16)
    cs = orc.Scene(scene.points.cpu().numpy(), scene.radii.cpu().numpy(), sc["density"], sc["normals"],
                   sc["texel_sites"], rgb.cpu().numpy(), sc["texel_height"], scene.walk.adjacency.cpu().numpy(),
                   scene.walk.offsets.cpu().numpy(), fh)
    rays = np.sort(np.random.default_rng(5).choice(1920 * 1080, 512, replace=False))
    c = orc.walk_forward(cs, words, rays=rays)
    assert c["status"] & ~rt_cycle() == 0
    assert np.array_equal(q.color.reshape(-1, 3).cpu().numpy()[rays], c["color"])


def rt_cycle():
    from powerfoam import _lib

    return _lib.PF_STATUS_CYCLE


def test_8_scenes_64gb_budget_admission_and_plain_equality():
    rng = np.random.default_rng(41)
    cells = np.exp(rng.uniform(np.log(1e5), np.log(2e6), 8)).astype(int)
    scenes = {f"s{k}": _served_scene(int(n), seed=100 + k)[0] for k, n in enumerate(cells)}
    # max_wait: held requests wait for memory however long the checks take
    srv = RenderServer(DEV, memory_budget=64 << 30, max_wait=1e9)
    for name, s in scenes.items():
        srv.add_scene(name, s)
    # requests until their reservations exceed the budget by a quarter
    reqs, demand = [], 0
    g = np.random.default_rng(42)
    while demand < 1.25 * srv.budget and len(reqs) < 400:
        name = f"s{int(g.integers(0, 8))}"
        w, h = RES[int(g.integers(0, 4))]
        cam = workloads.sphere_cameras(g, 1, w, h)[0]
        mode = "forward_backward" if g.uniform() < 0.5 else "forward"
        r = _req(name, cam, mode, g, len(reqs))
        reqs.append(r)
        demand += srv._request_bytes(r)
    plain = PlainRenderer(scenes)
    checked = []

    def check(r):  # each finished request against the plain path, then dropped
        p = plain.render(RenderRequest(scene=r.scene, camera=r.camera, mode=r.mode, target=r.target, rid=r.rid))
        ok = torch.equal(r.color, p.color)
        if r.mode == "forward_backward":
            ok = ok and all(_rel_dev(r.grads[k], p.grads[k]) <= 1e-4 for k in KEYS)
        checked.append((r.rid, ok))
        r.grads = None
        r.color = r.log_t = None

    srv.on_finish = check
    held = []
    _drain(srv, reqs, watch=lambda s: held.append(len(s.queue)))
    assert not srv.rejected
    assert max(held) > 0  # the budget held requests back while others ran
    assert len(checked) == len(reqs) and all(ok for _, ok in checked), [rid for rid, ok in checked if not ok]
    assert int(srv.status.item()) & ~rt_cycle() == 0


def test_scene_added_while_steps_are_in_flight(scene4k):
    sc = scene4k
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    rng = np.random.default_rng(12)
    sh = t((0.1 * rng.standard_normal((len(sc["radii"]), 16, 3))).astype(np.float32))
    pts, rad = t(sc["points"]), t(sc["radii"])
    adj, off = t(sc["adjacency"]), t(sc["offsets"])
    faces = rt.walk_faces(pts, rad, adj, off)
    a = ResidentScene(pts, rad, t(sc["density"]), t(sc["normals"]), t(sc["texel_sites"]), t(sc["texel_height"]),
                      adj, off, faces, sh=sh)
    b = ResidentScene(pts, rad, t(sc["density"]), t(sc["normals"]), t(sc["texel_sites"]), t(sc["texel_height"]),
                      adj, off, faces, sh=(sh * 0.5).contiguous())
    g = np.random.default_rng(3)
    first = [_req("a", workloads.sphere_cameras(g, 1, 96, 96)[0], "forward", g, k) for k in range(6)]
    srv = RenderServer(DEV, tile_capacity=64, quantum=16)
    srv.add_scene("a", a)
    for r in first:
        srv.submit(r)
    srv.step()
    srv.step()  # two steps in flight on the server's streams
    srv.add_scene("b", b)
    second = [_req("b", workloads.sphere_cameras(g, 1, 96, 96)[0], "forward", g, 10 + k) for k in range(3)]
    _drain(srv, second)
    plain = PlainRenderer({"a": a, "b": b})
    for r in first + second:
        assert r.done is not None
        p = plain.render(RenderRequest(scene=r.scene, camera=r.camera, mode="forward", rid=r.rid))
        assert torch.equal(r.color, p.color), r.rid
