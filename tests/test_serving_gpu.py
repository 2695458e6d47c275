"""Serving: each request's result is the same alone and in a mixed batch,
the same as the plain per-request path, and the same whether steps run as
captured graph replays or eager launches; admission rejects what cannot fit."""
import numpy as np
import pytest
import torch

from oracle import oracle as orc
from powerfoam import raytrace as rt
from powerfoam import workloads
from powerfoam.serving import PlainRenderer, RenderRequest, RenderServer, ResidentScene

DEV = "cuda"


def t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


@pytest.fixture(scope="module")
def served(scene4k):
    sc = scene4k
    rng = np.random.default_rng(12)
    sh = (0.1 * rng.standard_normal((len(sc["radii"]), 16, 3))).astype(np.float32)
    pts, rad = t(sc["points"]), t(sc["radii"])
    adj, off = t(sc["adjacency"]), t(sc["offsets"])
    faces = rt.walk_faces(pts, rad, adj, off)
    scene = ResidentScene(pts, rad, t(sc["density"]), t(sc["normals"]), t(sc["texel_sites"]), t(sc["texel_height"]),
                          adj, off, faces, sh=t(sh))
    return scene


def _requests(seed, n, with_bwd=True):
    rng = np.random.default_rng(seed)
    res = [(64, 64), (96, 80), (128, 128), (40, 24)]
    reqs = []
    for k in range(n):
        w, h = res[k % len(res)]
        cam = workloads.sphere_cameras(rng, 1, w, h)[0]
        mode = "forward_backward" if (with_bwd and k % 3 == 2) else "forward"
        tgt = torch.from_numpy(rng.uniform(0, 1, (h, w, 3)).astype(np.float32)) if mode != "forward" else None
        reqs.append(RenderRequest(scene="s", camera=cam, mode=mode, target=tgt, rid=k))
    return reqs


pytestmark = pytest.mark.gpu


def _drain(server, reqs):
    for r in reqs:
        server.submit(r)
    while server.queue or server.active:
        server.step()
    server.poll(block=True)
    torch.cuda.synchronize()


@pytest.mark.parametrize("use_graph,backward", [(False, "grouped"), (True, "grouped"), (True, "recorded")])
def test_alone_equals_mixed(served, use_graph, backward):
    alone = _requests(3, 6)
    for r in alone:
        srv = RenderServer(DEV, tile_capacity=64, quantum=16, use_graph=use_graph, backward=backward)
        srv.add_scene("s", served)
        _drain(srv, [r])
    mixed = _requests(3, 6)
    srv = RenderServer(DEV, tile_capacity=64, quantum=16, use_graph=use_graph, backward=backward)
    srv.add_scene("s", served)
    _drain(srv, mixed)
    assert int(srv.status.item()) == 0
    for a, b in zip(alone, mixed):
        assert b.done is not None and not b.rejected
        assert torch.equal(a.color, b.color) and torch.equal(a.log_t, b.log_t)
        if a.mode == "forward_backward":
            for k in ("spheres", "nsigmas", "texel_rgb"):
                x, y = a.grads[k].cpu().numpy(), b.grads[k].cpu().numpy()
                assert np.linalg.norm(x - y) <= 1e-4 * max(np.linalg.norm(x), 1e-30)


def test_served_equals_plain_path(served):
    reqs = _requests(5, 4)
    srv = RenderServer(DEV, tile_capacity=128, quantum=32)
    srv.add_scene("s", served)
    _drain(srv, reqs)
    plain = PlainRenderer({"s": served})
    for r in _requests(5, 4):
        plain.render(r)
        q = next(x for x in reqs if x.rid == r.rid)
        assert torch.equal(q.color, r.color)
        if r.mode == "forward_backward":
            x, y = q.grads["spheres"].cpu().numpy(), r.grads["spheres"].cpu().numpy()
            assert np.linalg.norm(x - y) <= 1e-4 * np.linalg.norm(y)


@pytest.mark.parametrize("record_cap,backward", [(0, "recorded"), (8, "recorded"), (8, "grouped"), (128, "grouped")])
def test_served_backward_with_and_without_record(served, record_cap, backward):
    """Forward+backward tiles differentiated by replay (record_cap 0), from a
    small record read tile by tile (8: rays with more contributing segments
    are replayed in the same wave), or grouped by cell over the step (rays
    past the cap replayed in the same launch) give the plain path's
    gradients."""
    reqs = [r for r in _requests(11, 9) if r.mode == "forward_backward"]
    srv = RenderServer(DEV, tile_capacity=256, quantum=64, record_cap=record_cap, backward=backward)
    srv.add_scene("s", served)
    _drain(srv, reqs)
    assert int(srv.status.item()) == 0
    assert srv.backward == (backward if record_cap else "recorded")
    plain = PlainRenderer({"s": served})
    for r in [r for r in _requests(11, 9) if r.mode == "forward_backward"]:
        plain.render(r)
        q = next(x for x in reqs if x.rid == r.rid)
        assert torch.equal(q.color, r.color)
        for k in ("spheres", "nsigmas", "texel_sites", "texel_rgb", "texel_height"):
            x, y = q.grads[k].cpu().numpy(), r.grads[k].cpu().numpy()
            assert np.linalg.norm(x - y) <= 1e-4 * max(np.linalg.norm(y), 1e-30), k
    if record_cap == 8:
        # the record held some rays and not others: their counts, from a
        # recording forward of the same requests
        from powerfoam import _lib, shading
        from powerfoam.camera import cameras_to_device
        from powerfoam.serving import _SceneView

        over = held = 0
        for r in reqs:
            cam = cameras_to_device([r.camera], None, DEV)
            rt.start_cells_device(served.points, served.radii, cam)
            rgb = shading.sh_forward(served.sh, 3, served.walk.texel_sites, cam)
            c = _lib.PfScene.from_buffer_copy(served.walk.c)
            c.texel_rgb = rgb.data_ptr()
            f = rt.launch_walk_forward(_SceneView(c, served.num_cells, served.K, rgb), cam, r.width, r.height,
                                       record_cap=512)
            counts = rt.record_counts(f["record"], 1, r.width, r.height)
            over += int((counts > record_cap).sum())
            held += int(((counts > 0) & (counts <= record_cap)).sum())
        assert over > 0 and held > 0


def test_graph_replays_match_eager_repeatedly(served):
    outs = []
    for use_graph in (False, True, True):
        reqs = _requests(8, 5, with_bwd=False)
        srv = RenderServer(DEV, tile_capacity=32, quantum=8, use_graph=use_graph)
        srv.add_scene("s", served)
        _drain(srv, reqs)
        outs.append([r.color.clone() for r in reqs])
        torch.cuda.synchronize()
    for a, b in zip(outs[0], outs[1]):
        assert torch.equal(a, b)
    for a, b in zip(outs[0], outs[2]):
        assert torch.equal(a, b)


def test_memory_admission_rejects_what_cannot_fit(served):
    prepass = served.num_cells * served.K * 12  # the request's own texel colours
    srv = RenderServer(DEV, memory_budget=served.nbytes() + 64 * 64 * 16 + prepass + 1024)
    srv.add_scene("s", served)
    small = _requests(9, 1, with_bwd=False)[0]
    big = RenderRequest(scene="s", camera=workloads.front_camera(1024, 1024))
    assert srv.submit(small)
    assert not srv.submit(big) and big.rejected
    _drain(srv, [])
    assert small.done is not None


def test_request_waits_for_memory_held_by_a_finished_step(served):
    """Budget for one request at a time: the second, submitted directly
    (arrival left at 0), waits for the first to release its memory and is
    then served; max_wait counts from submit."""
    prepass = served.num_cells * served.K * 12
    srv = RenderServer(DEV, memory_budget=served.nbytes() + 64 * 64 * 16 + prepass + 1024, max_wait=30.0)
    srv.add_scene("s", served)
    a, b = (RenderRequest(scene="s", camera=workloads.front_camera(64, 64), rid=k) for k in range(2))
    assert srv.submit(a) and srv.submit(b)
    _drain(srv, [])
    assert not srv.rejected and a.done is not None and b.done is not None
    assert b.admitted >= a.done
    assert torch.equal(a.color, b.color)
    assert srv.reserved == served.nbytes()


def test_bad_scenes_and_requests_are_refused(served):
    """A second scene under a resident name, an unknown scene or mode, or a
    forward+backward request without a matching target: ValueError at
    add_scene / submit, before anything reaches the request table."""
    with pytest.raises(ValueError):
        RenderServer(DEV, tile_capacity=0)
    with pytest.raises(ValueError):
        RenderServer(DEV, tile_capacity=(1 << 26) + 1)
    srv = RenderServer(DEV)
    srv.add_scene("s", served)
    reserved = srv.reserved
    with pytest.raises(ValueError):
        srv.add_scene("s", served)
    assert srv.reserved == reserved and srv.scene_index == {"s": 0}
    cam = workloads.front_camera(32, 24)
    with pytest.raises(ValueError):
        srv.submit(RenderRequest(scene="nope", camera=cam))
    with pytest.raises(ValueError):
        srv.submit(RenderRequest(scene="s", camera=cam, mode="backward"))
    with pytest.raises(ValueError):
        srv.submit(RenderRequest(scene="s", camera=cam, mode="forward_backward"))
    with pytest.raises(ValueError):
        srv.submit(RenderRequest(scene="s", camera=cam, mode="forward_backward", target=torch.zeros(32, 24, 3)))
    assert not srv.queue
    ok = RenderRequest(scene="s", camera=cam, mode="forward_backward", target=torch.zeros(24, 32, 3))
    _drain(srv, [ok])
    assert ok.done is not None and int(srv.status.item()) & ~4 == 0


def test_prepass_colour_equals_walk_colour(served):
    """Texel colours from the per-request pre-pass and SH evaluated inside
    the walk are the same numbers, so the images are identical."""
    outs = []
    for colour in ("prepass", "walk"):
        reqs = _requests(21, 4)
        srv = RenderServer(DEV, tile_capacity=64, quantum=16, colour=colour)
        srv.add_scene("s", served)
        _drain(srv, reqs)
        outs.append(reqs)
    for a, b in zip(*outs):
        assert torch.equal(a.color, b.color) and torch.equal(a.log_t, b.log_t)
        if a.mode == "forward_backward":
            x, y = a.grads["texel_rgb"].cpu().numpy(), b.grads["texel_rgb"].cpu().numpy()
            assert np.linalg.norm(x - y) <= 1e-4 * max(np.linalg.norm(x), 1e-30)


def test_multi_scene_admission_mix_equals_plain():
    """Several resident scenes, a mix of forward and forward+backward
    requests of several resolutions, and a memory budget that admits only
    part of the stream at once: every result equals the plain per-request
    path's (colour bit for bit, gradients within 1e-4 relative L2), and the
    waiting requests are admitted as memory frees up, none rejected."""
    import bench

    scenes = {f"s{k}": bench.build_served_scene(n, 40 + k) for k, n in enumerate((3000, 9000, 20000))}
    rng = np.random.default_rng(77)
    res = [(64, 64), (96, 80), (128, 128), (40, 24)]

    def make():
        reqs = []
        for k in range(18):
            w, h = res[k % 4]
            name = f"s{k % 3}"
            bwd = k % 3 == 1
            tgt = torch.from_numpy(rng.uniform(0, 1, (h, w, 3)).astype(np.float32)) if bwd else None
            reqs.append(RenderRequest(scene=name, camera=workloads.sphere_cameras(rng, 1, w, h)[0], rid=k,
                                      mode="forward_backward" if bwd else "forward", target=tgt))
        return reqs

    reqs = make()
    plain_reqs = [RenderRequest(scene=r.scene, camera=r.camera, rid=r.rid, mode=r.mode, target=r.target)
                  for r in reqs]
    srv = RenderServer(DEV, tile_capacity=96, quantum=24, max_wait=60.0)
    for name, sc in scenes.items():
        srv.add_scene(name, sc)
    # room for about three of the largest requests' buffers besides the scenes
    srv.budget = sum(s.nbytes() for s in scenes.values()) + 3 * max(srv._request_bytes(r) for r in reqs)
    for r in reqs:
        assert srv.submit(r)
    held = False
    while srv.queue or srv.active or srv.inflight:
        srv.step()
        held = held or (bool(srv.queue) and bool(srv.active))
        srv.poll()
    srv.poll(block=True)
    torch.cuda.synchronize()
    assert not srv.rejected and all(r.done is not None for r in reqs)
    assert held  # the budget kept part of the stream waiting while others ran
    plain = PlainRenderer(scenes)
    for q in plain_reqs:
        plain.render(q)
    torch.cuda.synchronize()
    for a, b in zip(reqs, plain_reqs):
        assert torch.equal(a.color, b.color), a.rid
        if a.mode == "forward_backward":
            for k in ("spheres", "nsigmas", "texel_sites", "texel_rgb", "texel_height"):
                x, y = a.grads[k].cpu().numpy(), b.grads[k].cpu().numpy()
                assert np.linalg.norm(x - y) <= 1e-4 * max(np.linalg.norm(y), 1e-30), (a.rid, k)


def test_grouped_step_limits_and_entry_budget(served):
    """The grouped serving step with one forward+backward request's cells per
    step (max_buckets = N), eight forward+backward tiles per step and a sort
    budget of two entries per slot (most rays with segments go to the
    replay list) still gives the plain path's colours and gradients, with
    forward requests in the same steps."""
    n = served.num_cells
    reqs = _requests(21, 9)
    srv = RenderServer(DEV, tile_capacity=64, quantum=16, bwd_capacity=8, max_buckets=n, entries_per_slot=2)
    srv.add_scene("s", served)
    with pytest.raises(ValueError):
        big = RenderServer(DEV, max_buckets=n - 1)
        big.add_scene("s", served)
    _drain(srv, reqs)
    assert int(srv.status.item()) == 0 and srv.backward == "grouped"
    plain = PlainRenderer({"s": served})
    for r in _requests(21, 9):
        plain.render(r)
        q = next(x for x in reqs if x.rid == r.rid)
        assert torch.equal(q.color, r.color) and torch.equal(q.log_t, r.log_t)
        if r.mode == "forward_backward":
            for k in ("spheres", "nsigmas", "texel_sites", "texel_rgb", "texel_height"):
                x, y = q.grads[k].cpu().numpy(), r.grads[k].cpu().numpy()
                assert np.linalg.norm(x - y) <= 1e-4 * max(np.linalg.norm(y), 1e-30), k
