"""The recorded backward: the forward writes each ray's contributing
segments and the backward differentiates them without replaying the walk,
ray by ray (pf_walk_backward_recorded) or sorted by cell
(pf_walk_backward_grouped).

Bar: the recording forward's outputs bit-equal to the plain forward's; the
record's cells an in-order subsequence of the visited sequence; gradients
within 1e-3 relative L2 of the oracle's, and within 1e-5 of the replay
backward's (the two differ only in the order of float additions), with the
record holding every ray or only some (cap 8: many replayed).
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


def _k(sc, K):
    out = dict(sc)
    for key in ("texel_sites", "texel_rgb", "texel_height"):
        out[key] = np.ascontiguousarray(sc[key][:, :K])
    return out


def _run(sc, cams, cap, exact=False):
    words = _cams(sc, cams)
    gs, faces = _gpu_scene(sc)
    cams_dev = torch.from_numpy(words).to(DEV)
    W, H = cams[0].width, cams[0].height
    plain = rt.launch_walk_forward(gs, cams_dev, W, H, want_sequences=True, seq_cap=4096)
    fwd = rt.launch_walk_forward(gs, cams_dev, W, H, want_sequences=True, seq_cap=4096, record_cap=cap)
    rng = np.random.default_rng(1)
    target = torch.from_numpy(rng.uniform(0, 1, (len(cams), H, W, 3)).astype(np.float32)).to(DEV)
    gc = 2.0 * (fwd["color"] - target)
    gl = -0.3 * torch.exp(fwd["log_t"])
    rec = rt.launch_walk_backward(gs, cams_dev, W, H, fwd["color"], fwd["log_t"], gc, gl, exact_texel=exact,
                                  record=fwd["record"], grouped=False)
    rep = rt.launch_walk_backward(gs, cams_dev, W, H, fwd["color"], fwd["log_t"], gc, gl, exact_texel=exact)
    grp = None
    if gs.K == 8:
        grp = rt.launch_walk_backward(gs, cams_dev, W, H, fwd["color"], fwd["log_t"], gc, gl, exact_texel=exact,
                                      record=fwd["record"], grouped=True)
    torch.cuda.synchronize()
    for key in ("color", "log_t", "depth", "visits", "seq_hash", "seq"):
        assert torch.equal(plain[key], fwd[key]), key
    assert int(rec["status"].item()) == 0
    if grp is not None:
        # the grouped backward: the same segments summed in cell order
        assert int(grp["status"].item()) == 0
        for key in KEYS:
            assert _rel(grp[key].cpu().numpy(), rep[key].cpu().numpy()) < 1e-5, key
    return words, gs, faces, fwd, gc, gl, rec, rep


@pytest.mark.parametrize("mixed", [False, True])
def test_recorded_backward_matches_cpu_and_replay(scene_textured, mixed):
    """mixed: a cap below the longest rays, so that the launch replays some
    rays and reads the others from the record."""
    sc = scene_textured
    cams = [workloads.front_camera(40, 40, hfov_deg=40.0)]
    cap = 512
    if mixed:
        words = _cams(sc, cams)
        gs, _ = _gpu_scene(sc)
        probe = rt.launch_walk_forward(gs, torch.from_numpy(words).to(DEV), 40, 40, record_cap=512)
        longest = int(probe["record"]["counts"].max().item())
        assert longest > 8
        cap = max(8, longest // 2 // 8 * 8)
    words, gs, faces, fwd, gc, gl, rec, rep = _run(sc, cams, cap)
    counts = rt.record_counts(fwd["record"], 1, 40, 40)
    if not mixed:
        assert int(counts.max().item()) <= cap
    else:
        assert bool((counts > cap).any()) and bool(((counts > 0) & (counts <= cap)).any())
    cs = _cpu_scene(sc, faces)
    c = orc.walk_forward(cs, words)
    cg = orc.walk_backward(cs, words, c["color"], gc.cpu().numpy(), gl.cpu().numpy())
    for key in KEYS:
        assert _rel(rec[key].cpu().numpy(), cg[key]) < 1e-3, key
        assert _rel(rec[key].cpu().numpy(), rep[key].cpu().numpy()) < 1e-5, key


def test_recorded_backward_exact_texel(scene_textured):
    words, gs, faces, fwd, gc, gl, rec, rep = _run(scene_textured, [workloads.front_camera(32, 32, hfov_deg=40.0)],
                                                   64, exact=True)
    for key in KEYS:
        assert _rel(rec[key].cpu().numpy(), rep[key].cpu().numpy()) < 1e-5, key


def test_recorded_backward_ragged_batch_and_k4(scene_textured):
    """Three cameras of 37x29 (tiles past the image edge) and K = 4 texels
    (the per-lane gradient path)."""
    sc = _k(scene_textured, 4)
    cams = workloads.sphere_cameras(np.random.default_rng(5), 3, 37, 29, radius=2.5)
    words, gs, faces, fwd, gc, gl, rec, rep = _run(sc, cams, 64)
    cs = _cpu_scene(sc, faces)
    c = orc.walk_forward(cs, words)
    cg = orc.walk_backward(cs, words, c["color"], gc.cpu().numpy(), gl.cpu().numpy())
    for key in KEYS:
        assert _rel(rec[key].cpu().numpy(), cg[key]) < 1e-3, key
        assert _rel(rec[key].cpu().numpy(), rep[key].cpu().numpy()) < 1e-5, key


def test_record_lists_contributing_cells_in_walk_order(scene_textured):
    cams = [workloads.front_camera(24, 24, hfov_deg=40.0)]
    words, gs, faces, fwd, gc, gl, rec, rep = _run(scene_textured, cams, 512)
    r = fwd["record"]
    S = r["counts"].numel()
    seg = r["segments"].cpu().numpy()
    cnt = r["counts"].cpu().numpy()
    assert r["cap"] == 512 and seg.shape == (512 * S, 4)
    seq = fwd["seq"].cpu().numpy()[0]
    off = scene_textured["offsets"]
    th = (24 + 7) // 8
    n_checked = 0
    for slot in range(S):
        tile, lane = divmod(slot, 64)
        i, j = (tile % th) * 8 + lane % 8, (tile // th) * 8 + lane // 8
        n = int(cnt[slot])
        visited = [int(v) for v in seq[i, j] if v >= 0]
        pos = 0
        for k in range(n):
            cell, tn_bits, tf_bits, ids = (int(x) for x in seg[rt.record_segment(r, k, slot)])
            while pos < len(visited) and visited[pos] != cell:
                pos += 1
            assert pos < len(visited), (slot, k, cell)
            pos += 1
            tn = np.int32(tn_bits).view(np.float32)
            tf = np.int32(tf_bits).view(np.float32)
            assert tn <= tf
            deg = int(off[cell + 1] - off[cell])
            lo, hi = ((ids & 0xFFFF) ^ 0x8000) - 0x8000, ids >> 16
            assert -2 <= lo < deg and -2 <= hi < deg
            n_checked += 1
    assert n_checked > 100


def test_autograd_record_and_replay_agree(scene_textured):
    """RayTracer.forward records by default; walk_record_cap = 0 replays."""
    import types

    sc = scene_textured
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    cam = workloads.front_camera(32, 32, hfov_deg=40.0)
    res = []
    for cap in (64, 0):
        params = [t(sc[k]).requires_grad_(True) for k in
                  ("points", "radii", "density", "normals", "texel_sites", "texel_rgb", "texel_height")]
        adj, off = t(sc["adjacency"]).int(), t(sc["offsets"]).int()
        faces = rt.walk_faces(params[0], params[1], adj, off)
        tracer = rt.RayTracer(types.SimpleNamespace(walk_record_cap=cap), DEV)
        color, alpha, _ = tracer.forward([cam], None, *params, adj, off, faces)
        ((color ** 2).sum() + 0.5 * alpha.sum()).backward()
        res.append([p.grad.cpu().numpy() for p in params])
    for a, b in zip(*res):
        assert _rel(a, b) < 1e-5


def test_record_of_another_walk_is_reported_not_read(scene_textured):
    """A recorded cell outside the scene is caught by the backward's bounds
    check: PF_STATUS_BAD_RECORD, the ray's remaining segments skipped."""
    from powerfoam import _lib

    sc = scene_textured
    cams = [workloads.front_camera(16, 16, hfov_deg=40.0)]
    words = _cams(sc, cams)
    gs, _ = _gpu_scene(sc)
    cams_dev = torch.from_numpy(words).to(DEV)
    fwd = rt.launch_walk_forward(gs, cams_dev, 16, 16, record_cap=64)
    r = fwd["record"]
    slot = int(torch.nonzero(r["counts"] > 0)[0, 0].item())
    r["segments"][rt.record_segment(r, 0, slot), 0] = sc["points"].shape[0]  # one past the last cell
    gc = torch.ones_like(fwd["color"])
    for grouped in (False, True):
        g = rt.launch_walk_backward(gs, cams_dev, 16, 16, fwd["color"], fwd["log_t"], gc, None, record=r,
                                    grouped=grouped)
        torch.cuda.synchronize()
        assert int(g["status"].item()) & _lib.PF_STATUS_BAD_RECORD
        assert all(bool(g[k].isfinite().all()) for k in KEYS)


@pytest.mark.parametrize("grouped", [False, True])
def test_record_stamp_of_other_cameras_refused(scene_textured, grouped):
    """A record written for other cameras (in-range cells, a stale record)
    fails the stamp check: PF_STATUS_BAD_RECORD and no gradient at all;
    the record's own cameras pass it."""
    from powerfoam import _lib

    sc = scene_textured
    cams = [workloads.front_camera(16, 16, hfov_deg=40.0)]
    words = _cams(sc, cams)
    gs, _ = _gpu_scene(sc)
    cams_dev = torch.from_numpy(words).to(DEV)
    fwd = rt.launch_walk_forward(gs, cams_dev, 16, 16, record_cap=64)
    other = words.copy()
    other[0, 0] = np.float32(other[0, :1].view(np.float32)[0] + 0.01).view(np.int32)  # the eye moved
    gc = torch.ones_like(fwd["color"])
    for w, bad in ((other, True), (words, False)):
        g = rt.launch_walk_backward(gs, torch.from_numpy(w).to(DEV), 16, 16, fwd["color"], fwd["log_t"], gc, None,
                                    record=fwd["record"], grouped=grouped)
        torch.cuda.synchronize()
        st = int(g["status"].item())
        assert bool(st & _lib.PF_STATUS_BAD_RECORD) == bad, st
        total = sum(float(g[k].abs().sum().item()) for k in KEYS)
        assert (total == 0.0) == bad, total


def test_grouped_backward_per_camera_colours(scene_textured):
    """Three cameras, each with its own texel colours (the colour pre-pass
    per camera): the grouped backward sums runs per (camera, cell), so each
    camera's colour gradients equal that camera's alone and the shared
    arrays equal the sum over the cameras."""
    sc = scene_textured
    cams = workloads.sphere_cameras(np.random.default_rng(3), 3, 24, 24, radius=2.5)
    words = _cams(sc, cams)
    rng = np.random.default_rng(4)
    rgb = rng.uniform(0, 1, (3,) + sc["texel_rgb"].shape).astype(np.float32)
    gs1, faces = _gpu_scene(sc)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    gsb = rt.WalkScene(t(sc["points"]), t(sc["radii"]), t(sc["density"]), t(sc["normals"]), t(sc["texel_sites"]),
                       t(rgb), t(sc["texel_height"]), t(sc["adjacency"]), t(sc["offsets"]), gs1.faces)
    cams_dev = t(words)
    fwd = rt.launch_walk_forward(gsb, cams_dev, 24, 24, record_cap=64)
    gc = 2.0 * (fwd["color"] - 0.4)
    grp = rt.launch_walk_backward(gsb, cams_dev, 24, 24, fwd["color"], fwd["log_t"], gc, None, record=fwd["record"],
                                  grouped=True)
    rec = rt.launch_walk_backward(gsb, cams_dev, 24, 24, fwd["color"], fwd["log_t"], gc, None, record=fwd["record"],
                                  grouped=False)
    assert tuple(grp["texel_rgb"].shape) == (3,) + sc["texel_rgb"].shape
    sums = {k: 0.0 for k in KEYS if k != "texel_rgb"}
    for b in range(3):
        gsk = rt.WalkScene(t(sc["points"]), t(sc["radii"]), t(sc["density"]), t(sc["normals"]),
                           t(sc["texel_sites"]), t(rgb[b]), t(sc["texel_height"]), t(sc["adjacency"]),
                           t(sc["offsets"]), gs1.faces)
        one = rt.launch_walk_backward(gsk, cams_dev[b:b + 1].contiguous(), 24, 24, fwd["color"][b:b + 1].contiguous(),
                                      fwd["log_t"][b:b + 1].contiguous(), gc[b:b + 1].contiguous(), None)
        torch.cuda.synchronize()
        assert _rel(grp["texel_rgb"][b].cpu().numpy(), one["texel_rgb"].cpu().numpy()) < 1e-5, b
        for k in sums:
            sums[k] = sums[k] + one[k].cpu().numpy()
    for k in sums:
        assert _rel(grp[k].cpu().numpy(), sums[k]) < 1e-5, k
    for k in KEYS:
        assert _rel(grp[k].cpu().numpy(), rec[k].cpu().numpy()) < 1e-5, k


@pytest.mark.parametrize("attr,f16,packed", [(torch.float16, True, True), (torch.float32, False, False),
                                             (torch.float16, False, False), (torch.float32, True, False),
                                             (torch.float16, True, False)])
def test_recorded_backward_dtypes(scene_textured, attr, f16, packed):
    """Every attribute dtype x face layout: the recording forward equals the
    plain one and the recorded backward equals the replay."""
    sc = scene_textured
    cams = [workloads.front_camera(40, 32, hfov_deg=40.0)]
    words = _cams(sc, cams)
    gs, _ = _gpu_scene(sc, attr, f16, packed)
    cams_dev = torch.from_numpy(words).to(DEV)
    plain = rt.launch_walk_forward(gs, cams_dev, 40, 32, want_sequences=True)
    fwd = rt.launch_walk_forward(gs, cams_dev, 40, 32, want_sequences=True, record_cap=16)
    for key in ("color", "log_t", "depth", "visits", "seq_hash"):
        assert torch.equal(plain[key], fwd[key]), key
    gc = 2.0 * (fwd["color"] - 0.5)
    rec = rt.launch_walk_backward(gs, cams_dev, 40, 32, fwd["color"], fwd["log_t"], gc, None, record=fwd["record"],
                                  grouped=False)
    grp = rt.launch_walk_backward(gs, cams_dev, 40, 32, fwd["color"], fwd["log_t"], gc, None, record=fwd["record"],
                                  grouped=True)
    rep = rt.launch_walk_backward(gs, cams_dev, 40, 32, fwd["color"], fwd["log_t"], gc, None)
    torch.cuda.synchronize()
    for key in KEYS:
        assert _rel(rec[key].cpu().numpy(), rep[key].cpu().numpy()) < 1e-5, key
        assert _rel(grp[key].cpu().numpy(), rep[key].cpu().numpy()) < 1e-5, key
