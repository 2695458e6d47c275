"""Parity of the gfx950 walk (through the C ABI) with the CPU restatement.

Bar: visited-cell sequences bit-exact (visit counts, FNV-1a-64 hashes and
the recorded sequences); RGB, log-transmittance and depth within 1e-4
max-abs; gradients within 1e-3 relative L2 per array.
"""
import numpy as np
import pytest
import torch

from oracle import fp64
from oracle import oracle as orc
from powerfoam import _lib
from powerfoam import raytrace as rt
from powerfoam import workloads
from powerfoam.camera import cameras_to_device, pack_cameras

pytestmark = pytest.mark.gpu
TOL = 1e-4
DEV = "cuda"


def _gpu_scene(sc, attr=torch.float32, face_f16=True, packed=False):
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    pts, rad = t(sc["points"]), t(sc["radii"])
    adj, off = t(sc["adjacency"]).int(), t(sc["offsets"]).int()
    faces = rt.walk_faces(pts, rad, adj, off, torch.float16 if face_f16 else torch.float32)
    s = rt.WalkScene(pts, rad, t(sc["density"]), t(sc["normals"]), t(sc["texel_sites"]), t(sc["texel_rgb"]),
                     t(sc["texel_height"]), adj, off, faces, attr_dtype=attr, pack_faces=packed)
    return s, faces


def _cpu_scene(sc, faces_dev, attr=torch.float32):
    # attributes go through the same rounding the GPU scene applied
    def r(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(attr).float().numpy()

    f = faces_dev.cpu()
    faces = f.view(torch.int16).numpy().view(np.uint16) if f.dtype == torch.float16 else f.numpy()
    return orc.Scene(sc["points"], sc["radii"], r(sc["density"]), r(sc["normals"]), r(sc["texel_sites"]),
                     r(sc["texel_rgb"]), r(sc["texel_height"]), sc["adjacency"], sc["offsets"], faces)


def _cams(sc, cams):
    starts = [orc.start_cell(sc["points"], sc["radii"], c.eye.numpy()) for c in cams]
    return pack_cameras(cams, starts)


def _compare_forward(sc, cams, attr=torch.float32, face_f16=True, seq_cap=64, packed=False):
    gs, faces = _gpu_scene(sc, attr, face_f16, packed)
    words = _cams(sc, cams)
    cams_dev = torch.from_numpy(words).to(DEV)
    W, H = cams[0].width, cams[0].height
    g = rt.launch_walk_forward(gs, cams_dev, W, H, want_sequences=True, seq_cap=seq_cap)
    torch.cuda.synchronize()
    c = orc.walk_forward(_cpu_scene(sc, faces, attr), words, seq_cap=seq_cap)
    assert int(g["status"].item()) == 0 and c["status"] == 0
    assert np.array_equal(g["visits"].cpu().numpy(), c["visits"])
    assert np.array_equal(g["seq_hash"].cpu().numpy().view(np.uint64), c["seq_hash"])
    assert np.array_equal(g["seq"].cpu().numpy(), c["seq"])
    for k in ("color", "log_t", "depth"):
        err = np.abs(g[k].cpu().numpy() - c[k]).max()
        assert err <= TOL, (k, err)
    return g, c


def test_start_cells_match(scene4k):
    sc = scene4k
    eyes = [[0.0, -3.0, 0.0], [0.1, 0.2, 0.3], [2.0, 2.0, -2.0], [-0.5, 0.9, 0.0]]
    cams = [workloads.front_camera(16, 16)] * len(eyes)
    words = pack_cameras(cams)
    words.view(np.float32)[:, 0:3] = np.asarray(eyes, np.float32)
    cams_dev = torch.from_numpy(words).to(DEV)
    rt.start_cells_device(torch.from_numpy(sc["points"]).to(DEV), torch.from_numpy(sc["radii"]).to(DEV), cams_dev)
    got = cams_dev.cpu().numpy()[:, 9]
    want = [orc.start_cell(sc["points"], sc["radii"], np.asarray(e, np.float32)) for e in eyes]
    assert got.tolist() == want


@pytest.mark.parametrize("f16", [True, False])
def test_walk_faces_bit_equal(scene4k, f16):
    sc = scene4k
    _, faces = _gpu_scene(sc, face_f16=f16)
    want = orc.walk_faces(sc["points"], sc["radii"], sc["adjacency"], sc["offsets"], f16=f16)
    got = faces.cpu()
    got = got.view(torch.int16).numpy().view(np.uint16) if f16 else got.numpy()
    assert np.array_equal(got, want)


def _permuted(sc, order):
    """The scene with cell order[i] renumbered i, its power-diagram CSR
    rebuilt in the new order."""
    from powerfoam.adjacency import build_power_adjacency

    out = {k: np.ascontiguousarray(sc[k][order]) for k in ("points", "radii", "density", "normals",
                                                           "texel_sites", "texel_rgb", "texel_height")}
    adj, off = build_power_adjacency(torch.from_numpy(out["points"]), torch.from_numpy(out["radii"]))
    out["adjacency"], out["offsets"] = adj.numpy(), off.numpy()
    return out


@pytest.mark.parametrize("packed", [False, True])
def test_rows_at_the_end_of_the_face_array(scene4k, packed):
    """The cells a frame visits most, renumbered last: their CSR rows sit at
    the end of the face array, where the paired face loads run past the
    array's end (and read zeros) rather than into the next row.  Same bar
    as every walk: sequences bit-exact against the oracle."""
    sc = scene4k
    cams = [workloads.front_camera(64, 48)]
    _, faces = _gpu_scene(sc)
    c = orc.walk_forward(_cpu_scene(sc, faces), _cams(sc, cams), seq_cap=256)
    seq = c["seq"][c["seq"] >= 0]
    hits = np.bincount(seq, minlength=len(sc["radii"]))
    order = np.argsort(hits, kind="stable")  # most visited last
    assert hits[order[-1]] > 0 and hits[order[0]] == 0
    ps = _permuted(sc, order)
    E = ps["adjacency"].shape[0]
    last = int(np.searchsorted(ps["offsets"], E - 1, side="right")) - 1  # the cell owning the last row
    g, cp = _compare_forward(ps, cams, packed=packed, seq_cap=256)
    visited = cp["seq"][cp["seq"] >= 0]
    assert last == len(ps["radii"]) - 1 and last in set(visited.tolist())


def test_walk_uniform_4k_128px_bit_exact(scene4k):
    """The 4,096-cell, 128x128, 60 deg, seed-0 workload."""
    g, c = _compare_forward(scene4k, [workloads.front_camera(128, 128)])
    assert np.array_equal(g["color"].cpu().numpy(), c["color"])


@pytest.mark.parametrize("attr,f16,packed", [(torch.float16, True, True), (torch.float32, False, False),
                                             (torch.float16, False, False), (torch.float32, True, False),
                                             (torch.float16, True, False)])
def test_walk_dtypes(scene4k, attr, f16, packed):
    """Attribute dtypes x face layouts (fp16 packed records, fp16, fp32)."""
    gs, _ = _gpu_scene(scene4k, attr, f16, packed)
    want = _lib.PF_F16_PACKED if (f16 and packed) else (_lib.PF_F16 if f16 else _lib.PF_F32)
    assert gs.c.face_dtype == want
    _compare_forward(scene4k, [workloads.front_camera(64, 48)], attr=attr, face_f16=f16, packed=packed)


def test_packed_faces_match_reference_layout(scene4k):
    """Packing reference-layout fp16 faces gives the records pf_walk_faces
    writes directly (same fp16 bits, neighbour alongside)."""
    sc = scene4k
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    pts, rad = t(sc["points"]), t(sc["radii"])
    adj, off = t(sc["adjacency"]).int(), t(sc["offsets"]).int()
    f16 = rt.walk_faces(pts, rad, adj, off, torch.float16)
    packed = rt.pack_walk_faces(f16, adj)
    direct = torch.empty_like(packed)
    rc = _lib.load().pf_walk_faces(_lib.ptr(pts), _lib.ptr(rad), pts.shape[0], _lib.ptr(adj), _lib.ptr(off),
                                   adj.numel(), _lib.ptr(direct), _lib.PF_F16_PACKED, _lib.stream_ptr())
    _lib.check(rc, "pf_walk_faces")
    torch.cuda.synchronize()
    assert torch.equal(packed, direct)
    assert torch.equal(packed[:, 2], adj)
    assert torch.equal(packed[:, :2].contiguous().view(torch.float16).view(-1, 4), f16)


def test_walk_batch_ragged_tiles(scene4k):
    """Several cameras, width and height not multiples of the 8x8 tile."""
    rng = np.random.default_rng(5)
    cams = workloads.sphere_cameras(rng, 3, 37, 29)
    _compare_forward(scene4k, cams)


def test_walk_textured_cells(scene_textured):
    _compare_forward(scene_textured, [workloads.front_camera(72, 72)])


def test_walk_single_cell_no_edges():
    pts = np.zeros((1, 3), np.float32)
    sc = dict(points=pts, radii=np.array([0.5], np.float32), density=np.array([2.0], np.float32),
              normals=np.array([[0, 0, 1]], np.float32), texel_sites=np.repeat(pts[:, None], 8, 1),
              texel_rgb=np.full((1, 8, 3), 0.5, np.float32), texel_height=np.ones((1, 8), np.float32),
              adjacency=np.zeros(0, np.int32), offsets=np.zeros(2, np.int32))
    _compare_forward(sc, [workloads.front_camera(16, 16, hfov_deg=20.0)])


def test_walk_step_cap_flags_status(scene4k):
    gs, _ = _gpu_scene(scene4k)
    cams = [workloads.front_camera(16, 16)]
    cams_dev = torch.from_numpy(_cams(scene4k, cams)).to(DEV)
    g = rt.launch_walk_forward(gs, cams_dev, 16, 16, max_steps=3, want_sequences=True)
    assert int(g["status"].item()) & 1
    assert int(g["visits"].max().item()) == 3


def test_walk_deterministic(scene4k):
    gs, _ = _gpu_scene(scene4k)
    cams = [workloads.front_camera(64, 64)]
    cams_dev = torch.from_numpy(_cams(scene4k, cams)).to(DEV)
    a = rt.launch_walk_forward(gs, cams_dev, 64, 64, want_sequences=True)
    b = rt.launch_walk_forward(gs, cams_dev, 64, 64, want_sequences=True)
    assert torch.equal(a["color"], b["color"]) and torch.equal(a["seq_hash"], b["seq_hash"])


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("exact", [False, True])
def test_walk_backward_matches_cpu(scene_textured, exact):
    sc = scene_textured
    cams = [workloads.front_camera(40, 40, hfov_deg=40.0)]
    words = _cams(sc, cams)
    gs, faces = _gpu_scene(sc)
    cams_dev = torch.from_numpy(words).to(DEV)
    fwd = rt.launch_walk_forward(gs, cams_dev, 40, 40)
    rng = np.random.default_rng(1)
    target = torch.from_numpy(rng.uniform(0, 1, (1, 40, 40, 3)).astype(np.float32)).to(DEV)
    gc = 2.0 * (fwd["color"] - target)
    gl = -0.3 * torch.exp(fwd["log_t"])
    grads = rt.launch_walk_backward(gs, cams_dev, 40, 40, fwd["color"], fwd["log_t"], gc, gl, exact_texel=exact)
    cs = _cpu_scene(sc, faces)
    c = orc.walk_forward(cs, words)
    cg = orc.walk_backward(cs, words, c["color"], gc.cpu().numpy(), gl.cpu().numpy(), exact=exact)
    assert int(grads["status"].item()) == 0
    for k in ("spheres", "nsigmas", "texel_sites", "texel_rgb", "texel_height"):
        assert _rel(grads[k].cpu().numpy(), cg[k]) < 1e-3, k


def test_walk_autograd_function(scene_textured):
    """RayTracer.forward is differentiable; its gradients equal the direct
    backward launch (the glue maps spheres/nsigmas grads positionally)."""
    sc = scene_textured
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    pts = t(sc["points"]).requires_grad_(True)
    rad = t(sc["radii"]).requires_grad_(True)
    dens = t(sc["density"]).requires_grad_(True)
    nrm = t(sc["normals"]).requires_grad_(True)
    tsites = t(sc["texel_sites"]).requires_grad_(True)
    trgb = t(sc["texel_rgb"]).requires_grad_(True)
    th = t(sc["texel_height"]).requires_grad_(True)
    adj, off = t(sc["adjacency"]).int(), t(sc["offsets"]).int()
    faces = rt.walk_faces(pts, rad, adj, off)
    tracer = rt.RayTracer(None, DEV)
    cam = workloads.front_camera(32, 32, hfov_deg=40.0)
    color, alpha, depth = tracer.forward([cam], None, pts, rad, dens, nrm, tsites, trgb, th, adj, off, faces)
    loss = (color ** 2).sum() + 0.5 * alpha.sum()
    loss.backward()
    gs, _ = _gpu_scene(sc)
    words = _cams(sc, [cam])
    cams_dev = torch.from_numpy(words).to(DEV)
    fwd = rt.launch_walk_forward(gs, cams_dev, 32, 32)
    assert torch.equal(fwd["color"], color.detach())
    g = rt.launch_walk_backward(gs, cams_dev, 32, 32, fwd["color"], fwd["log_t"], 2 * fwd["color"],
                                -0.5 * torch.exp(fwd["log_t"]))
    assert _rel(pts.grad.cpu().numpy(), g["spheres"][:, :3].cpu().numpy()) < 1e-3
    assert _rel(rad.grad.cpu().numpy(), g["spheres"][:, 3].cpu().numpy()) < 1e-3
    assert _rel(dens.grad.cpu().numpy(), g["nsigmas"][:, 3].cpu().numpy()) < 1e-3
    assert _rel(trgb.grad.cpu().numpy(), g["texel_rgb"].cpu().numpy()) < 1e-3


def test_walk_gpu_gradients_match_fp64_finite_differences(scene_textured):
    """GPU backward (exact texel mode, f32 faces) vs FP64 central differences
    (step 1e-4) of the FP64 forward on 48 random scalar parameters."""
    sc = scene_textured
    cams = [workloads.front_camera(24, 24, hfov_deg=40.0)]
    words = _cams(sc, cams)
    gs, _ = _gpu_scene(sc, face_f16=False)
    cams_dev = torch.from_numpy(words).to(DEV)
    fwd = rt.launch_walk_forward(gs, cams_dev, 24, 24)
    target = np.random.default_rng(3).uniform(0, 1, (1, 24, 24, 3)).astype(np.float32)
    w_alpha = 0.3
    gc = 2.0 * (fwd["color"] - torch.from_numpy(target).to(DEV))
    gl = -w_alpha * torch.exp(fwd["log_t"])
    g = rt.launch_walk_backward(gs, cams_dev, 24, 24, fwd["color"], fwd["log_t"], gc, gl, exact_texel=True)
    g = {k: v.cpu().numpy() for k, v in g.items() if v is not None}
    P = dict(points=sc["points"].astype(np.float64), radii=sc["radii"].astype(np.float64),
             nsigmas=np.concatenate([sc["normals"], sc["density"][:, None]], 1).astype(np.float64),
             texel_sites=sc["texel_sites"].astype(np.float64), texel_rgb=sc["texel_rgb"].astype(np.float64),
             texel_height=sc["texel_height"].astype(np.float64))
    an = dict(points=g["spheres"][:, :3], radii=g["spheres"][:, 3], nsigmas=g["nsigmas"],
              texel_sites=g["texel_sites"], texel_rgb=g["texel_rgb"], texel_height=g["texel_height"])
    rng = np.random.default_rng(11)
    touched = np.flatnonzero(np.abs(g["nsigmas"][:, 3]) > 1e-6)
    a, fd = [], []
    for name in P:
        for _ in range(8):
            cell = int(rng.choice(touched))
            idx = (cell,) + tuple(int(rng.integers(0, s)) for s in P[name].shape[1:])
            Pp = {k: v.copy() for k, v in P.items()}
            Pm = {k: v.copy() for k, v in P.items()}
            Pp[name][idx] += 1e-4
            Pm[name][idx] -= 1e-4
            lp = orc.walk_loss_f64(Pp, sc["adjacency"], sc["offsets"], words, target, w_alpha=w_alpha)
            lm = orc.walk_loss_f64(Pm, sc["adjacency"], sc["offsets"], words, target, w_alpha=w_alpha)
            fd.append((lp - lm) / 2e-4)
            a.append(float(an[name][idx]))
    assert _rel(np.array(a), np.array(fd)) < 1e-3


def test_sh_forward_backward_match_fp64():
    from powerfoam import shading

    rng = np.random.default_rng(2)
    n, K = 300, 8
    coeffs = (0.3 * rng.standard_normal((n, 16, 3))).astype(np.float32)
    sites = rng.uniform(-1, 1, (n, K, 3)).astype(np.float32)
    cams = workloads.sphere_cameras(rng, 2, 8, 8)
    cams_dev = cameras_to_device(cams, None, DEV)
    c_t = torch.from_numpy(coeffs).to(DEV)
    s_t = torch.from_numpy(sites).to(DEV)
    rgb = shading.sh_forward(c_t, 3, s_t, cams_dev)
    Y = fp64.sh_basis_fp64(sites[None] - np.stack([c.eye.numpy() for c in cams])[:, None, None, :], 3)
    ref = np.maximum(0.0, 0.5 + np.einsum("bnkl,nlc->bnkc", Y, coeffs.astype(np.float64)))
    assert np.abs(rgb.cpu().numpy() - ref).max() < 1e-5
    g = rng.standard_normal(ref.shape).astype(np.float32)
    gc = shading.sh_backward(c_t, 3, s_t, cams_dev, torch.from_numpy(g).to(DEV))
    mask = (0.5 + np.einsum("bnkl,nlc->bnkc", Y, coeffs.astype(np.float64))) > 0
    ref_g = np.einsum("bnkl,bnkc->nlc", Y, g * mask)
    assert _rel(gc.cpu().numpy(), ref_g) < 1e-5


@pytest.mark.parametrize("hit_b", [False, True])
def test_loop_guard_matches_cpu(hit_b):
    """Hand-built three-cell loop: without attenuation both sides stop at the
    same step with status 4; with attenuation both walk it to the threshold."""
    from test_oracle_walk import _loop_scene

    S, words = _loop_scene(hit_b)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    pts, rad = t(S.spheres[:, :3]), t(S.spheres[:, 3])
    gs = rt.WalkScene(pts, rad, t(S.nsigmas[:, 3]), t(S.nsigmas[:, :3]), t(S.texel_sites), t(S.texel_rgb),
                      t(S.texel_height), t(S.adjacency), t(S.offsets), t(S.faces))
    g = rt.launch_walk_forward(gs, t(words), 3, 3, want_sequences=True, seq_cap=64, want_ray_status=True)
    torch.cuda.synchronize()
    c = orc.walk_forward(S, words, seq_cap=64)
    assert int(g["status"].item()) == c["status"] == (0 if hit_b else 4)
    # each ray's own bits: the rays the guard stopped, counted alike
    assert np.array_equal(g["ray_status"].cpu().numpy(), c["ray_status"])
    assert np.bitwise_or.reduce(c["ray_status"].reshape(-1)) == c["status"]
    assert np.array_equal(g["visits"].cpu().numpy(), c["visits"])
    assert np.array_equal(g["seq"].cpu().numpy(), c["seq"])
    assert np.array_equal(g["log_t"].cpu().numpy(), c["log_t"])
    gg = rt.launch_walk_backward(gs, t(words), 3, 3, g["color"], g["log_t"], torch.ones_like(g["color"]), None)
    torch.cuda.synchronize()
    assert int(gg["status"].item()) == (0 if hit_b else 4)
