"""Parity at the benchmark configurations' full sizes, on sampled rays.

* 1,000,000-cell Gaussian-mixture scene, one 1920x1080 camera, SH degree 3:
  the whole frame renders on the GPU; 2,048 sampled rays are walked by the
  oracle and must match bit for bit (visits, sequence hashes, colour, log T),
  and the backward of an L2 loss on those rays must match the oracle's
  within 1e-3 relative L2 per gradient array.
* 262,144 cells, 16 cameras at 800x800, each with its own texel colours
  (its SH pre-pass): with fp16 faces, the grouped, recorded and replay
  backwards of a loss on 1,024 sampled rays against the oracle's backward
  on every parameter; with f32 faces in exact texel mode, also against
  FP64 central differences (step 1e-4) of the FP64 forward on 64 random
  scalars.
  A scalar within a step of a jump (an exit face switches) or a bend (the
  min / max of two faces' t switches) is drawn again: there the central
  quotient disagrees with the half-step one or the one-sided slopes with
  each other.

Rays outside the sample get a zero incoming gradient, so both backwards
replay only the sampled rays.  Each case also runs the recorded backward
(the forward's record of contributing segments, read instead of a replay)
and checks it against the oracle and the replay; in the 16-camera case at a
cap that some rays exceed, so that the same launch replays those.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as orc
from powerfoam import raytrace as rt
from powerfoam import shading, workloads
from powerfoam.camera import pack_cameras
from powerfoam.power import build_power_adjacency_gpu

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def rt_status_cycle():
    from powerfoam import _lib

    return _lib.PF_STATUS_CYCLE  # the frame holds a looping ray (stopped, as in the forward)


def _setup(n, cams_fn, seed, f16=True):
    sc = workloads.gmm_scene(n, seed=seed)
    cams = cams_fn(sc["rng"])
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    pts, rad = t(sc["points"]), t(sc["radii"])
    adj, off = build_power_adjacency_gpu(pts, rad)
    faces = rt.walk_faces(pts, rad, adj, off, torch.float16 if f16 else torch.float32)
    words = pack_cameras(cams, [orc.start_cell(sc["points"], sc["radii"], c.eye.numpy()) for c in cams])
    cams_dev = torch.from_numpy(words).to(DEV)
    sites = t(sc["texel_sites"])
    K = sites.shape[1]
    rgb = torch.empty(len(cams), n, K, 3, device=DEV)
    shading.sh_forward(t(sc["sh"]), 3, sites, cams_dev, out=rgb)
    return sc, adj, off, faces, words, cams_dev, rgb, None


def _gpu_scene(sc, adj, off, faces, rgb, n_cams, n):
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    return rt.WalkScene(t(sc["points"]), t(sc["radii"]), t(sc["density"]), t(sc["normals"]), t(sc["texel_sites"]),
                        rgb, t(sc["texel_height"]), adj, off, faces)


def _cpu_scene(sc, adj, off, faces, rgb):
    f = faces.cpu()
    fz = f.view(torch.int16).numpy().view(np.uint16) if f.dtype == torch.float16 else f.numpy()
    return orc.Scene(sc["points"], sc["radii"], sc["density"], sc["normals"], sc["texel_sites"],
                     rgb, sc["texel_height"], adj.cpu().numpy(), off.cpu().numpy(), fz)


def test_1m_cells_1080p_sampled_rays():
    W, H = 1920, 1080
    n = 1_000_000
    sc, adj, off, faces, words, cams_dev, rgb, _ = _setup(n, lambda r: workloads.sphere_cameras(r, 1, W, H), 0)
    gs = _gpu_scene(sc, adj, off, faces, rgb[0], 1, n)
    g = rt.launch_walk_forward(gs, cams_dev, W, H, want_sequences=True)
    torch.cuda.synchronize()
    rays = np.sort(np.random.default_rng(7).choice(W * H, 2048, replace=False))
    cs = _cpu_scene(sc, adj, off, faces, rgb[0].cpu().numpy())
    c = orc.walk_forward(cs, words, rays=rays)
    assert np.array_equal(g["visits"].reshape(-1)[rays].cpu().numpy(), c["visits"])
    assert np.array_equal(g["seq_hash"].reshape(-1)[rays].cpu().numpy().view(np.uint64), c["seq_hash"])
    assert np.abs(g["color"].reshape(-1, 3)[rays].cpu().numpy() - c["color"]).max() <= 1e-4
    assert np.abs(g["log_t"].reshape(-1)[rays].cpu().numpy() - c["log_t"]).max() <= 1e-4
    # backward of an L2 loss on the sampled rays
    target = torch.rand(1, H, W, 3, generator=torch.Generator().manual_seed(3)).to(DEV)
    mask = torch.zeros(W * H, dtype=torch.bool, device=DEV)
    mask[torch.from_numpy(rays).to(DEV)] = True
    gc = (2.0 * (g["color"] - target)) * mask.view(1, H, W, 1)
    gg = rt.launch_walk_backward(gs, cams_dev, W, H, g["color"], g["log_t"], gc, None)
    # the benchmark step's form: the forward's record, read by the backward
    gr = rt.launch_walk_forward(gs, cams_dev, W, H, want_depth=False, record_cap=128)
    assert torch.equal(gr["color"], g["color"])
    grec = rt.launch_walk_backward(gs, cams_dev, W, H, g["color"], g["log_t"], gc, None, record=gr["record"])
    torch.cuda.synchronize()
    assert int(grec["status"].item()) & ~rt_status_cycle() == 0
    cb = orc.walk_backward(cs, words, g["color"].cpu().numpy(), gc.cpu().numpy())
    for k in ("spheres", "nsigmas", "texel_sites", "texel_rgb", "texel_height"):
        assert _rel(gg[k].cpu().numpy(), cb[k]) < 1e-3, k
        assert _rel(grec[k].cpu().numpy(), cb[k]) < 1e-3, k
        assert _rel(grec[k].cpu().numpy(), gg[k].cpu().numpy()) < 1e-5, k


def _per_camera_oracle_backward(sc, adj, off, faces, rgb, words, color, gc, gl, exact):
    """The oracle's backward of a batch whose cameras have their own texel
    colours: one camera at a time, shared arrays summed, colour gradients
    per camera."""
    B = words.shape[0]
    out = {k: 0.0 for k in ("spheres", "nsigmas", "texel_sites", "texel_height")}
    out["texel_rgb"] = []
    for b in range(B):
        cs = _cpu_scene(sc, adj, off, faces, rgb[b])
        cb = orc.walk_backward(cs, words[b:b + 1], color[b:b + 1], gc[b:b + 1], gl[b:b + 1], exact=exact)
        print(f"oracle backward: camera {b + 1} of {B}", flush=True)
        for k in ("spheres", "nsigmas", "texel_sites", "texel_height"):
            out[k] = out[k] + cb[k]
        out["texel_rgb"].append(cb["texel_rgb"])
    out["texel_rgb"] = np.stack(out["texel_rgb"])
    return out


def _c3_batch(f16):
    """262,144 cells, 16 cameras at 800x800, each camera with the texel
    colours of its own SH pre-pass (texel_rgb (16, N, K, 3)), and an L2 +
    alpha loss on 1,024 sampled rays."""
    W = H = 800
    n, B = 262_144, 16
    sc, adj, off, faces, words, cams_dev, rgb, _ = _setup(n, lambda r: workloads.sphere_cameras(r, B, W, H), 1,
                                                         f16=f16)
    gs = _gpu_scene(sc, adj, off, faces, rgb, B, n)
    assert gs.rgb_cams == B
    fwd = rt.launch_walk_forward(gs, cams_dev, W, H, record_cap=32)
    rng = np.random.default_rng(5)
    rays = np.sort(rng.choice(B * W * H, 1024, replace=False))
    target = rng.uniform(0, 1, (B, H, W, 3)).astype(np.float32)
    w_alpha = 0.3
    mask = torch.zeros(B * W * H, device=DEV)
    mask[torch.from_numpy(rays).to(DEV)] = 1.0
    gc = 2.0 * (fwd["color"] - torch.from_numpy(target).to(DEV)) * mask.view(B, H, W, 1)
    gl = -w_alpha * torch.exp(fwd["log_t"]) * mask.view(B, H, W)
    return sc, adj, off, faces, words, cams_dev, rgb, gs, fwd, rays, target, w_alpha, gc, gl, (W, H, n, B)


def test_262k_cells_16_cameras_per_camera_colours_fp16_faces():
    """The differentiation configuration as benchmarked: fp16 walk
    faces, per-camera colours, the reference's texel approximations; the
    grouped backward (per-camera colour buckets), the ray-by-ray recorded
    backward and the replay against the oracle on every parameter."""
    sc, adj, off, faces, words, cams_dev, rgb, gs, fwd, rays, target, w_alpha, gc, gl, dims = _c3_batch(True)
    W, H, n, B = dims
    grp = rt.launch_walk_backward(gs, cams_dev, W, H, fwd["color"], fwd["log_t"], gc, gl, record=fwd["record"],
                                  grouped=True)
    rec = rt.launch_walk_backward(gs, cams_dev, W, H, fwd["color"], fwd["log_t"], gc, gl, record=fwd["record"],
                                  grouped=False)
    rep = rt.launch_walk_backward(gs, cams_dev, W, H, fwd["color"], fwd["log_t"], gc, gl)
    torch.cuda.synchronize()
    assert tuple(grp["texel_rgb"].shape) == (B, n, 8, 3)
    assert int(grp["status"].item()) & ~rt_status_cycle() == 0
    cb = _per_camera_oracle_backward(sc, adj, off, faces, rgb.cpu().numpy(), words, fwd["color"].cpu().numpy(),
                                     gc.cpu().numpy(), gl.cpu().numpy(), exact=False)
    for k in ("spheres", "nsigmas", "texel_sites", "texel_rgb", "texel_height"):
        assert _rel(grp[k].cpu().numpy(), cb[k]) < 1e-3, k
        assert _rel(grp[k].cpu().numpy(), rep[k].cpu().numpy()) < 1e-5, k
        assert _rel(rec[k].cpu().numpy(), rep[k].cpu().numpy()) < 1e-5, k


def test_262k_cells_16_cameras_finite_differences():
    """Exact texel mode, f32 faces (so that the FP64 forward's exact faces
    are the kernels' faces to rounding), per-camera colours: the grouped and
    recorded backwards against the replay and the oracle on every
    parameter, and against FP64 central differences on 64 scalars."""
    sc, adj, off, faces, words, cams_dev, rgb, gs, fwd, rays, target, w_alpha, gc, gl, dims = _c3_batch(False)
    W, H, n, B = dims
    g = rt.launch_walk_backward(gs, cams_dev, W, H, fwd["color"], fwd["log_t"], gc, gl, exact_texel=True)
    gr = rt.launch_walk_backward(gs, cams_dev, W, H, fwd["color"], fwd["log_t"], gc, gl, exact_texel=True,
                                 record=fwd["record"], grouped=True)
    grr = rt.launch_walk_backward(gs, cams_dev, W, H, fwd["color"], fwd["log_t"], gc, gl, exact_texel=True,
                                  record=fwd["record"], grouped=False)
    torch.cuda.synchronize()
    assert bool((fwd["record"]["counts"] > 32).any())  # some rays replayed in the same launch
    for k in ("spheres", "nsigmas", "texel_sites", "texel_rgb", "texel_height"):
        assert _rel(gr[k].cpu().numpy(), g[k].cpu().numpy()) < 1e-5, k
        assert _rel(grr[k].cpu().numpy(), g[k].cpu().numpy()) < 1e-5, k
    g = {k: v.cpu().numpy() for k, v in g.items() if v is not None and k != "status"}
    rgb_np = rgb.cpu().numpy()
    cb = _per_camera_oracle_backward(sc, adj, off, faces, rgb_np, words, fwd["color"].cpu().numpy(),
                                     gc.cpu().numpy(), gl.cpu().numpy(), exact=True)
    for k in ("spheres", "nsigmas", "texel_sites", "texel_rgb", "texel_height"):
        assert _rel(g[k], cb[k]) < 1e-3, k
    # FP64 central differences of the loss on the sampled rays, camera by
    # camera with that camera's colours
    P = dict(points=sc["points"].astype(np.float64), radii=sc["radii"].astype(np.float64),
             nsigmas=np.concatenate([sc["normals"], sc["density"][:, None]], 1).astype(np.float64),
             texel_sites=sc["texel_sites"].astype(np.float64), texel_height=sc["texel_height"].astype(np.float64))
    rgb64 = rgb_np.astype(np.float64)
    an = dict(points=g["spheres"][:, :3], radii=g["spheres"][:, 3], nsigmas=g["nsigmas"],
              texel_sites=g["texel_sites"], texel_rgb=g["texel_rgb"], texel_height=g["texel_height"])
    adj_np, off_np = adj.cpu().numpy(), off.cpu().numpy()

    def loss():
        # every camera with its own colours, in one pass over the sampled rays
        return orc.walk_loss_f64(dict(P, texel_rgb=rgb64), adj_np, off_np, words, target, w_alpha=w_alpha,
                                 rays=rays)

    touched = np.flatnonzero(np.abs(g["nsigmas"][:, 3]) > 1e-6)
    names = list(P) + ["texel_rgb"]
    a, fd = [], []
    tries = 0
    while len(a) < 64 and tries < 256:
        name = names[tries % len(names)]
        tries += 1
        cell = int(rng_choice(touched, tries))
        if name == "texel_rgb":
            arr = rgb64
            b = int(np.random.default_rng(tries).integers(0, B))
            idx = (b, cell) + tuple(int(np.random.default_rng(tries + 1).integers(0, s)) for s in arr.shape[2:])
        else:
            arr = P[name]
            idx = (cell,) + tuple(int(np.random.default_rng(tries + 1).integers(0, s)) for s in arr.shape[1:])
        x0 = arr[idx]
        l0 = loss()
        v = {}
        for h in (1e-4, 5e-5):
            arr[idx] = x0 + h
            v[h] = loss()
            arr[idx] = x0 - h
            v[-h] = loss()
            arr[idx] = x0
        central = (v[1e-4] - v[-1e-4]) / 2e-4
        half = (v[5e-5] - v[-5e-5]) / 1e-4
        right, left = (v[1e-4] - l0) / 1e-4, (l0 - v[-1e-4]) / 1e-4
        # within a step of an exit-face switch the loss jumps (the two
        # steps' quotients disagree); within a step of a min / max switch
        # between two faces' t it bends (the one-sided slopes disagree):
        # the derivative there is one-sided, and the scalar is drawn again
        tol = abs(central) + 1e-6
        if abs(central - half) > 2e-3 * tol or abs(right - left) > 1e-2 * tol:
            continue
        a.append(float(an[name][idx]))
        fd.append(central)
        print(f"finite differences: {len(a)} of 64 scalars ({tries} drawn)", flush=True)
    assert len(a) == 64, f"only {len(a)} of {tries} scalars were away from a jump or bend"
    assert _rel(np.array(a), np.array(fd)) < 1e-3


def rng_choice(arr, k):
    return arr[np.random.default_rng(1000 + k).integers(0, arr.size)]
