"""Pins the CPU restatement of the tile rasterizer (oracle/pf_oracle.cpp)
against FP64 brute force: binning completeness and order, the forward
composite, and the backward against FP64 central differences.  CPU only."""
import numpy as np
import pytest

from oracle import fp64
from oracle import oracle as orc
from powerfoam import workloads
from powerfoam.camera import pack_cameras


def _rscene(sc, f16=False):
    """The rasterizer clips with the Čech complex (bvh.py), as the
    reference's training path does (scene.py:235-237)."""
    sph = np.concatenate([sc["points"], sc["radii"][:, None]], 1).astype(np.float32)
    adj, off = orc.cech(sc["points"], sc["radii"])
    faces = orc.raster_faces(sph, adj, off, f16=f16)
    return orc.RasterScene(sc["points"], sc["radii"], sc["density"], sc["normals"], sc["texel_sites"],
                           sc["texel_rgb"], sc["texel_height"], adj, off, faces), sph


def test_cech_brute_force_invariants(scene_textured):
    sc = scene_textured
    adj, off = orc.cech(sc["points"], sc["radii"])
    src = np.repeat(np.arange(len(sc["radii"])), np.diff(off))
    d = np.linalg.norm(sc["points"][adj].astype(np.float64) - sc["points"][src], axis=1)
    rs = sc["radii"][adj].astype(np.float64) + sc["radii"][src]
    assert np.all(d < rs + 1e-6) and np.all(src != adj)
    pairs = set(zip(src.tolist(), adj.tolist()))
    assert all((b, a) in pairs for a, b in pairs)
    # completeness: every overlapping pair (FP64) with a clear margin is listed
    p = sc["points"].astype(np.float64)
    r = sc["radii"].astype(np.float64)
    D = np.linalg.norm(p[:, None] - p[None], axis=-1)
    want = (D < (r[:, None] + r[None]) - 1e-5) & ~np.eye(len(r), dtype=bool)
    assert all((a, b) in pairs for a, b in zip(*np.nonzero(want)))


def test_raster_faces_definition(scene4k):
    sc = scene4k
    sph = np.concatenate([sc["points"], sc["radii"][:, None]], 1).astype(np.float32)
    f = orc.raster_faces(sph, sc["adjacency"], sc["offsets"], f16=False)
    src = np.repeat(np.arange(len(sc["radii"])), np.diff(sc["offsets"]))
    ref = np.concatenate([sph[sc["adjacency"], :3] - sph[src, :3], sph[sc["adjacency"], 3:]], 1)
    assert np.array_equal(f, ref)
    assert np.array_equal(orc.raster_faces(sph, sc["adjacency"], sc["offsets"], f16=True), orc.f32_to_f16_bits(ref))


def test_binning_is_complete_and_ordered(scene4k):
    """Every cell whose sphere a pixel ray hits (FP64) is in that pixel's
    tile list; every list is sorted by power distance from the eye."""
    sc = scene4k
    cam = workloads.front_camera(64, 48)
    w = pack_cameras([cam])[0]
    _, sph = _rscene(sc)
    offsets, prims = orc.raster_bin(sph, w)
    eye, dirs = fp64.camera_rays(w)
    p = sc["points"].astype(np.float64)
    r = sc["radii"].astype(np.float64)
    pd = np.sum((p - eye) ** 2, 1) - r * r
    H, W = 48, 64
    tiles_h = (H + 7) // 8
    for t in range((offsets.size - 1)):
        lst = prims[offsets[t]:offsets[t + 1]]
        assert np.all(np.diff(pd[lst]) >= 0)
    missing = 0
    oc = eye[None] - p
    for k in range(dirs.shape[0]):
        i, j = divmod(k, W)
        d = dirs[k]
        b = oc @ d
        h = r * r - np.sum((oc - b[:, None] * d) ** 2, 1)
        hit = (h > 1e-9) & (-b + np.sqrt(np.maximum(h, 0)) > 0)
        hit &= np.linalg.norm(p - eye, axis=1) >= 4 * r
        t = (i // 8) + (j // 8) * tiles_h
        lst = set(prims[offsets[t]:offsets[t + 1]].tolist())
        missing += len(set(np.flatnonzero(hit).tolist()) - lst)
    assert missing == 0


def test_raster_forward_matches_fp64(scene_textured):
    """Composite of sphere-clipped power cells in power-distance order equals
    the FP64 envelope composite (faces in f32; tiny threshold so the
    per-tile and per-ray early stops agree)."""
    sc = scene_textured
    cam = workloads.front_camera(40, 40)
    w = pack_cameras([cam])[0]
    S, sph = _rscene(sc, f16=False)
    offsets, prims = orc.raster_bin(sph, w)
    out = orc.raster_forward(S, w, offsets, prims, thr=1e-7)
    ref = fp64.render(sc, w, thr=1e-7)
    err = np.abs(out["color"].reshape(-1, 3) - ref["color"])
    aerr = np.abs(np.exp(out["log_t"].reshape(-1)) - np.exp(ref["log_t"]))
    assert err.mean() < 1e-5 and aerr.mean() < 1e-5
    assert err.max() < 1e-3 and aerr.max() < 1e-3


def test_raster_quantile_depths_match_fp64(scene4k):
    sc = scene4k
    cam = workloads.front_camera(32, 32)
    w = pack_cameras([cam])[0]
    S, sph = _rscene(sc, f16=False)
    offsets, prims = orc.raster_bin(sph, w)
    dq = np.tile(np.array([0.5], np.float32), (32, 32, 1))
    out = orc.raster_forward(S, w, offsets, prims, thr=1e-7, depth_q=dq)
    ref = fp64.render(sc, w, thr=1e-7, depth_q=0.5)
    got = out["qdepth"][..., 0].reshape(-1)
    hit = got >= 0
    assert hit.sum() > 100
    assert np.median(np.abs(got[hit] - ref["depth"][hit])) < 1e-5


def test_raster_early_stop_is_per_tile(scene4k):
    sc = scene4k
    cam = workloads.front_camera(32, 32)
    w = pack_cameras([cam])[0]
    S, sph = _rscene(sc, f16=True)
    offsets, prims = orc.raster_bin(sph, w)
    out = orc.raster_forward(S, w, offsets, prims, thr=0.97)
    lens = np.diff(offsets)
    assert np.all(out["early"] <= lens) and np.any(out["early"] < lens)


def test_raster_backward_matches_fp64_central_differences(scene4k):
    """Plain cells (texels at the site, no dipole): the reference's texel
    approximations vanish, so the analytic backward must match FD of the
    FP64 forward."""
    sc = scene4k
    cam = workloads.front_camera(24, 24, hfov_deg=30.0)
    w = pack_cameras([cam], [orc.start_cell(sc["points"], sc["radii"], cam.eye.numpy())])
    S, sph = _rscene(sc, f16=False)
    offsets, prims = orc.raster_bin(sph, w[0])
    thr = 1e-7
    out = orc.raster_forward(S, w[0], offsets, prims, thr=thr)
    rng = np.random.default_rng(3)
    target = rng.uniform(0, 1, (1, 24, 24, 3)).astype(np.float32)
    w_alpha = 0.3
    gc = 2.0 * (out["color"] - target[0])
    gl = -w_alpha * np.exp(out["log_t"])
    g = orc.raster_backward(S, w[0], offsets, prims, out["early"], out["log_t"], gc, gl, nthreads=1)
    P = dict(points=sc["points"].astype(np.float64), radii=sc["radii"].astype(np.float64),
             nsigmas=np.concatenate([sc["normals"], sc["density"][:, None]], 1).astype(np.float64),
             texel_sites=sc["texel_sites"].astype(np.float64), texel_rgb=sc["texel_rgb"].astype(np.float64),
             texel_height=sc["texel_height"].astype(np.float64))
    an = dict(points=g["spheres"][:, :3], radii=g["spheres"][:, 3], nsigmas=g["nsigmas"], texel_rgb=g["texel_rgb"])
    touched = np.flatnonzero(np.abs(g["nsigmas"][:, 3]) > 1e-6)
    a, fd = [], []
    for name in ("points", "radii", "nsigmas", "texel_rgb"):
        for _ in range(10):
            cell = int(rng.choice(touched))
            idx = (cell,) + tuple(int(rng.integers(0, s)) for s in P[name].shape[1:])
            if name == "nsigmas":
                idx = (cell, 3)
            Pp = {k: v.copy() for k, v in P.items()}
            Pm = {k: v.copy() for k, v in P.items()}
            Pp[name][idx] += 1e-4
            Pm[name][idx] -= 1e-4
            lp = orc.walk_loss_f64(Pp, sc["adjacency"], sc["offsets"], w, target, thr=thr, w_alpha=w_alpha)
            lm = orc.walk_loss_f64(Pm, sc["adjacency"], sc["offsets"], w, target, thr=thr, w_alpha=w_alpha)
            fd.append((lp - lm) / 2e-4)
            a.append(float(an[name][idx]))
    a, fd = np.array(a), np.array(fd)
    rel = np.linalg.norm(a - fd) / np.linalg.norm(fd)
    assert rel < 1e-3, rel


def test_raster_benchmark_and_visualize_agree_with_forward(scene4k):
    """The no-grad variants composite the same cells: with a tiny threshold
    and raster faces, visualize's colour equals the forward's."""
    sc = scene4k
    cam = workloads.front_camera(32, 32)
    w = pack_cameras([cam])[0]
    S, sph = _rscene(sc, f16=True)
    offsets, prims = orc.raster_bin(sph, w)
    fwd = orc.raster_forward(S, w, offsets, prims, thr=1e-7)
    vis = orc.raster_plain(S, w, offsets, prims, thr=1e-7, vis=True, max_isect=1 << 20)
    assert np.abs(vis["color"] - fwd["color"]).max() < 1e-5
    assert np.abs(vis["alpha"] - (1 - np.exp(fwd["log_t"]))).max() < 1e-5
    assert np.array_equal(vis["isect"].max() <= np.diff(offsets).max() + 1, True)
