"""Pins the CPU restatement of the walk (oracle/pf_oracle.cpp) without the
reference (Warp is not installable here): against an independent FP64
brute force (oracle/fp64.py), closed forms, IEEE conversions, invariants and
FP64 central finite differences.  CPU only."""
import numpy as np
import pytest
import torch

from oracle import fp64
from oracle import oracle as orc
from powerfoam import workloads
from powerfoam.adjacency import build_power_adjacency, csr_is_symmetric
from powerfoam.camera import pack_cameras


def _scene(sc, faces):
    return orc.Scene(sc["points"], sc["radii"], sc["density"], sc["normals"], sc["texel_sites"], sc["texel_rgb"],
                     sc["texel_height"], sc["adjacency"], sc["offsets"], faces)


def _cam_words(sc, cam):
    return pack_cameras([cam], [orc.start_cell(sc["points"], sc["radii"], cam.eye.numpy())])


# ---------------------------------------------------------------- scalars --
def test_expf_within_one_ulp_of_fp64():
    x = np.concatenate([np.linspace(-103.0, 0.0, 200001, dtype=np.float32),
                        np.linspace(0.0, 88.0, 20001, dtype=np.float32)])
    got = orc.expf(x).astype(np.float64)
    ref = np.exp(x.astype(np.float64))
    ulp = np.spacing(ref.astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(got - ref) / ulp) <= 1.0
    sp = orc.expf(np.array([0.0, -np.inf, np.inf, np.nan, -200.0, 100.0], np.float32))
    assert sp[0] == 1.0 and sp[1] == 0.0 and np.isinf(sp[2]) and np.isnan(sp[3]) and sp[4] == 0.0 and np.isinf(sp[5])


def test_logf_within_three_ulp_of_fp64():
    # used only for the depth output (tolerance 1e-4), never for a walk decision
    x = np.exp(np.linspace(-80, 80, 200001)).astype(np.float32)
    got = orc.logf(x).astype(np.float64)
    ref = np.log(x.astype(np.float64))
    ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(got - ref) / ulp) <= 3.0
    assert orc.logf(np.array([1.0], np.float32))[0] == 0.0


def test_half_conversion_is_ieee_round_to_nearest_even():
    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.standard_normal(200000).astype(np.float32) * np.float32(10.0) ** rng.integers(-9, 6, 200000),
        np.array([0.0, -0.0, 65504.0, 65519.996, 65520.0, 1e9, -1e9, 2.0**-24, 2.0**-25, 3 * 2.0**-26,
                  2.0**-14, 2.0**-14 - 2.0**-25, 1.0 + 2.0**-11, 1.0 + 3 * 2.0**-11, np.inf, -np.inf], np.float32),
    ]).astype(np.float32)
    got = orc.f32_to_f16_bits(x)
    ref = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, ref)
    back = orc.f16_bits_to_f32(ref)
    assert np.array_equal(back, ref.view(np.float16).astype(np.float32))


# ------------------------------------------------------------ camera, start --
def test_ray_dirs_match_fp64_and_project_to_their_pixel():
    cam = workloads.front_camera(48, 32)
    w = pack_cameras([cam])[0]
    d32 = orc.ray_dirs(w)
    eye, d64 = fp64.camera_rays(w)
    assert np.max(np.abs(d32.reshape(-1, 3) - d64)) < 1e-6
    # closed form: the point eye + t d projects back onto pixel (j, i) under
    # the intrinsics of camera.py:251-263 and the extrinsics of :265-282
    P = cam.projection_matrix().double().numpy()
    X = np.concatenate([eye + 2.5 * d64, np.ones((d64.shape[0], 1))], 1)
    uvw = X @ P.T
    ij = np.stack(np.meshgrid(np.arange(32), np.arange(48), indexing="ij"), -1).reshape(-1, 2)
    # w2c's second axis is f x r = -up, so image rows grow downwards: v == i
    u = uvw[:, 0] / uvw[:, 2]
    v = uvw[:, 1] / uvw[:, 2]
    assert np.max(np.abs(u - ij[:, 1])) < 1e-4
    assert np.max(np.abs(v - ij[:, 0])) < 1e-4


def test_start_cell_is_the_power_owner_of_the_eye(scene4k):
    sc = scene4k
    for eye in ([0.0, -3.0, 0.0], [0.1, 0.2, 0.3], [2.0, 2.0, -2.0]):
        eye = np.asarray(eye, np.float32)
        k = orc.start_cell(sc["points"], sc["radii"], eye)
        pw = np.sum((sc["points"].astype(np.float64) - eye) ** 2, 1) - sc["radii"].astype(np.float64) ** 2
        assert k == int(np.argmin(pw))


def test_start_cell_first_index_on_ties():
    pts = np.array([[1, 0, 0], [0, 1, 0], [-1, 0, 0]], np.float32)
    assert orc.start_cell(pts, np.zeros(3, np.float32), np.zeros(3, np.float32)) == 0


# ------------------------------------------------------------ adjacency --
def test_power_adjacency_invariants(scene4k):
    ok, msg = csr_is_symmetric(scene4k["adjacency"], scene4k["offsets"])
    assert ok, msg
    deg = np.diff(scene4k["offsets"])
    assert 10 < deg.mean() < 30


def test_walk_faces_match_their_definition(scene4k):
    sc = scene4k
    f32 = orc.walk_faces(sc["points"], sc["radii"], sc["adjacency"], sc["offsets"], f16=False)
    src = np.repeat(np.arange(len(sc["radii"])), np.diff(sc["offsets"]))
    p = sc["points"].astype(np.float64)
    r = sc["radii"].astype(np.float64)
    pm = 0.5 * (np.sum(p * p, 1) - r * r)
    ref = np.concatenate([p[sc["adjacency"]] - p[src], (pm[sc["adjacency"]] - pm[src])[:, None]], 1)
    assert np.max(np.abs(f32 - ref)) < 1e-6
    f16 = orc.walk_faces(sc["points"], sc["radii"], sc["adjacency"], sc["offsets"], f16=True)
    assert np.array_equal(f16, orc.f32_to_f16_bits(f32))


# ------------------------------------------------------------- the walk --
def _fp64_compare(sc, cam, check_seq=True):
    w = _cam_words(sc, cam)
    faces = orc.walk_faces(sc["points"], sc["radii"], sc["adjacency"], sc["offsets"], f16=False)
    out = orc.walk_forward(_scene(sc, faces), w, seq_cap=512)
    ref = fp64.render(sc, w[0])
    n_mis = 0
    if check_seq:
        seqs = out["seq"].reshape(-1, 512)
        for k in range(seqs.shape[0]):
            s = seqs[k][seqs[k] >= 0].tolist()
            if s != ref["seq"][k] and ref["margin"][k] > 1e-6:
                n_mis += 1
    col = out["color"].reshape(-1, 3)
    err = np.abs(col - ref["color"])
    aerr = np.abs(np.exp(out["log_t"].reshape(-1)) - np.exp(ref["log_t"]))
    return n_mis, err, aerr, out, ref


def test_walk_sequences_equal_fp64_power_owners(scene4k):
    cam = workloads.front_camera(64, 64)
    n_mis, err, aerr, out, _ = _fp64_compare(scene4k, cam)
    assert out["status"] == 0
    assert n_mis == 0
    assert err.mean() < 1e-5 and aerr.mean() < 1e-5
    assert err.max() < 1e-3 and aerr.max() < 1e-3


def test_walk_textured_cells_match_fp64(scene_textured):
    cam = workloads.front_camera(48, 48)
    n_mis, err, aerr, out, _ = _fp64_compare(scene_textured, cam)
    assert n_mis == 0
    assert err.mean() < 1e-4 and aerr.mean() < 1e-4


def test_a_missing_delaunay_edge_is_caught(scene4k):
    """Planted mistake: drop the CSR edges of the start cell's most visited
    exit; the oracle's sequences must then disagree with the FP64 owners."""
    sc = dict(scene4k)
    cam = workloads.front_camera(32, 32)
    w = _cam_words(sc, cam)
    start = int(w[0, 9])
    off, adj = sc["offsets"], sc["adjacency"]
    row = adj[off[start]:off[start + 1]]
    faces = orc.walk_faces(sc["points"], sc["radii"], adj, off, f16=False)
    seq = orc.walk_forward(_scene(sc, faces), w, seq_cap=4)["seq"].reshape(-1, 4)
    victim = np.bincount(seq[:, 1][seq[:, 1] >= 0]).argmax()
    keep = np.ones(adj.size, bool)
    keep[off[start]:off[start + 1]][row == victim] = False
    src = np.repeat(np.arange(len(off) - 1), np.diff(off))
    keep &= ~((src == victim) & (adj == start))
    new_adj = adj[keep]
    new_off = np.concatenate([[0], np.cumsum(np.bincount(src[keep], minlength=len(off) - 1))]).astype(np.int32)
    sc["adjacency"], sc["offsets"] = new_adj, new_off
    n_mis, _, _, _, _ = _fp64_compare(sc, cam)
    assert n_mis > 0


def test_identical_texels_give_the_texel_colour():
    """Closed form: K identical texels => colour equals that texel's colour,
    and a single isolated cell composites to (1 - exp(-sigma chord)) c."""
    pts = np.array([[0.0, 0.0, 0.0]], np.float32)
    r = np.array([0.5], np.float32)
    sc = dict(points=pts, radii=r, density=np.array([2.0], np.float32),
              normals=np.array([[0, 0, 1]], np.float32), texel_sites=np.repeat(pts[:, None], 8, 1),
              texel_rgb=np.tile(np.array([0.2, 0.5, 0.9], np.float32), (1, 8, 1)),
              texel_height=np.full((1, 8), 1.0, np.float32), adjacency=np.zeros(0, np.int32),
              offsets=np.zeros(2, np.int32))
    cam = workloads.front_camera(9, 9, hfov_deg=10.0)
    w = pack_cameras([cam], [0])
    out = orc.walk_forward(_scene(sc, np.zeros((0, 4), np.float32)), w)
    c = out["color"][0, 4, 4]
    expect = (1 - np.exp(-2.0 * 1.0)) * np.array([0.2, 0.5, 0.9])
    assert np.allclose(c, expect, atol=2e-6)
    assert np.isclose(out["depth"][0, 4, 4], 2.5 + np.log(2.0) / 2.0, atol=1e-5)


def test_step_cap_is_reported_not_hung(scene4k):
    sc = scene4k
    cam = workloads.front_camera(16, 16)
    faces = orc.walk_faces(sc["points"], sc["radii"], sc["adjacency"], sc["offsets"], f16=True)
    out = orc.walk_forward(_scene(sc, faces), _cam_words(sc, cam), max_steps=3)
    assert out["status"] & 1
    assert out["visits"].max() == 3


# ------------------------------------------------------ backward vs FD --
def _fd_setup(sc, W=24, H=24):
    cam = workloads.front_camera(W, H, hfov_deg=40.0)
    w = _cam_words(sc, cam)
    rng = np.random.default_rng(3)
    target = rng.uniform(0, 1, (1, H, W, 3)).astype(np.float32)
    return w, target


def test_walk_backward_matches_fp64_central_differences(scene_textured):
    sc = scene_textured
    w, target = _fd_setup(sc)
    faces = orc.walk_faces(sc["points"], sc["radii"], sc["adjacency"], sc["offsets"], f16=False)
    S = _scene(sc, faces)
    fwd = orc.walk_forward(S, w)
    w_alpha = 0.3
    gc = 2.0 * (fwd["color"] - target)
    gl = -w_alpha * np.exp(fwd["log_t"])  # d/dlog_t of w_alpha (1 - exp(log_t))
    g = orc.walk_backward(S, w, fwd["color"], gc, gl, exact=True)
    assert g["status"] == 0
    P = dict(points=sc["points"].astype(np.float64), radii=sc["radii"].astype(np.float64),
             nsigmas=np.concatenate([sc["normals"], sc["density"][:, None]], 1).astype(np.float64),
             texel_sites=sc["texel_sites"].astype(np.float64), texel_rgb=sc["texel_rgb"].astype(np.float64),
             texel_height=sc["texel_height"].astype(np.float64))
    analytic = dict(points=g["spheres"][:, :3], radii=g["spheres"][:, 3], nsigmas=g["nsigmas"],
                    texel_sites=g["texel_sites"], texel_rgb=g["texel_rgb"], texel_height=g["texel_height"])
    rng = np.random.default_rng(11)
    touched = np.flatnonzero(np.abs(g["nsigmas"][:, 3]) > 1e-6)
    picks = []
    for name in ("points", "radii", "nsigmas", "texel_sites", "texel_rgb", "texel_height"):
        for _ in range(8):
            cell = int(rng.choice(touched))
            idx = (cell,) + tuple(int(rng.integers(0, s)) for s in P[name].shape[1:])
            picks.append((name, idx))
    h = 1e-4
    a, fd = [], []
    for name, idx in picks:
        Pp = {k: v.copy() for k, v in P.items()}
        Pm = {k: v.copy() for k, v in P.items()}
        Pp[name][idx] += h
        Pm[name][idx] -= h
        lp = orc.walk_loss_f64(Pp, sc["adjacency"], sc["offsets"], w, target, w_alpha=w_alpha)
        lm = orc.walk_loss_f64(Pm, sc["adjacency"], sc["offsets"], w, target, w_alpha=w_alpha)
        fd.append((lp - lm) / (2 * h))
        a.append(float(analytic[name][idx]))
    a, fd = np.array(a), np.array(fd)
    rel = np.linalg.norm(a - fd) / np.linalg.norm(fd)
    assert rel < 1e-3, (rel, list(zip(picks, a, fd)))


def test_fp64_loss_with_per_camera_colours_is_the_sum_over_cameras(scene_textured):
    """One FP64 loss call over several cameras, each with its own texel
    colours (texel_rgb (B, N, K, 3)), on sampled rays of all of them equals
    the sum of one call per camera with that camera's colours; with
    identical colours it equals the shared-colour call."""
    sc = scene_textured
    rng = np.random.default_rng(3)
    W = H = 24
    cams = workloads.sphere_cameras(rng, 3, W, H)
    words = np.concatenate([_cam_words(sc, c) for c in cams])
    B, N, K = 3, len(sc["radii"]), 8
    P = dict(points=sc["points"].astype(np.float64), radii=sc["radii"].astype(np.float64),
             nsigmas=np.concatenate([sc["normals"], sc["density"][:, None]], 1).astype(np.float64),
             texel_sites=sc["texel_sites"].astype(np.float64), texel_height=sc["texel_height"].astype(np.float64))
    rgb = rng.uniform(0, 1, (B, N, K, 3))
    target = rng.uniform(0, 1, (B, H, W, 3)).astype(np.float32)
    rays = np.sort(rng.choice(B * W * H, 300, replace=False))
    both = orc.walk_loss_f64(dict(P, texel_rgb=rgb), sc["adjacency"], sc["offsets"], words, target, w_alpha=0.3,
                             rays=rays)
    per = 0.0
    for b in range(B):
        rb = rays[(rays >= b * W * H) & (rays < (b + 1) * W * H)] - b * W * H
        per += orc.walk_loss_f64(dict(P, texel_rgb=rgb[b]), sc["adjacency"], sc["offsets"], words[b:b + 1],
                                 target[b:b + 1], w_alpha=0.3, rays=rb)
    assert abs(both - per) <= 1e-12 * abs(per)
    same = orc.walk_loss_f64(dict(P, texel_rgb=np.stack([rgb[0]] * B)), sc["adjacency"], sc["offsets"], words,
                             target, w_alpha=0.3, rays=rays)
    shared = orc.walk_loss_f64(dict(P, texel_rgb=rgb[0]), sc["adjacency"], sc["offsets"], words, target,
                               w_alpha=0.3, rays=rays)
    assert abs(same - shared) <= 1e-12 * abs(shared)  # the threads' sums may add in another order
    with pytest.raises(ValueError):
        orc.walk_loss_f64(dict(P, texel_rgb=rgb[:2]), sc["adjacency"], sc["offsets"], words, target, rays=rays)


# ------------------------------------------------------ walks that loop --
def _loop_scene(hit_b):
    """Three cells whose exit faces send the rays of a narrow 3x3 camera
    A -> B -> C -> A (exit t = 1, 2, 1.5 along +y from the eye (0,-3,0)).
    With hit_b, cell B's sphere (centre t = 1.5, r = 0.3) is hit, so every
    pass through the loop attenuates and the walk ends on transmittance."""
    K = 8
    pts = np.array([[5, 5, 5], [0, -1.5, 0] if hit_b else [5, 5, 6], [5, 6, 5]], np.float32)
    rad = np.array([0.01, 0.3 if hit_b else 0.01, 0.01], np.float32)
    normals = np.tile(np.array([0, 0, 1], np.float32), (3, 1))
    sites = np.repeat(pts[:, None, :], K, 1)
    adj = np.array([1, 2, 0], np.int32)
    off = np.array([0, 1, 2, 3], np.int32)
    faces = np.array([[0, 1, 0, 1 - 3], [0, 1, 0, 2 - 3], [0, 1, 0, 1.5 - 3]], np.float32)
    S = orc.Scene(pts, rad, np.ones(3, np.float32), normals, sites, np.full((3, K, 3), 0.5, np.float32),
                  np.repeat(2 * rad[:, None], K, 1), adj, off, faces)
    w = pack_cameras([workloads.front_camera(3, 3, hfov_deg=1.0)], [0])  # all 9 rays within 0.02 of the axis
    return S, w


def test_non_attenuating_loop_stops_at_brent_step():
    S, w = _loop_scene(hit_b=False)
    out = orc.walk_forward(S, w, seq_cap=16)
    seq = out["seq"][0, 1, 1]
    assert out["status"] == 4
    # states A B C A B C: Brent saves A (step 0), C (step 2) and finds C at step 5
    assert out["visits"][0, 1, 1] == 5 and list(seq[:5]) == [0, 1, 2, 0, 1]
    assert out["log_t"][0, 1, 1] == 0
    # the centre ray's own bits name the guard (the OR over rays is status)
    assert out["ray_status"][0, 1, 1] == 4
    assert np.bitwise_or.reduce(out["ray_status"].reshape(-1)) == out["status"]
    g = orc.walk_backward(S, w, out["color"], np.ones_like(out["color"]))
    assert g["status"] == 4


def test_attenuating_loop_is_walked_to_the_threshold():
    """The reference's loop ends on such a ray, so the walk must not stop it."""
    S, w = _loop_scene(hit_b=True)
    out = orc.walk_forward(S, w, seq_cap=256)
    v = out["visits"][0, 1, 1]
    assert out["status"] == 0 and np.all(np.exp(out["log_t"]) < 1e-3)
    assert v > 9 and list(out["seq"][0, 1, 1][:9]) == [0, 1, 2] * 3


def test_large_scene_loops_only_where_state_repeats(scene4k):
    """On a real scene the guard changes nothing: every ray's walk equals the
    FP64 owner sequence test's walk, so no state repeats and no bit 4."""
    sc = scene4k
    faces = orc.walk_faces(sc["points"], sc["radii"], sc["adjacency"], sc["offsets"], f16=True)
    out = orc.walk_forward(_scene(sc, faces), _cam_words(sc, workloads.front_camera(64, 64)), seq_cap=512)
    assert out["status"] == 0
    for s, n in zip(out["seq"].reshape(-1, 512), out["visits"].reshape(-1)):
        s = s[:min(n, 512)]
        assert len(np.unique(s)) == len(s)
