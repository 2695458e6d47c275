"""Pins the rest of the tile rasterizer's restatement (oracle/pf_oracle.cpp):
its side outputs against the independent FP64 envelope oracle
(oracle/fp64.py raster_render), and every incoming-gradient chain and every
gradient family of its backward against FP64 central differences of the
FP64 raster functional (oracle.raster_loss_f64).

Reference semantics pinned here (rasterize.py):
- forward side outputs :1088-1129 (normal distance sum (n.d)^2 alpha T for
  n.d > 0, normal sum alpha T n, quantile depths, per-cell contrib and
  point_err with the 1/(H W) scale);
- backward chains :1371-1426 (colour, log T, normal, normal distance,
  quantile depths, contrib), the texel backward texture.py:74-186.

The reference's backward differentiates the first power of n.d for the
normal distance (:1383, :1393, :1424) while its forward accumulates the
second power (:1114).  The restatement reproduces that; this file pins the
backward as the exact gradient of the first-power functional and the
forward as the second-power sum.  CPU only."""
import numpy as np
import pytest

from oracle import fp64
from oracle import oracle as orc
from powerfoam import workloads
from powerfoam.camera import pack_cameras

QS = np.array([0.9, 0.6, 0.3], np.float32)


def _rscene(sc):
    sph = np.concatenate([sc["points"], sc["radii"][:, None]], 1).astype(np.float32)
    adj, off = orc.cech(sc["points"], sc["radii"])
    faces = orc.raster_faces(sph, adj, off, f16=False)
    S = orc.RasterScene(sc["points"], sc["radii"], sc["density"], sc["normals"], sc["texel_sites"],
                        sc["texel_rgb"], sc["texel_height"], adj, off, faces)
    return S, sph


def _params(sc):
    return dict(points=sc["points"].astype(np.float64), radii=sc["radii"].astype(np.float64),
                nsigmas=np.concatenate([sc["normals"], sc["density"][:, None]], 1).astype(np.float64),
                texel_sites=sc["texel_sites"].astype(np.float64), texel_rgb=sc["texel_rgb"].astype(np.float64),
                texel_height=sc["texel_height"].astype(np.float64))


def _tight_texels(sc, spread=0.01, seed=11):
    """The textured scene with texel sites within a fraction of the radius of
    the cell's site: every texel weight stays far above 1e-3 of the sum, so
    the reference's texel skip never fires."""
    rng = np.random.default_rng(seed)
    out = dict(sc)
    r = sc["radii"][:, None, None]
    K = sc["texel_sites"].shape[1]
    out["texel_sites"] = (sc["points"][:, None, :] + spread * r * rng.standard_normal((len(r), K, 3))).astype(
        np.float32)
    return out


def test_raster_side_outputs_match_fp64(scene_textured):
    """normal_distance, normal, quantile depths, contrib and point_err of the
    restated forward against the FP64 envelope composite."""
    sc = scene_textured
    H = W = 32
    cam = workloads.front_camera(W, H)
    w = pack_cameras([cam])[0]
    S, sph = _rscene(sc)
    offsets, prims = orc.raster_bin(sph, w)
    rng = np.random.default_rng(2)
    gt = rng.uniform(0, 1, (H, W, 3)).astype(np.float32)
    dq = np.tile(QS, (H, W, 1))
    out = orc.raster_forward(S, w, offsets, prims, thr=1e-7, depth_q=dq, ray_gt=gt, point_err=True)
    ref = fp64.raster_render(sc, w, quantiles=QS, ray_gt=gt, thr=1e-7)
    # rays whose envelope has a near tie may composite a sliver differently
    clear = ref["margin"] > 1e-4
    assert clear.mean() > 0.95
    for key, ok, shape in (("color", "color", (-1, 3)), ("normal_distance", "ndist", (-1,)),
                           ("normal", "normal", (-1, 3))):
        got = out[ok].reshape(shape)[clear]
        want = ref[key][clear]
        err = np.abs(got - want)
        assert err.max() < 1e-4 and err.mean() < 1e-6, key
    # the normal distance is the second power of n.d: the first power differs
    assert np.abs(ref["normal_distance"]).sum() > 1.0
    qd = out["qdepth"].reshape(-1, QS.size)[clear]
    qref = ref["qdepth"][clear]
    assert np.array_equal(qd < 0, qref < 0)
    reached = qref >= 0
    assert reached.sum() > 500 and (~reached).sum() > 50
    assert np.abs(qd[reached] - qref[reached]).max() < 1e-4
    # per-cell sums over every pixel (1/(H W) scale); sliver rays included
    for key, ok in (("contrib", "contrib"), ("point_err", "perr")):
        got, want = out[ok], ref[key]
        assert np.abs(want).sum() > 0.1, key
        assert np.abs(got - want).sum() < 1e-3 * np.abs(want).sum(), key
    assert np.abs(out["contrib"] - ref["contrib"]).max() < 1e-5


def _fd_check(loss, P, an, names, cells, rng, n_scalars, idx_of=None):
    """FP64 central differences at step 1e-4 on random scalars of the given
    families, skipping scalars within a step of a jump (the two steps'
    quotients disagree) or a bend (the one-sided slopes disagree)."""
    a, fd = [], []
    tries = 0
    while len(a) < n_scalars and tries < 8 * n_scalars:
        name = names[tries % len(names)]
        tries += 1
        cell = int(rng.choice(cells))
        idx = (cell,) + tuple(int(rng.integers(0, s)) for s in P[name].shape[1:])
        if idx_of is not None:
            idx = idx_of(name, idx)
        x0 = P[name][idx]
        l0 = loss(P)
        v = {}
        for h in (1e-4, 5e-5):
            P[name][idx] = x0 + h
            v[h] = loss(P)
            P[name][idx] = x0 - h
            v[-h] = loss(P)
            P[name][idx] = x0
        central = (v[1e-4] - v[-1e-4]) / 2e-4
        half = (v[5e-5] - v[-5e-5]) / 1e-4
        right, left = (v[1e-4] - l0) / 1e-4, (l0 - v[-1e-4]) / 1e-4
        tol = abs(central) + 1e-6
        if abs(central - half) > 2e-3 * tol or abs(right - left) > 1e-2 * tol:
            continue
        a.append(float(an[name][idx]))
        fd.append(central)
    assert len(a) == n_scalars, f"only {len(a)} of {tries} scalars were away from a jump or bend"
    a, fd = np.array(a), np.array(fd)
    return np.linalg.norm(a - fd) / np.linalg.norm(fd)


def _fd_per_family(loss, P, an, names, cells, rng, per_family):
    """Relative L2 error of each gradient family on its own scalars, so a
    family with small gradients cannot hide behind a large one."""
    return {name: _fd_check(loss, P, an, [name], cells, rng, per_family) for name in names}


def _setup_chains(sc, H=24, W=24, seed=3):
    cam = workloads.front_camera(W, H, hfov_deg=30.0)
    w = pack_cameras([cam], [orc.start_cell(sc["points"], sc["radii"], cam.eye.numpy())])
    S, sph = _rscene(sc)
    offsets, prims = orc.raster_bin(sph, w[0])
    dq = np.tile(QS, (H, W, 1))
    out = orc.raster_forward(S, w[0], offsets, prims, thr=1e-7, depth_q=dq)
    rng = np.random.default_rng(seed)
    g = dict(
        target=rng.uniform(0, 1, (H, W, 3)).astype(np.float32),
        w_alpha=0.3,
        g_nd=rng.uniform(-1, 1, (H, W)),
        g_n=rng.uniform(-1, 1, (H, W, 3)),
        g_c=rng.uniform(-50, 50, len(sc["radii"])),
        g_q=rng.uniform(-0.5, 0.5, (H, W, QS.size)),
    )
    # a gradient on a quantile the forward never reached is routed to the
    # ray's last segment by the reference (rasterize.py:1409-1420 runs while
    # T > q); only reached quantiles have a depth to differentiate
    g["g_q"][out["qdepth"] < 0] = 0.0
    return w, S, offsets, prims, dq, out, g, rng


def _analytic(S, w, offsets, prims, dq, out, g, exact):
    gc = 2.0 * (out["color"] - g["target"])
    gl = -g["w_alpha"] * np.exp(out["log_t"])
    return orc.raster_backward(S, w[0], offsets, prims, out["early"], out["log_t"], gc, gl,
                               g_ndist=g["g_nd"].astype(np.float32), g_normal=g["g_n"].astype(np.float32),
                               depth_q=dq, g_qdepth=g["g_q"].astype(np.float32),
                               g_contrib=g["g_c"].astype(np.float32), exact=exact, nthreads=1)


def _loss(sc, w, dq, g, nd_pow):
    def f(P):
        return orc.raster_loss_f64(P, sc["adjacency"], sc["offsets"], w[0], g["target"], w_alpha=g["w_alpha"],
                                   g_ndist=g["g_nd"], nd_pow=nd_pow, g_normal=g["g_n"], depth_q=dq,
                                   g_qdepth=g["g_q"], g_contrib=g["g_c"])
    return f


def _an_arrays(gr):
    return dict(points=gr["spheres"][:, :3], radii=gr["spheres"][:, 3], nsigmas=gr["nsigmas"],
                texel_sites=gr["texel_sites"], texel_rgb=gr["texel_rgb"], texel_height=gr["texel_height"])


def test_raster_backward_every_chain_matches_fp64_central_differences(scene_textured):
    """Exact-texel backward with all six incoming gradients non-zero, on
    textured cells, against FP64 central differences of the functional
    (normal distance at the first power, as the reference's backward
    differentiates it): every gradient family, 12 scalars each."""
    sc = scene_textured
    w, S, offsets, prims, dq, out, g, rng = _setup_chains(sc)
    assert (out["qdepth"] >= 0).sum() > 200
    gr = _analytic(S, w, offsets, prims, dq, out, g, exact=True)
    an = _an_arrays(gr)
    touched = np.flatnonzero(np.abs(gr["nsigmas"][:, 3]) > 1e-6)
    assert touched.size > 20
    P = _params(sc)

    rel = _fd_per_family(_loss(sc, w, dq, g, nd_pow=1), P, an, list(P), touched, rng, 12)
    assert all(v < 1e-3 for v in rel.values()), rel


@pytest.mark.parametrize("chain", ["normal", "normal_distance", "quantile", "contrib"])
def test_raster_backward_single_chain_matches_fp64(scene_textured, chain):
    """One incoming gradient at a time (colour and alpha off), so an error
    in one chain cannot hide behind the colour gradient's size."""
    sc = scene_textured
    w, S, offsets, prims, dq, out, g, rng = _setup_chains(sc, seed=4)
    g["target"] = out["color"].astype(np.float32)  # zero colour gradient
    g["w_alpha"] = 0.0
    keep = dict(normal="g_n", normal_distance="g_nd", quantile="g_q", contrib="g_c")[chain]
    for k in ("g_n", "g_nd", "g_q", "g_c"):
        if k != keep:
            g[k] = np.zeros_like(g[k])
    gr = _analytic(S, w, offsets, prims, dq, out, g, exact=True)
    an = _an_arrays(gr)
    touched = np.flatnonzero(np.abs(gr["nsigmas"]).sum(1) > 1e-7)
    assert touched.size > 10
    P = _params(sc)
    names = ["points", "radii", "nsigmas"] + (["texel_height", "texel_sites"] if chain != "normal" else [])
    rel = _fd_per_family(_loss(sc, w, dq, g, nd_pow=1), P, an, names, touched, rng, 8)
    assert all(v < 1e-3 for v in rel.values()), (chain, rel)


def test_raster_normal_distance_backward_is_first_power(scene_textured):
    """The reference's normal-distance backward is the gradient of
    sum (n.d) alpha T, not of the forward's sum (n.d)^2 alpha T: the FD of
    the second-power functional disagrees with it."""
    sc = scene_textured
    w, S, offsets, prims, dq, out, g, rng = _setup_chains(sc, seed=6)
    g["target"] = out["color"].astype(np.float32)
    g["w_alpha"] = 0.0
    for k in ("g_n", "g_q", "g_c"):
        g[k] = np.zeros_like(g[k])
    gr = _analytic(S, w, offsets, prims, dq, out, g, exact=True)
    an = _an_arrays(gr)
    touched = np.flatnonzero(np.abs(gr["nsigmas"][:, :3]).sum(1) > 1e-7)
    P = _params(sc)

    def idx_of(name, idx):
        return (idx[0], int(rng.integers(0, 3)))

    rel1 = _fd_check(_loss(sc, w, dq, g, nd_pow=1), P, an, ["nsigmas"], touched, np.random.default_rng(9), 16,
                     idx_of)
    rel2 = _fd_check(_loss(sc, w, dq, g, nd_pow=2), P, an, ["nsigmas"], touched, np.random.default_rng(9), 16,
                     idx_of)
    assert rel1 < 1e-3 and rel2 > 0.1, (rel1, rel2)


def test_raster_default_backward_equals_exact_without_skipped_texels(scene_textured):
    """The default backward (the reference's texel approximations) equals the
    exact one on every array but the radius when no texel weight falls
    below 1e-3 of the sum; the radius then differs by the weights' radius
    term alone, which the exact mode adds."""
    sc = _tight_texels(scene_textured)
    w, S, offsets, prims, dq, out, g, rng = _setup_chains(sc, seed=8)
    de = _analytic(S, w, offsets, prims, dq, out, g, exact=False)
    ex = _analytic(S, w, offsets, prims, dq, out, g, exact=True)
    for k in ("nsigmas", "texel_sites", "texel_rgb", "texel_height"):
        np.testing.assert_array_equal(de[k], ex[k], err_msg=k)
    np.testing.assert_array_equal(de["spheres"][:, :3], ex["spheres"][:, :3])
    diff = np.abs(de["spheres"][:, 3] - ex["spheres"][:, 3])
    assert diff.max() > 1e-6  # the radius term is real on textured cells
    # and the exact mode's radius gradient is the FD one
    an = _an_arrays(ex)
    touched = np.flatnonzero(diff > 1e-7)
    rel = _fd_check(_loss(sc, w, dq, g, nd_pow=1), _params(sc), an, ["radii"], touched, rng, 16)
    assert rel < 1e-3, rel


def test_raster_default_backward_skips_small_texels(scene_textured):
    """With texels spread over the cell some weights fall under 1e-3 of the
    sum: the default mode drops their site/colour/height terms (texture.py:
    118-119, :158-159) and the exact mode keeps them."""
    sc = scene_textured
    w, S, offsets, prims, dq, out, g, rng = _setup_chains(sc, seed=10)
    de = _analytic(S, w, offsets, prims, dq, out, g, exact=False)
    ex = _analytic(S, w, offsets, prims, dq, out, g, exact=True)
    skipped = (de["texel_rgb"] == 0).all(-1) & (ex["texel_rgb"] != 0).any(-1)
    assert skipped.sum() > 0
    # a skipped texel's colour gradient is below 1e-3 of its cell's largest
    rows = np.nonzero(skipped)
    big = np.abs(ex["texel_rgb"][rows[0]]).max(axis=(1, 2))
    assert np.all(np.abs(ex["texel_rgb"][rows]).max(-1) < 2e-3 * big + 1e-12)
