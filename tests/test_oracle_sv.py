"""Spherical-Voronoi colour pre-pass restatement (powerfoam/color_fn.py:54-184),
pinned by closed forms (out of view, one lobe, coincident lobes), by the skip
and clamp rules, and by FP64 central differences of the forward."""
import math

import numpy as np
import pytest

from oracle import oracle as orc

EYE = np.array([0.0, -3.0, 0.0], np.float32)
FWD = np.array([0.0, 1.0, 0.0], np.float32)
CUT = math.cos(math.atan(math.sqrt(2 * 0.577 ** 2)) + math.radians(10.0))


def _lobes(rng, n, dof, temp=(2.0, 6.0)):
    ax = rng.standard_normal((dof, n, 3)).astype(np.float32)
    ax /= np.linalg.norm(ax, axis=-1, keepdims=True)
    vals = (0.3 * rng.standard_normal((dof, n, 3))).astype(np.float32)
    temps = rng.uniform(*temp, (dof, n)).astype(np.float32)
    return ax, vals, temps


def _sites(rng, n):
    return (rng.uniform(-0.5, 0.5, (n, 3)) * [1, 0.2, 1]).astype(np.float32)


def test_out_of_view_is_half_grey():
    rng = np.random.default_rng(0)
    sites = np.array([[0, -4.0, 0], [3.0, -3.0, 0.0]], np.float32)  # behind / beside the eye
    ax, vals, temps = _lobes(rng, 2, 4)
    v, w = orc.sv_forward(sites, EYE, FWD, CUT, ax, vals, temps)
    assert np.all(v == 0.5) and np.all(w == 0)
    g = orc.sv_backward(sites, EYE, FWD, CUT, ax, vals, temps, w, v, np.ones((2, 3)))
    assert all(np.all(x == 0) for x in g)


def test_single_lobe_closed_form():
    rng = np.random.default_rng(1)
    n = 64
    sites = _sites(rng, n)
    ax, vals, temps = _lobes(rng, n, 1)
    v, w = orc.sv_forward(sites, EYE, FWD, CUT, ax, vals, temps, f64=True)
    d = sites.astype(np.float64) - EYE
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w_ref = np.exp(-temps[0] * np.linalg.norm(d - ax[0], axis=1))
    np.testing.assert_allclose(w, w_ref, rtol=1e-12)
    np.testing.assert_allclose(v, np.maximum(vals[0].astype(np.float64) + 0.5, 0), rtol=1e-12, atol=1e-15)


def test_coincident_lobes_average():
    rng = np.random.default_rng(2)
    n, dof = 32, 5
    sites = _sites(rng, n)
    ax, vals, temps = _lobes(rng, n, dof)
    ax[:] = ax[0]
    temps[:] = temps[0]
    v, _ = orc.sv_forward(sites, EYE, FWD, CUT, ax, vals, temps, f64=True)
    np.testing.assert_allclose(v, np.maximum(vals.astype(np.float64).mean(0) + 0.5, 0), rtol=1e-12, atol=1e-15)


def test_f32_tracks_f64():
    rng = np.random.default_rng(3)
    sites = _sites(rng, 500)
    ax, vals, temps = _lobes(rng, 500, 8)
    v32, w32 = orc.sv_forward(sites, EYE, FWD, CUT, ax, vals, temps)
    v64, w64 = orc.sv_forward(sites, EYE, FWD, CUT, ax, vals, temps, f64=True)
    assert np.abs(v32 - v64).max() < 2e-6 and np.abs(w32 / w64 - 1).max() < 2e-6


@pytest.mark.parametrize("seed", [4, 5])
def test_backward_matches_central_differences(seed):
    rng = np.random.default_rng(seed)
    n, dof = 12, 4
    sites = _sites(rng, n)
    ax, vals, temps = _lobes(rng, n, dof, temp=(0.5, 1.5))  # every lobe above the 1e-3 skip
    vals[:, :2, 0] = -2.0  # two sites clamp their red channel at 0
    g = rng.standard_normal((n, 3)).astype(np.float32)
    v, w = orc.sv_forward(sites, EYE, FWD, CUT, ax, vals, temps, f64=True)
    assert np.all(v[:2, 0] == 0)
    ga, gv, gt = orc.sv_backward(sites, EYE, FWD, CUT, ax, vals, temps, w, v, g, f64=True)

    def loss(a, b, c):
        # the lobe arrays are f32; a step of 2^-12 around them is exact in f32
        val, _ = orc.sv_forward(sites, EYE, FWD, CUT, a, b, c, f64=True)
        return float((val * g).sum())

    h = 2.0 ** -12
    for arr, grad in ((ax, ga), (vals, gv), (temps, gt)):
        flat, gflat = arr.reshape(-1), grad.reshape(-1)
        idx = rng.choice(flat.size, 24, replace=False)
        for k in idx:
            old = flat[k]
            flat[k] = old + h
            lp = loss(ax, vals, temps)
            flat[k] = old - h
            lm = loss(ax, vals, temps)
            flat[k] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - gflat[k]) <= 1e-5 * max(1.0, abs(fd)), (k, fd, gflat[k])


def test_skip_rule_zeroes_negligible_lobes():
    rng = np.random.default_rng(6)
    n = 16
    sites = _sites(rng, n)
    ax, vals, temps = _lobes(rng, n, 2, temp=(1.0, 1.0))
    vals = np.abs(vals)  # no clamped channel: every lobe-0 gradient is nonzero
    d = sites - EYE
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ax[0] = d  # lobe 0 points at the site: weight 1
    ax[1] = -d  # lobe 1 points away, temperature 20: weight exp(-40)
    temps[1] = 20.0
    v, w = orc.sv_forward(sites, EYE, FWD, CUT, ax, vals, temps)
    ga, gv, gt = orc.sv_backward(sites, EYE, FWD, CUT, ax, vals, temps, w, v, np.ones((n, 3), np.float32))
    assert np.all(ga[1] == 0) and np.all(gv[1] == 0) and np.all(gt[1] == 0)
    assert np.all(gv[0] != 0)


def test_planted_sign_error_is_caught():
    """Flipping dp - dp_out (a plausible slip) breaks the FD agreement."""
    rng = np.random.default_rng(7)
    n, dof = 6, 3
    sites = _sites(rng, n)
    ax, vals, temps = _lobes(rng, n, dof, temp=(0.5, 1.5))
    g = rng.standard_normal((n, 3)).astype(np.float32)
    v, w = orc.sv_forward(sites, EYE, FWD, CUT, ax, vals, temps, f64=True)
    _, _, gt = orc.sv_backward(sites, EYE, FWD, CUT, ax, vals, temps, w, v, g, f64=True)
    h = 2.0 ** -12
    k = 0
    t2 = temps.copy().reshape(-1)
    t2[k] += h
    lp = float((orc.sv_forward(sites, EYE, FWD, CUT, ax, vals, t2.reshape(temps.shape), f64=True)[0] * g).sum())
    t2[k] -= 2 * h
    lm = float((orc.sv_forward(sites, EYE, FWD, CUT, ax, vals, t2.reshape(temps.shape), f64=True)[0] * g).sum())
    fd = (lp - lm) / (2 * h)
    assert abs(fd - gt.reshape(-1)[k]) < 1e-6 * max(1, abs(fd))
    assert abs(fd + gt.reshape(-1)[k]) > 1e-3 * abs(fd)
