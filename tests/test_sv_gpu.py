"""Spherical-Voronoi colour pre-pass on the GPU (through the C ABI) against the
CPU restatement: f32 bit-exact, f16 storage equal to the restatement's f32
value (on the f16-rounded lobes) rounded to f16; autograd through
SphericalVoronoi equals the restatement's backward bit for bit."""
import types

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from powerfoam import workloads
from powerfoam.color_fn import SphericalVoronoi, camera_frame

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _case(seed, n=4096, dof=6):
    rng = np.random.default_rng(seed)
    sites = rng.uniform(-1.0, 1.0, (n, 3)).astype(np.float32)
    sites[: n // 8, 1] -= 4.0  # behind the eye at y = -3: out of view
    ax = rng.standard_normal((dof, n, 3)).astype(np.float32)
    ax /= np.linalg.norm(ax, axis=-1, keepdims=True)
    vals = (0.4 * rng.standard_normal((dof, n, 3))).astype(np.float32)
    temps = rng.uniform(0.5, 8.0, (dof, n)).astype(np.float32)
    cam = workloads.front_camera(64, 48, hfov_deg=50.0)
    return sites, ax, vals, temps, cam


@pytest.mark.parametrize("seed", [0, 1])
def test_sv_forward_backward_f32_bit_exact(seed):
    sites, ax, vals, temps, cam = _case(seed)
    sv = SphericalVoronoi(types.SimpleNamespace(sv_dof=ax.shape[0]), DEV)
    t = lambda a: torch.from_numpy(a).to(DEV)
    A, V, T = (t(a).requires_grad_(True) for a in (ax, vals, temps))
    out = sv.forward(t(sites), cam, A, V, T)
    eye, fwd, cut = camera_frame(cam)
    v_ref, w_ref = orc.sv_forward(sites, eye.numpy(), fwd.numpy(), cut, ax, vals, temps)
    assert 0 < int((w_ref == 0).sum()) < len(w_ref)  # some sites out of view
    assert np.array_equal(out.detach().cpu().numpy(), v_ref)
    g = np.random.default_rng(seed + 10).standard_normal(v_ref.shape).astype(np.float32)
    out.backward(t(g))
    ga, gv, gt = orc.sv_backward(sites, eye.numpy(), fwd.numpy(), cut, ax, vals, temps, w_ref, v_ref, g)
    assert np.array_equal(A.grad.cpu().numpy(), ga)
    assert np.array_equal(V.grad.cpu().numpy(), gv)
    assert np.array_equal(T.grad.cpu().numpy(), gt)


def test_sv_half_storage():
    sites, ax, vals, temps, cam = _case(2)
    h = lambda a: torch.from_numpy(a).to(DEV).half()
    sv = SphericalVoronoi(types.SimpleNamespace(sv_dof=ax.shape[0]), DEV, attr_dtype="half")
    out = sv.forward(torch.from_numpy(sites).to(DEV), cam, h(ax), h(vals), h(temps))
    assert out.dtype == torch.float16
    eye, fwd, cut = camera_frame(cam)
    q = lambda a: torch.from_numpy(a).half().float().numpy()
    v_ref, _ = orc.sv_forward(sites, eye.numpy(), fwd.numpy(), cut, q(ax), q(vals), q(temps))
    want = torch.from_numpy(v_ref).half().float().numpy()  # RNE to f16, as the kernel stores
    assert np.array_equal(out.float().cpu().numpy(), want)


def test_sv_rejects_bad_shapes():
    sites, ax, vals, temps, cam = _case(3, n=64)
    sv = SphericalVoronoi(types.SimpleNamespace(sv_dof=ax.shape[0]), DEV)
    t = lambda a: torch.from_numpy(a).to(DEV)
    with pytest.raises(ValueError):
        sv.forward(t(sites), cam, t(ax[:, :32]), t(vals), t(temps))
    with pytest.raises(ValueError):
        SphericalVoronoi(types.SimpleNamespace(sv_dof=2), DEV, attr_dtype="bf16")
