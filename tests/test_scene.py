"""Scene glue (powerfoam/scene.py:239-337, 772-821) on the CPU: activations
against independent closed forms (scipy's rotation matrices, the softplus
definition), the lobe-major colour layout against explicit indexing,
world texel positions, the checkpoint round trip with a tensors-only load,
and the launcher's error conventions (warp_interop.py:14-15)."""
import types

import numpy as np
import pytest
import torch
from scipy.spatial.transform import Rotation

from powerfoam.scene import CHECKPOINT_KEYS, PowerfoamScene
from powerfoam.warp_interop import launch_kernel_from_torch

ARGS = types.SimpleNamespace(num_texel_sites=8, sv_dof=3, render_objective="volume", is_pinhole=True)


def _scene(n=37, seed=0):
    g = torch.Generator().manual_seed(seed)
    K, D = ARGS.num_texel_sites, ARGS.sv_dof
    s = PowerfoamScene(ARGS)
    t = dict(points=torch.randn(n, 3, generator=g), radii=0.05 * torch.randn(n, generator=g),
             density=0.02 * torch.randn(n, generator=g), quaternions=torch.randn(n, 4, generator=g),
             texel_sites=0.1 * torch.randn(n, K, 2, generator=g),
             texel_sv_axis=2 * torch.randn(n, K, 3 * D, generator=g),
             texel_sv_rgb=0.3 * torch.randn(n, K, 3 * D, generator=g), texel_height=torch.randn(n, K, generator=g))
    s._set_params(t, "cpu")
    s.adjacency = torch.arange(n, dtype=torch.int32)
    s.adjacency_offsets = torch.arange(n + 1, dtype=torch.int32)
    return s


def test_softplus_beta_100():
    s = _scene()
    x = s.radii.detach().double()
    np.testing.assert_allclose(s.get_radii().detach().double(), np.log1p(np.exp(100 * x)) / 100, rtol=1e-6,
                               atol=1e-7)
    assert torch.all(s.get_density() > 0)


def test_frame_is_rotation_matrix_rows():
    s = _scene()
    q = (s.quaternions / s.quaternions.norm(dim=-1, keepdim=True)).detach().double().numpy()
    R = Rotation.from_quat(q[:, [1, 2, 3, 0]]).as_matrix()  # scipy order (x, y, z, w)
    n = s.get_normals().detach().double().numpy()
    t, b = (a.detach().double().numpy() for a in s.get_tangents())
    np.testing.assert_allclose(n, R[:, 0, :], atol=1e-6)
    np.testing.assert_allclose(t, R[:, 1, :], atol=1e-6)
    np.testing.assert_allclose(b, R[:, 2, :], atol=1e-6)
    np.testing.assert_allclose(np.cross(n, t), b, atol=1e-5)  # right-handed orthonormal frame


def test_att_sv_lobe_major_layout():
    s = _scene(n=5)
    K, D = ARGS.num_texel_sites, ARGS.sv_dof
    axis, rgb, temp = (a.detach().numpy() for a in s.get_att_sv())
    ax_raw = s.texel_sv_axis.detach().numpy().reshape(5, K, D, 3)
    rgb_raw = s.texel_sv_rgb.detach().numpy().reshape(5, K, D, 3)
    for i in range(5):
        for k in range(K):
            for j in range(D):
                v = ax_raw[i, k, j]
                np.testing.assert_allclose(temp[j, i * K + k], np.linalg.norm(v), rtol=1e-6)
                np.testing.assert_allclose(axis[j, i * K + k], v / np.linalg.norm(v), rtol=1e-5, atol=1e-6)
                np.testing.assert_array_equal(rgb[j, i * K + k], rgb_raw[i, k, j])


def test_world_texels():
    s = _scene(n=4)
    sites, h = (a.detach().numpy() for a in s.world_texels())
    r = s.get_radii().detach().numpy()
    t, b = (a.detach().numpy() for a in s.get_tangents())
    p = s.points.detach().numpy()
    uv = s.texel_sites.detach().numpy()
    want = p[:, None] + r[:, None, None] * (uv[..., :1] * t[:, None] + uv[..., 1:] * b[:, None])
    np.testing.assert_allclose(sites, want, rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(h, s.texel_height.detach().numpy() * r[:, None], rtol=1e-6)


def test_checkpoint_round_trip(tmp_path):
    s = _scene()
    path = tmp_path / "model.pt"
    s.save_pt(path)
    data = torch.load(path, weights_only=True)
    assert set(data) == set(CHECKPOINT_KEYS)
    s2 = PowerfoamScene(ARGS).load_pt(path, device="cpu")
    for k in CHECKPOINT_KEYS[:8]:
        assert torch.equal(getattr(s2, k).detach(), getattr(s, k).detach())
    assert torch.equal(s2.adjacency, s.adjacency) and torch.equal(s2.adjacency_offsets, s.adjacency_offsets)


def test_checkpoint_missing_key_rejected(tmp_path):
    path = tmp_path / "bad.pt"
    torch.save({"points": torch.zeros(2, 3)}, path)
    with pytest.raises(ValueError):
        PowerfoamScene(ARGS).load_pt(path, device="cpu")


def test_unknown_dtype_rejected():
    with pytest.raises(ValueError):
        PowerfoamScene(ARGS, attr_dtype="bf16")


def test_launcher_rejects_host_tensors_and_unknown_names():
    with pytest.raises(ValueError, match="Only CUDA tensors"):
        launch_kernel_from_torch("cuda", "pf_cech_bounds", 1, [torch.zeros(3, 3), torch.zeros(3), 3], 256)
    with pytest.raises(ValueError):
        launch_kernel_from_torch("cuda", "not_a_kernel", 1, [], 256)
