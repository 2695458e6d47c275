"""Synthetic scenes and cameras for the benchmark workloads.

All randomness comes from ``numpy.random.default_rng(seed)`` drawn in a fixed
order (sites, weights, density, colour, camera), so the CPU oracle and the
GPU see identical inputs.  A power weight w is read as w = r^2 (the
reference's power is |x - p|^2 - r^2, powerfoam/rendering_math.py:79-84), so
radii = sqrt(w).  Cells carry no dipole detail: normals (0, 0, 1), all K
texel sites at the cell site, texel heights 2r (the displaced plane then
lies outside the sphere and never clips), every texel the cell's colour.

Workloads:
  * ``uniform_4k``: 4,096 sites uniform in [-1,1]^3, w ~ U[0, 1e-2],
    density ~ lognormal(0, 1), RGB ~ U[0,1] (SH degree 0), one 128x128
    camera at (0, -3, 0) looking at the origin, 60 deg horizontal FOV.
  * ``gmm``: sites from an 8-component Gaussian mixture (centres U[-1,1]^3,
    sigma 0.2), w ~ U[0, 1e-4], density ~ lognormal(0, 1), SH degree 3
    coefficients N(0, 0.1) (48 per cell), cameras on the radius-3 sphere.
"""
import numpy as np
import torch

K_TEXELS = 8


def _fill(points, radii, density, rgb, K=K_TEXELS):
    n = points.shape[0]
    normals = np.zeros((n, 3), np.float32)
    normals[:, 2] = 1.0
    texel_sites = np.repeat(points[:, None, :], K, axis=1).astype(np.float32)
    texel_height = np.repeat((2.0 * radii)[:, None], K, axis=1).astype(np.float32)
    texel_rgb = None if rgb is None else np.repeat(rgb[:, None, :], K, axis=1).astype(np.float32)
    return normals, texel_sites, texel_height, texel_rgb


def uniform_scene(n=4096, seed=0, w_max=1e-2):
    rng = np.random.default_rng(seed)
    points = rng.uniform(-1.0, 1.0, (n, 3)).astype(np.float32)
    radii = np.sqrt(rng.uniform(0.0, w_max, n)).astype(np.float32)
    density = rng.lognormal(0.0, 1.0, n).astype(np.float32)
    rgb = rng.uniform(0.0, 1.0, (n, 3)).astype(np.float32)
    normals, sites, height, texel_rgb = _fill(points, radii, density, rgb)
    return dict(points=points, radii=radii, density=density, normals=normals, texel_sites=sites,
                texel_rgb=texel_rgb, texel_height=height, sh=None, rng=rng)


def gmm_scene(n, seed=0, w_max=1e-4, sh_degree=3):
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-1.0, 1.0, (8, 3))
    comp = rng.integers(0, 8, n)
    points = (centres[comp] + 0.2 * rng.standard_normal((n, 3))).astype(np.float32)
    radii = np.sqrt(rng.uniform(0.0, w_max, n)).astype(np.float32)
    density = rng.lognormal(0.0, 1.0, n).astype(np.float32)
    nco = (sh_degree + 1) ** 2
    sh = (0.1 * rng.standard_normal((n, nco, 3))).astype(np.float32)
    normals, sites, height, _ = _fill(points, radii, density, None)
    return dict(points=points, radii=radii, density=density, normals=normals, texel_sites=sites, texel_rgb=None,
                texel_height=height, sh=sh, rng=rng)


def sphere_cameras(rng, count, width, height, radius=3.0, hfov_deg=60.0):
    from .camera import look_at_camera

    cams = []
    for _ in range(count):
        v = rng.standard_normal(3)
        v /= np.linalg.norm(v)
        eye = radius * v
        up = np.array([0.0, 0.0, 1.0]) if abs(v[2]) < 0.95 else np.array([0.0, 1.0, 0.0])
        cams.append(look_at_camera(eye, [0.0, 0.0, 0.0], up, width, height, hfov_deg))
    return cams


def front_camera(width=128, height=128, hfov_deg=60.0):
    from .camera import look_at_camera

    return look_at_camera([0.0, -3.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 1.0], width, height, hfov_deg)


def to_torch(scene, device):
    out = {}
    for k, v in scene.items():
        if isinstance(v, np.ndarray):
            out[k] = torch.from_numpy(v).to(device)
        else:
            out[k] = v
    return out
