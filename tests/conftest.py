import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "dynamic-powerfoam_amd")
for p in (PKG, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


def gpu_available():
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def scene4k():
    """The 4,096-cell uniform scene with its power-diagram CSR (numpy)."""
    import torch
    from powerfoam import workloads
    from powerfoam.adjacency import build_power_adjacency

    sc = workloads.uniform_scene(4096, seed=0)
    adj, off = build_power_adjacency(torch.from_numpy(sc["points"]), torch.from_numpy(sc["radii"]))
    sc["adjacency"] = adj.numpy()
    sc["offsets"] = off.numpy()
    return sc


@pytest.fixture(scope="session")
def scene_textured():
    """512 cells with general dipole detail: random normals, texel sites
    spread over the cell, random heights and colours (exercises the plane
    clip and the texel interpolation of raytrace.py:42-113)."""
    import torch
    from powerfoam import workloads
    from powerfoam.adjacency import build_power_adjacency

    rng = np.random.default_rng(7)
    n, K = 512, 8
    pts = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    radii = np.sqrt(rng.uniform(0.01, 0.06, n)).astype(np.float32)
    nrm = rng.standard_normal((n, 3))
    nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    sites = (pts[:, None, :] + 0.5 * radii[:, None, None] * rng.standard_normal((n, K, 3))).astype(np.float32)
    heights = (radii[:, None] * rng.uniform(-0.5, 0.5, (n, K))).astype(np.float32)
    rgb = rng.uniform(0, 1, (n, K, 3)).astype(np.float32)
    dens = rng.lognormal(0.5, 1.0, n).astype(np.float32)
    adj, off = build_power_adjacency(torch.from_numpy(pts), torch.from_numpy(radii))
    return dict(points=pts, radii=radii, density=dens, normals=nrm, texel_sites=sites, texel_rgb=rgb,
                texel_height=heights, adjacency=adj.numpy(), offsets=off.numpy())
