"""GPU Morton codes are the oracle's bit for bit; morton_sort is their
stable order; a Morton-ordered scene renders the same image."""
import numpy as np
import pytest
import torch

from oracle import oracle as orc
from powerfoam import geometry
from powerfoam import raytrace as rt
from powerfoam import workloads
from powerfoam.camera import cameras_to_device
from powerfoam.power import build_power_adjacency_gpu

pytestmark = pytest.mark.gpu
DEV = "cuda"


@pytest.mark.parametrize("n,seed", [(1, 0), (1000, 1), (200000, 2)])
def test_codes_and_sort(n, seed):
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal((n, 3)) * [1.0, 0.3, 2.0]).astype(np.float32)
    p[: n // 10] = p[0]  # duplicates: ties keep their input order
    codes = geometry.morton_codes(torch.from_numpy(p).to(DEV)).cpu().numpy()
    assert np.array_equal(codes, orc.morton_codes(p))
    perm = geometry.morton_sort(torch.from_numpy(p).to(DEV))
    assert perm.dtype == torch.int32
    assert np.array_equal(perm.cpu().numpy(), np.argsort(codes, kind="stable"))


def test_sorted_scene_renders_the_same():
    sc = workloads.uniform_scene(4096, seed=5)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    cams = cameras_to_device([workloads.front_camera(96, 80)], None, DEV)
    images = []
    for sort in (False, True):
        arrs = [t(sc[k]) for k in ("points", "radii", "density", "normals", "texel_sites", "texel_rgb", "texel_height")]
        if sort:
            arrs = list(geometry.permute_cells(geometry.morton_sort(arrs[0]), *arrs))
        pts, rad, dens, nrm, sites, rgb, height = arrs
        adj, off = build_power_adjacency_gpu(pts, rad)
        s = rt.WalkScene(pts, rad, dens, nrm, sites, rgb, height, adj, off, rt.walk_faces(pts, rad, adj, off))
        c = cams.clone()
        rt.start_cells_device(pts, rad, c)
        images.append(rt.launch_walk_forward(s, c, 96, 80, want_sequences=True))
    torch.cuda.synchronize()
    a, b = images
    assert torch.equal(a["visits"], b["visits"])
    # same cells in another numbering: colour equal up to the face order of ties
    assert torch.equal(a["color"], b["color"])
