"""GPU power-diagram adjacency against the Qhull lift (benchmark.py:25-83
restated in powerfoam/adjacency.py): identical edge sets on uniform and
clustered scenes (hidden sites included), and walks through either CSR
visit the same cells and render the same images, though the lift's rows
repeat neighbours and the GPU rows do not."""
import numpy as np
import pytest
import torch

from powerfoam import raytrace as rt
from powerfoam import workloads
from powerfoam.adjacency import build_power_adjacency, csr_is_symmetric
from powerfoam.camera import cameras_to_device
from powerfoam.power import build_power_adjacency_gpu

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _edges(adj, off):
    adj = np.asarray(adj.cpu()).astype(np.int64)
    off = np.asarray(off.cpu()).astype(np.int64)
    n = off.size - 1
    return np.unique(np.repeat(np.arange(n), np.diff(off)) * n + adj)


@pytest.mark.parametrize("kind,n,seed", [("uniform", 4096, 0), ("uniform", 2048, 3), ("gmm", 20000, 1),
                                         ("gmm", 200000, 2)])
def test_same_edges_as_qhull(kind, n, seed):
    sc = workloads.uniform_scene(n, seed=seed) if kind == "uniform" else workloads.gmm_scene(n, seed=seed)
    pts, rad = torch.from_numpy(sc["points"]).to(DEV), torch.from_numpy(sc["radii"]).to(DEV)
    # both lift on the GPU, as benchmark.py:29-31 does with CUDA tensors: a
    # lift computed on the host can differ in the last bit of |p|, which moves
    # faces of near-degenerate configurations
    q_adj, q_off = build_power_adjacency(pts, rad)
    g_adj, g_off = build_power_adjacency_gpu(pts, rad)
    assert g_adj.dtype == torch.int32 and g_off.dtype == torch.int32 and g_off.numel() == n + 1
    assert np.array_equal(_edges(g_adj, g_off), _edges(q_adj, q_off))
    # the lift's CSR lists an edge twice when tetrahedra hold it in both
    # orientations (benchmark.py:47-66); the GPU rows hold each neighbour once
    assert g_adj.numel() == _edges(q_adj, q_off).size <= q_adj.numel()
    ok, msg = csr_is_symmetric(g_adj, g_off)
    assert ok, msg
    rows = np.split(g_adj.cpu().numpy(), g_off.cpu().numpy()[1:-1])
    assert all(np.all(np.diff(r) > 0) for r in rows)  # rows sorted ascending


def test_walks_agree(scene4k):
    sc = scene4k
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    pts, rad = t(sc["points"]), t(sc["radii"])
    outs = []
    for adj, off in ((t(sc["adjacency"]), t(sc["offsets"])), build_power_adjacency_gpu(pts, rad)):
        faces = rt.walk_faces(pts, rad, adj, off)
        s = rt.WalkScene(pts, rad, t(sc["density"]), t(sc["normals"]), t(sc["texel_sites"]), t(sc["texel_rgb"]),
                         t(sc["texel_height"]), adj, off, faces)
        cams = cameras_to_device([workloads.front_camera(96, 80)], None, DEV)
        rt.start_cells_device(pts, rad, cams)
        outs.append(rt.launch_walk_forward(s, cams, 96, 80, want_sequences=True))
    torch.cuda.synchronize()
    a, b = outs
    assert torch.equal(a["visits"], b["visits"]) and torch.equal(a["seq_hash"], b["seq_hash"])
    assert torch.equal(a["color"], b["color"])


def test_rejects_host_tensors():
    with pytest.raises(ValueError):
        build_power_adjacency_gpu(torch.zeros(4, 3), torch.zeros(4))


def test_rejects_mismatched_or_non_finite_sites():
    pts = torch.rand(64, 3, device=DEV)
    with pytest.raises(ValueError, match="radii"):
        build_power_adjacency_gpu(pts, torch.full((63,), 0.01, device=DEV))
    with pytest.raises(ValueError, match="points"):
        build_power_adjacency_gpu(pts[:, :2].contiguous(), torch.full((64,), 0.01, device=DEV))
    bad = pts.clone()
    bad[5, 1] = float("nan")
    with pytest.raises(ValueError, match="finite"):
        build_power_adjacency_gpu(bad, torch.full((64,), 0.01, device=DEV))
