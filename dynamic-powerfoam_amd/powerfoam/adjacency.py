"""Power-diagram adjacency (scene preparation, host side).

``build_power_adjacency`` restates ``benchmark.py:25-83`` of the reference:
lift each site to 4-D (x, y, z, |p|^2 - r^2), take the convex hull with
Qhull (scipy), keep the lower-hull simplices (weighted Delaunay tetrahedra),
take their six edges, deduplicate exactly as ``torch.unique(edges, dim=0)``
does (rows are not canonicalised, so (a, b) and (b, a) can both survive),
direct them both ways and build a CSR sorted by source.  The one deliberate
difference: the reference's ``torch.argsort`` by source is unstable, so its
neighbour order within a row is unspecified; here the sort is stable, which
fixes the order (and so the tie-breaking of the walk's strict argmin).

This is offline scene preparation on the CPU, as in the reference; the hot
path consumes its CSR on the GPU.
"""
import numpy as np
import torch


def build_power_adjacency(points: torch.Tensor, radii: torch.Tensor, alpha_complex=False):
    from scipy.spatial import ConvexHull

    norms = points.detach().norm(dim=-1)
    weights = norms**2 - radii.detach() ** 2
    lifted = torch.cat([points.detach(), weights[:, None]], dim=-1).cpu().numpy()
    hull = ConvexHull(lifted)
    simp = hull.simplices[hull.equations[:, 3] < 0].astype(np.int64)
    n = points.shape[0]
    pairs = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
    a = np.concatenate([simp[:, i] for i, _ in pairs])
    b = np.concatenate([simp[:, j] for _, j in pairs])
    code = np.unique(a * n + b)  # == torch.unique(dim=0): sorted unique rows
    ua, ub = code // n, code % n
    if alpha_complex:
        p = points.detach().cpu().numpy().astype(np.float32)
        r = radii.detach().cpu().numpy().astype(np.float32)
        d = np.linalg.norm(p[ua] - p[ub], axis=-1)
        keep = d < (r[ua] + r[ub])
        ua, ub = ua[keep], ub[keep]
    src = np.concatenate([ua, ub])
    dst = np.concatenate([ub, ua])
    order = np.argsort(src, kind="stable")
    adjacency = dst[order].astype(np.int32)
    counts = np.bincount(src, minlength=n)
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    dev = points.device
    return torch.from_numpy(adjacency).to(dev), torch.from_numpy(offsets).to(dev)


def csr_is_symmetric(adjacency, offsets):
    """Check the CSR invariants the walk relies on: every edge i->j has j->i
    and there are no self loops.  Returns (ok, message)."""
    adj = np.asarray(adjacency.cpu() if isinstance(adjacency, torch.Tensor) else adjacency, dtype=np.int64)
    off = np.asarray(offsets.cpu() if isinstance(offsets, torch.Tensor) else offsets, dtype=np.int64)
    n = off.size - 1
    src = np.repeat(np.arange(n), np.diff(off))
    if np.any(src == adj):
        return False, "self loop"
    fwd = set(zip(src.tolist(), adj.tolist()))
    for s, d in fwd:
        if (d, s) not in fwd:
            return False, f"edge {s}->{d} has no reverse"
    return True, "ok"
