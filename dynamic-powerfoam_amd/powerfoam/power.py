"""Power-diagram adjacency built on the GPU: the graph
``build_power_adjacency`` (benchmark.py:25-83) obtains from a Qhull hull of
the lifted sites, from two gfx950 kernels instead.

1. ``pf_power_cells``: every site's power cell clipped to the sites'
   bounding box B (with a margin), in FP64: first by its 32 nearest sites,
   nearest first, then by any site found in the empty-sphere ball of a
   vertex, until every vertex is certified.  This gives every face that
   meets B.
2. ``pf_power_hull``: outside B only the sites whose cells reach B's
   boundary own points, so their faces outside B are faces of the diagram
   of those few sites alone; a brute-force pass over them adds the faces
   with a vertex outside B.

The bisectors use the lifted coordinate exactly as the reference computes
it (``points.norm(dim=-1)**2 - radii**2`` in f32, on the points' device),
so the diagram is the one Qhull triangulates.  The result equals the Qhull
adjacency as a set of edges, except for faces farther than ``far`` times
the scene extent and for degenerate (co-spherical) configurations, which
Qhull triangulates one way and the clipping may resolve another.  Each row
lists a neighbour once, in ascending order; the reference's rows list a
neighbour twice when its tetrahedra hold the edge in both orientations
(benchmark.py:47-66), which the walk reads as the same face twice.
"""
import ctypes

import torch

from . import _lib

MAX_NEIGHBOURS = 58  # the kernels keep at most 64 planes, six of them the box


def lifted_coordinate(points, radii):
    """L = |p|^2 - r^2 as benchmark.py:29-32 computes it (f32)."""
    norms = points.norm(dim=-1)
    return (norms ** 2 - radii ** 2).contiguous()


def grid_shape(points, lo, hi, sites_per_cell, max_cells=1 << 25, max_dim=1 << 20):
    """Cell size and dims of the bucket grid.  The size is set so that a site
    shares its bucket with about ``sites_per_cell`` sites on average over the
    sites (not over the buckets): clustered scenes get buckets sized for
    their clusters, and the empty space costs empty buckets only."""
    N = points.shape[0]
    spans = [max(h - l, 1e-6) for l, h in zip(lo, hi)]
    cell = max((spans[0] * spans[1] * spans[2] / N * sites_per_cell) ** (1.0 / 3.0), 1e-6)
    origin = torch.tensor(lo, dtype=torch.float32, device=points.device)

    def shape(c):
        d = [int(s // c) + 1 for s in spans]
        while d[0] * d[1] * d[2] > max_cells or max(d) > max_dim:
            c *= 1.25
            d = [int(s // c) + 1 for s in spans]
        return c, d

    for _ in range(2):
        cell, dims = shape(cell)
        top = torch.tensor(dims, device=points.device) - 1
        idx = torch.minimum(torch.floor((points - origin) / cell).long().clamp_min(0), top)
        key = (idx[:, 2] * dims[1] + idx[:, 1]) * dims[0] + idx[:, 0]
        occupancy = float(torch.bincount(key, minlength=dims[0] * dims[1] * dims[2])[key].double().mean().item())
        cell *= (sites_per_cell / max(occupancy, 1.0)) ** (1.0 / 3.0)
    return shape(cell)


def build_power_adjacency_gpu(points, radii, margin=None, sites_per_cell=3.0, far=1e4):
    """(adjacency int32 (E,), offsets int32 (N+1,)) on points' device."""
    if points.device.type != "cuda":
        raise ValueError("Only CUDA tensors are supported as inputs")
    lib = _lib.load()
    dev = points.device
    pts = points.detach().to(torch.float32).contiguous()
    rad = radii.detach().to(torch.float32).contiguous()
    N = pts.shape[0]
    if pts.dim() != 2 or pts.shape[1] != 3 or tuple(rad.shape) != (N,) or rad.device != dev:
        raise ValueError("points must be (N, 3) and radii (N,) on one device")
    # the bucket grid is sized from the sites' bounds
    if N and not bool(torch.isfinite(pts).all() and torch.isfinite(rad).all()):
        raise ValueError("points and radii must be finite")
    if N < 2:
        return torch.zeros(0, dtype=torch.int32, device=dev), torch.zeros(N + 1, dtype=torch.int32, device=dev)
    L = lifted_coordinate(pts, rad)
    exact = (pts.double() ** 2).sum(-1) - rad.double() ** 2
    slack = float((L.double() - exact).abs().max().item()) * 2.0 + 1e-12
    lo, hi = pts.min(0).values, pts.max(0).values
    lo_h, hi_h = lo.cpu().tolist(), hi.cpu().tolist()
    rmax = float(rad.max().item())
    ext = max(h - l for l, h in zip(lo_h, hi_h))
    m = margin if margin is not None else max(0.1 * ext, 4.0 * rmax, 1e-6)
    box = torch.tensor([l - m for l in lo_h] + [h + m for h in hi_h], dtype=torch.float32)
    cell, dims = grid_shape(pts, lo_h, hi_h, sites_per_cell)
    g = _lib.PfCechGrid()
    for k in range(3):
        g.origin[k] = lo_h[k]
        g.dims[k] = dims[k]
    g.cell = cell
    G = dims[0] * dims[1] * dims[2]
    gcount = torch.zeros(G + 1, dtype=torch.int32, device=dev)
    gfill = torch.zeros(G, dtype=torch.int32, device=dev)
    key = torch.empty(N, dtype=torch.int32, device=dev)
    gstart = torch.empty(G + 1, dtype=torch.int64, device=dev)
    items = torch.empty(N, dtype=torch.int32, device=dev)
    ws_bytes = int(lib.pf_power_workspace_bytes(ctypes.byref(g)))
    if ws_bytes < 0:
        raise ValueError(f"bucket grid {dims} out of range")
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    lists = torch.empty(N, MAX_NEIGHBOURS, dtype=torch.int32, device=dev)
    counts = torch.empty(N, dtype=torch.int32, device=dev)
    flags = torch.empty(N, dtype=torch.int32, device=dev)
    P = _lib.ptr
    rc = lib.pf_power_cells(P(pts), P(L), N, ctypes.byref(g), P(gcount), P(gfill), P(key), P(gstart), P(items),
                            P(ws), ws_bytes, box.data_ptr(), rmax, slack, MAX_NEIGHBOURS, P(lists), P(counts), P(flags),
                            _lib.stream_ptr())
    _lib.check(rc, "pf_power_cells")
    if bool(((flags & 2) != 0).any().item()):
        raise RuntimeError("power cell with more than 58 neighbours: adjacency capacity exceeded")
    hull = torch.nonzero(flags & 1).flatten().to(torch.int32).contiguous()
    nh = hull.numel()
    hlists = torch.empty(max(nh, 1), MAX_NEIGHBOURS, dtype=torch.int32, device=dev)
    hcounts = torch.zeros(max(nh, 1), dtype=torch.int32, device=dev)
    hflags = torch.zeros(max(nh, 1), dtype=torch.int32, device=dev)
    rc = lib.pf_power_hull(P(pts), P(L), P(hull) if nh else None, nh, box.data_ptr(), float(far * ext),
                           MAX_NEIGHBOURS, P(hlists), P(hcounts), P(hflags), _lib.stream_ptr())
    _lib.check(rc, "pf_power_hull")
    if nh and bool(((hflags[:nh] & 2) != 0).any().item()):
        raise RuntimeError("hull cell with more than 58 neighbours: adjacency capacity exceeded")
    cols = torch.arange(MAX_NEIGHBOURS, device=dev)
    src = torch.repeat_interleave(torch.arange(N, device=dev), counts.long())
    dst = lists[cols[None, :] < counts[:, None]].long()
    if nh:
        src = torch.cat([src, torch.repeat_interleave(hull.long(), hcounts[:nh].long())])
        dst = torch.cat([dst, hlists[:nh][cols[None, :] < hcounts[:nh, None]].long()])
    keys = torch.unique(torch.cat([src * N + dst, dst * N + src]))
    src, dst = keys // N, keys % N
    offsets = torch.zeros(N + 1, dtype=torch.int64, device=dev)
    offsets[1:] = torch.cumsum(torch.bincount(src, minlength=N), 0)
    return dst.to(torch.int32).contiguous(), offsets.to(torch.int32).contiguous()
