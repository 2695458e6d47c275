"""Čech-complex adjacency — the drop-in for ``powerfoam/bvh.py``'s ``AABBTree``.

``AABBTree(device)``, ``.update(centers, radii)`` and
``.build_cech_complex() -> (adjacency int32 (E,), offsets int32 (N+1,))``
keep the reference's interface (bvh.py:72-123).  The Warp BVH is replaced by
a grid of levels built and queried by gfx950 kernels (``csrc/cech.hip``):
level 0's cells are twice the mean radius, each level doubles them until
they hold the largest sphere, and a sphere is listed in the finest level
whose cells hold it, so a few large spheres do not coarsen the grid for all
(a single grid sized by the largest radius crowds the small spheres of a
heavy-tailed distribution into few cells and turns the build quadratic).
Like the reference, the build reads the total edge count back to the host
once to size the CSR (bvh.py:110).
"""
import ctypes
import math

import torch

from . import _lib

_MAX_GRID_CELLS = 1 << 25
_MAX_LEVELS = 16


class AABBTree:
    def __init__(self, device):
        self.device = device
        _lib.load()

    def update(self, centers, radii):
        if centers.device.type != "cuda":
            raise ValueError("Only CUDA tensors are supported as inputs")
        self.centers = centers.detach().to(torch.float32).contiguous()
        self.radii = radii.detach().to(torch.float32).contiguous()
        lib = _lib.load()
        ws = torch.empty(8, dtype=torch.int32, device=self.centers.device)
        bounds = torch.empty(8, dtype=torch.float32, device=self.centers.device)
        _lib.check(lib.pf_cech_bounds(_lib.ptr(self.centers), _lib.ptr(self.radii), self.centers.shape[0],
                                      _lib.ptr(ws), _lib.ptr(bounds), _lib.stream_ptr()), "pf_cech_bounds")
        b = bounds.cpu().tolist()
        lo, hi, rmax, rmean = b[0:3], b[3:6], b[6], b[7]
        top = max(2.0 * rmax * 1.001, 1e-6)  # the last level holds the largest sphere
        cell = min(max(2.0 * rmean * 1.001, 1e-6), top)
        ext = [max(hi[k] - lo[k], 0.0) for k in range(3)]
        while True:
            dims = [int(math.floor(e / cell)) + 1 for e in ext]
            levels = 1
            while cell * 2 ** (levels - 1) < top:
                levels += 1
            if dims[0] * dims[1] * dims[2] <= _MAX_GRID_CELLS and levels <= _MAX_LEVELS:
                break
            cell *= 1.5
        g = _lib.PfCechGrid()
        for k in range(3):
            g.origin[k] = lo[k]
            g.dims[k] = dims[k]
        g.cell = cell
        g.levels = levels
        self.grid = g

    def build_cech_complex(self):
        lib = _lib.load()
        dev = self.centers.device
        n = self.centers.shape[0]
        g = self.grid
        G = lib.pf_cech_grid_size(ctypes.byref(g))
        if G <= 0:
            raise ValueError("Čech grid out of range")
        gcount = torch.zeros(G + 1, dtype=torch.int32, device=dev)
        gfill = torch.zeros(G, dtype=torch.int32, device=dev)
        key = torch.empty(n, dtype=torch.int32, device=dev)
        gstart = torch.empty(G + 1, dtype=torch.int64, device=dev)
        items = torch.empty(n, dtype=torch.int32, device=dev)
        rcount = torch.zeros(n + 1, dtype=torch.int32, device=dev)
        roff = torch.empty(n + 1, dtype=torch.int64, device=dev)
        st = _lib.stream_ptr()
        _lib.check(lib.pf_cech_count(_lib.ptr(self.centers), _lib.ptr(self.radii), n, ctypes.byref(g),
                                     _lib.ptr(gcount), _lib.ptr(gfill), _lib.ptr(key), _lib.ptr(gstart),
                                     _lib.ptr(items), _lib.ptr(rcount), _lib.ptr(roff), st), "pf_cech_count")
        total = int(roff[-1].item())
        if total >= 2**31:
            raise ValueError(f"Čech complex has {total} edges; int32 offsets overflow")
        adjacent = torch.zeros(max(total, 1), dtype=torch.int32, device=dev)
        offsets = torch.empty(n + 1, dtype=torch.int32, device=dev)
        wsb = lib.pf_cech_write_workspace_bytes(total, n)
        if wsb < 0:
            raise ValueError("Čech complex too large to sort its rows")
        ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
        _lib.check(lib.pf_cech_write(_lib.ptr(self.centers), _lib.ptr(self.radii), n, ctypes.byref(g),
                                     _lib.ptr(gstart), _lib.ptr(items), _lib.ptr(roff), total, _lib.ptr(adjacent),
                                     _lib.ptr(offsets), _lib.ptr(ws), wsb, st), "pf_cech_write")
        return adjacent[:total], offsets
