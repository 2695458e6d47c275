"""Cell ordering: Morton (Z-order) codes and the sort by them.

``morton_sort`` keeps the signature and result of geometry.py:50-85 (int32
permutation of the points, stable in the code).  The reference applies it
in training (``sort_points``, scene.py:197-233), so trained scenes have
their cells in this order; ``permute_cells`` applies a permutation to a
scene's per-cell arrays (before its adjacency is built).
"""
import ctypes

import torch

from . import _lib


def morton_codes(points):
    """(N,) int64 codes of points (N, 3) over their bounding box, as
    morton_code_kernel (geometry.py:18-47) computes them."""
    if points.device.type != "cuda":
        raise ValueError("Only CUDA tensors are supported as inputs")
    lib = _lib.load()
    pts = points.detach().to(torch.float32).contiguous()
    n = pts.shape[0]
    codes = torch.empty(n, dtype=torch.int64, device=pts.device)
    if n == 0:
        return codes
    lo = (ctypes.c_float * 3)(*pts.min(0).values.tolist())
    hi = (ctypes.c_float * 3)(*pts.max(0).values.tolist())
    _lib.check(lib.pf_morton_codes(_lib.ptr(pts), n, lo, hi, _lib.ptr(codes), _lib.stream_ptr()),
               "pf_morton_codes")
    return codes


def morton_sort(points):
    """int32 permutation sorting points by Morton code (geometry.py:50-85;
    a stable sort, as the reference's radix sort is)."""
    if points.shape[0] == 0:
        return torch.zeros(0, dtype=torch.int32, device=points.device)
    codes = morton_codes(points)
    return torch.sort(codes, stable=True).indices.to(torch.int32)


def permute_cells(perm, *arrays):
    """Each per-cell array (first dimension N) reordered by perm; None stays None."""
    idx = perm.long()
    return tuple(None if a is None else a[idx.to(a.device)] for a in arrays)
