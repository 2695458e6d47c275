"""MI355X (gfx950) implementation of Power Foam's power-diagram ray tracing
path, with the module layout and call signatures of the reference package
``powerfoam`` (IsaacSuo/dynamic-powerfoam): ``camera``, ``raytrace``,
``rasterize``, ``bvh``, ``warp_interop``.  All rendering runs in hand-written
HIP kernels (``libpf_hip.so``) behind a C ABI; see ``include/pf_hip.h``.
"""
from . import _lib  # noqa: F401

__all__ = ["_lib"]
