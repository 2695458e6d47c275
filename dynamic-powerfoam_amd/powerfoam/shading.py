"""View-dependent colour pre-pass (spherical harmonics) on the GPU.

Produces the ``texel_rgb`` the walk and the rasterizer consume, per camera,
the way ``SphericalVoronoi.forward`` (powerfoam/color_fn.py:188-246) does
for the reference's spherical-Voronoi colour: one value per texel site, from
the direction site - eye.  Kernels: ``csrc/sh.hip``.
"""
import torch

from . import _lib
from .raytrace import _expect


def _check_inputs(coeffs, degree, texel_sites, cams_dev):
    n = texel_sites.shape[0]
    K = texel_sites.shape[1] if texel_sites.dim() == 3 else -1
    _expect(coeffs, (n, (degree + 1) ** 2, 3), "coeffs")
    _expect(texel_sites, (n, K, 3), "texel_sites", device=coeffs.device)
    _expect(cams_dev, (cams_dev.shape[0], 12), "cams_dev", dtype=torch.int32, device=coeffs.device)
    return n, K, cams_dev.shape[0]


def sh_forward(coeffs, degree, texel_sites, cams_dev, out=None, dtype=torch.float32):
    """coeffs (N, (D+1)^2, 3) f32, texel_sites (N, K, 3) f32, cams_dev (B, 12)
    pf_camera words -> texel_rgb (B, N, K, 3)."""
    lib = _lib.load()
    n, K, B = _check_inputs(coeffs, degree, texel_sites, cams_dev)
    if out is None:
        out = torch.empty(B, n, K, 3, dtype=dtype, device=coeffs.device)
    _expect(out, (B, n, K, 3), "out", dtype=out.dtype, device=coeffs.device)
    rc = lib.pf_sh_forward(_lib.ptr(coeffs), degree, _lib.ptr(texel_sites), n, K, _lib.ptr(cams_dev), B,
                           _lib.ptr(out), _lib.dtype_code(out.dtype), n * K * 3, _lib.stream_ptr())
    _lib.check(rc, "pf_sh_forward")
    return out


def sh_backward(coeffs, degree, texel_sites, cams_dev, grad_rgb, grad_coeffs=None):
    """grad_rgb (B, N, K, 3) f32 -> dL/dcoeffs (N, (D+1)^2, 3) (accumulated)."""
    lib = _lib.load()
    n, K, B = _check_inputs(coeffs, degree, texel_sites, cams_dev)
    if grad_coeffs is None:
        grad_coeffs = torch.zeros_like(coeffs)
    _expect(grad_coeffs, tuple(coeffs.shape), "grad_coeffs", device=coeffs.device)
    grad_rgb = grad_rgb.contiguous()  # held until the launch is enqueued
    _expect(grad_rgb, (B, n, K, 3), "grad_rgb", device=coeffs.device)
    rc = lib.pf_sh_backward(_lib.ptr(coeffs), degree, _lib.ptr(texel_sites), n, K, _lib.ptr(cams_dev), B,
                            _lib.ptr(grad_rgb), n * K * 3, _lib.ptr(grad_coeffs), _lib.stream_ptr())
    _lib.check(rc, "pf_sh_backward")
    return grad_coeffs


class SHColorFn(torch.autograd.Function):
    """texel_rgb = SH(coeffs; texel site directions), differentiable in the
    coefficients (the sites are detached, as powerfoam/scene.py:333-335
    passes texel_sites.detach() to the spherical-Voronoi pre-pass)."""

    @staticmethod
    def forward(ctx, coeffs, degree, texel_sites, cams_dev):
        ctx.save_for_backward(coeffs, texel_sites, cams_dev)
        ctx.degree = degree
        return sh_forward(coeffs, degree, texel_sites, cams_dev)

    @staticmethod
    def backward(ctx, grad_rgb):
        coeffs, sites, cams_dev = ctx.saved_tensors
        return sh_backward(coeffs, ctx.degree, sites, cams_dev, grad_rgb.contiguous()), None, None, None
