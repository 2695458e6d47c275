"""Spherical-Voronoi colour pre-pass on the GPU, with the interface of
``powerfoam/color_fn.py`` (``SphericalVoronoi``, ``SphericalVoronoiGradFn``).

Turns a trained checkpoint's per-texel lobes (axis, rgb, temperature) into
the ``texel_rgb`` for one camera, the step right before the walk or the
rasterizer.  Kernels: ``csrc/sv.hip`` (``pf_sv_forward`` /
``pf_sv_backward``).
"""
import math
import types

import torch

from . import _lib

# Angular margin beyond the frustum diagonal (powerfoam/color_fn.py:7).
_FOV_BUFFER_RAD = math.radians(10.0)


def _host3(t):
    return torch.as_tensor(t, dtype=torch.float32).detach().reshape(3).cpu().contiguous()


def camera_frame(camera):
    """(eye, forward, fov_cos_cutoff) as color_fn.py:203-212 derives them:
    forward = normalize(up x right), cutoff = cos(atan(sqrt(|right|^2 +
    |up|^2)) + 10 degrees).  Computed on the host from the camera's
    tensors (no device round trip when the camera lives on the CPU)."""
    eye, right, up = _host3(camera.eye), _host3(camera.right), _host3(camera.up)
    fwd = torch.cross(up, right, dim=-1)
    fwd = fwd / torch.norm(fwd, dim=-1, keepdim=True)
    rn, un = float(torch.linalg.norm(right)), float(torch.linalg.norm(up))
    cutoff = math.cos(math.atan(math.sqrt(rn ** 2 + un ** 2)) + _FOV_BUFFER_RAD)
    return eye, fwd.contiguous(), cutoff


class SphericalVoronoiGradFn(torch.autograd.Function):
    """Autograd wrapper (powerfoam/color_fn.py:10-24)."""

    @staticmethod
    def forward(ctx, spherical_voronoi, points, camera, att_sites, att_values, att_temps):
        return spherical_voronoi._forward(ctx, points, camera, att_sites, att_values, att_temps)

    @staticmethod
    def backward(ctx, *args):
        return ctx.spherical_voronoi._backward(ctx, *args)


class SphericalVoronoi:
    """Per-texel view-dependent colour from ``sv_dof`` lobes.

    Lobe arrays are lobe-major, as ``PowerfoamScene.get_att_sv`` lays them
    out (powerfoam/scene.py:290-312): axes (sv_dof, N*K, 3), values
    (sv_dof, N*K, 3), temperatures (sv_dof, N*K), all of ``attr_dtype``.
    """

    @staticmethod
    def compute_fov_cos_cutoff(camera):
        """Cutoff from the intrinsics (powerfoam/color_fn.py:29-37)."""
        K = camera.intrinsics_matrix()
        fx, fy = K[0, 0], K[1, 1]
        cx, cy = K[0, 2], K[1, 2]
        x_max = max(float(cx), float(camera.width) - float(cx)) / float(fx)
        y_max = max(float(cy), float(camera.height) - float(cy)) / float(fy)
        tan_theta = float((x_max ** 2 + y_max ** 2) ** 0.5 * 1.1)
        return 1.0 / ((1.0 + tan_theta ** 2) ** 0.5)

    def __init__(self, args, device, attr_dtype="float"):
        if attr_dtype == "float":
            self.tscalar = torch.float32
        elif attr_dtype == "half":
            self.tscalar = torch.float16
        else:
            raise ValueError(f"Unsupported attribute dtype: {attr_dtype}")
        self.device = torch.device(device)
        self.args = args
        self.sv_dof = int(args.sv_dof)

    def _check(self, points, att_sites, att_values, att_temps):
        n = points.shape[0]
        if points.dtype != torch.float32 or points.dim() != 2 or points.shape[1] != 3:
            raise ValueError("points must be (N*K, 3) float32 texel sites")
        for name, a, shape in (("att_sites", att_sites, (self.sv_dof, n, 3)),
                               ("att_values", att_values, (self.sv_dof, n, 3)),
                               ("att_temps", att_temps, (self.sv_dof, n))):
            if tuple(a.shape) != shape:
                raise ValueError(f"{name} must have shape {shape}, got {tuple(a.shape)}")
            if a.dtype != self.tscalar:
                raise ValueError(f"{name} must be {self.tscalar}")

    def _forward(self, ctx, points, camera, att_sites, att_values, att_temps):
        self._check(points, att_sites, att_values, att_temps)
        lib = _lib.load()
        pts = points.detach().contiguous()
        axes, vals, temps = (a.detach().contiguous() for a in (att_sites, att_values, att_temps))
        eye, fwd, cutoff = camera_frame(camera)
        n = pts.shape[0]
        values_out = torch.empty(n, 3, device=pts.device, dtype=self.tscalar)
        wsum_out = torch.empty(n, device=pts.device, dtype=self.tscalar)
        if n > 0:
            rc = lib.pf_sv_forward(_lib.ptr(pts), n, eye.data_ptr(), fwd.data_ptr(), cutoff, _lib.ptr(axes),
                                   _lib.ptr(vals), _lib.ptr(temps), self.sv_dof, _lib.dtype_code(self.tscalar),
                                   _lib.ptr(values_out), _lib.ptr(wsum_out), _lib.stream_ptr())
            _lib.check(rc, "pf_sv_forward")
        if ctx is not None:
            ctx.spherical_voronoi = self
            ctx.frame = (eye, fwd, cutoff)
            ctx.save_for_backward(pts, axes, vals, temps, values_out, wsum_out)
        return values_out

    def forward(self, points, camera, att_sites, att_values, att_temps):
        """texel_rgb (N*K, 3) for ``camera``; differentiable in the lobes."""
        return SphericalVoronoiGradFn.apply(self, points, camera, att_sites, att_values, att_temps)

    def _backward(self, ctx, grad_values):
        pts, axes, vals, temps, values_in, wsum_in = ctx.saved_tensors
        eye, fwd, cutoff = ctx.frame
        g_axes, g_vals, g_temps = torch.zeros_like(axes), torch.zeros_like(vals), torch.zeros_like(temps)
        n = pts.shape[0]
        if n > 0:
            gin = grad_values.to(self.tscalar).contiguous()
            rc = _lib.load().pf_sv_backward(_lib.ptr(pts), n, eye.data_ptr(), fwd.data_ptr(), cutoff,
                                            _lib.ptr(axes), _lib.ptr(vals), _lib.ptr(temps), self.sv_dof,
                                            _lib.dtype_code(self.tscalar), _lib.ptr(wsum_in), _lib.ptr(values_in),
                                            _lib.ptr(gin), _lib.ptr(g_axes), _lib.ptr(g_vals), _lib.ptr(g_temps),
                                            _lib.stream_ptr())
            _lib.check(rc, "pf_sv_backward")
        return None, None, None, g_axes, g_vals, g_temps


class _SavedCtx:
    """What SphericalVoronoi._forward saves, kept for a backward run outside
    autograd (the serving path colours a request once and differentiates it
    after its last tile)."""

    def save_for_backward(self, *tensors):
        self.saved_tensors = tensors


def sv_prepass(points, camera, att_sites, att_values, att_temps, keep_for_backward=False):
    """texel_rgb (N*K, 3) of the lobes for one camera, outside autograd.
    With keep_for_backward, returns (texel_rgb, ctx) and ctx feeds
    sv_prepass_backward."""
    dtype = "half" if att_sites.dtype == torch.float16 else "float"
    sv = SphericalVoronoi(types.SimpleNamespace(sv_dof=att_sites.shape[0]), points.device, dtype)
    ctx = _SavedCtx() if keep_for_backward else None
    out = sv._forward(ctx, points, camera, att_sites, att_values, att_temps)
    return (out, ctx) if keep_for_backward else out


def sv_prepass_backward(ctx, grad_values):
    """(d axes, d values, d temperatures) of the lobes from dL/dtexel_rgb
    (N*K, 3), for a pre-pass run with keep_for_backward."""
    return ctx.spherical_voronoi._backward(ctx, grad_values)[3:]
