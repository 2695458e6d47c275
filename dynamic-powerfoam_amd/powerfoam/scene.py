"""The caller side of the render path: ``PowerfoamScene`` with the parameter
layout, activations, checkpoint format and ``forward`` of
``powerfoam/scene.py`` (the reference's training recipe, densification,
dataset initialisation and point-cloud export are not part of this package).

Activation glue (all plain torch elementwise work, as in the reference):
  radii   = softplus(raw, beta=100)                    scene.py:287-288
  density = softplus(raw, beta=100)                    scene.py:239-240
  normal / tangent / bitangent = rows 0 / 1 / 2 of the rotation matrix of
            the normalised quaternion (w, x, y, z)     scene.py:242-285
  texel sites (world) = p + r (u t + v b)               scene.py:324-329
  texel heights = h r                                  scene.py:337
  spherical-Voronoi lobes, lobe-major                  scene.py:290-312
``forward`` colours the texels with ``SphericalVoronoi`` (sites detached,
scene.py:331-335) and returns ``Rasterizer.forward``'s 9-tuple
(scene.py:339-353).  ``render_walk`` renders the same scene with the
power-cell walk on a power-diagram adjacency.
"""
import torch
import torch.nn as nn
import torch.nn.functional as F

from .bvh import AABBTree
from .color_fn import SphericalVoronoi
from .rasterize import Rasterizer
from .raytrace import RayTracer, walk_faces

CHECKPOINT_KEYS = ("points", "density", "radii", "quaternions", "texel_sites", "texel_sv_axis", "texel_sv_rgb",
                   "texel_height", "adjacency", "adjacency_offsets")


class PowerfoamScene(nn.Module):
    """Parameters (raw, before activation):
    points (N,3), radii (N,), density (N,), quaternions (N,4) [w,x,y,z],
    texel_sites (N,K,2) tangent-plane offsets in units of r,
    texel_sv_axis (N,K,3*D) lobe axes scaled by temperature,
    texel_sv_rgb (N,K,3*D) lobe colours, texel_height (N,K) in units of r;
    D = args.sv_dof, K = args.num_texel_sites."""

    def __init__(self, args, attr_dtype="float"):
        super().__init__()
        self.args = args
        if attr_dtype == "float":
            self.tscalar = torch.float32
        elif attr_dtype == "half":
            self.tscalar = torch.float16
        else:
            raise ValueError(f"Unsupported attribute dtype: {attr_dtype}")
        self.attr_dtype = attr_dtype
        self.adjacency = None
        self.adjacency_offsets = None

    # ---------------------------------------------------------- set-up --
    def _set_params(self, tensors, device):
        for k in CHECKPOINT_KEYS[:8]:
            t = tensors[k].to(device)
            t = t.float() if k == "points" else t.to(self.tscalar)
            setattr(self, k, nn.Parameter(t.contiguous()))
        self._renderers(device)

    def _renderers(self, device):
        self.rasterizer = Rasterizer(self.args, device, self.attr_dtype)
        self.raytracer = RayTracer(self.args, device, self.attr_dtype)
        self.sv = SphericalVoronoi(self.args, device, self.attr_dtype)
        self.aabb_tree = AABBTree(device)

    def initialize_random(self, num_points, device, seed=0, extent=1.0):
        """Random scene in [-extent, extent]^3 with the reference's initial
        parameter statistics (scene.py:122-176); radii from the mean
        distance to 8 nearest neighbours is replaced by the mean spacing."""
        g = torch.Generator().manual_seed(seed)
        n, K, D = num_points, self.args.num_texel_sites, self.args.sv_dof
        spacing = 2.0 * extent / max(num_points, 1) ** (1.0 / 3.0)
        t = dict(
            points=(torch.rand(n, 3, generator=g) * 2 - 1) * extent,
            radii=torch.full((n,), 0.5 * spacing),
            density=torch.full((n,), 0.1),
            quaternions=torch.nn.functional.normalize(torch.randn(n, 4, generator=g), dim=-1),
            texel_sites=0.1 * torch.randn(n, K, 2, generator=g),
            texel_sv_axis=2.0 * torch.randn(n, K, 3 * D, generator=g),
            texel_sv_rgb=0.3 * torch.randn(n, K, 3 * D, generator=g),
            texel_height=torch.zeros(n, K),
        )
        self._set_params(t, device)
        self.rebuild_adjacency()
        return self

    @property
    def device(self):
        return self.points.device

    def rebuild_adjacency(self):
        """Cech complex of the activated spheres (scene.py:235-237)."""
        self.aabb_tree.update(self.points.detach(), self.get_radii().detach().float())
        self.adjacency, self.adjacency_offsets = self.aabb_tree.build_cech_complex()

    # ----------------------------------------------------- activations --
    def get_density(self):
        return F.softplus(self.density, beta=100)

    def get_radii(self):
        return F.softplus(self.radii, beta=100)

    def _unit_quaternions(self):
        q = self.quaternions / self.quaternions.norm(dim=-1, keepdim=True)
        return q[:, 0], q[:, 1], q[:, 2], q[:, 3]

    def get_normals(self):
        w, x, y, z = self._unit_quaternions()
        n = torch.stack([1 - 2 * (y ** 2 + z ** 2), 2 * (x * y - z * w), 2 * (x * z + y * w)], dim=-1)
        return n / n.norm(dim=-1, keepdim=True)

    def get_tangents(self):
        w, x, y, z = self._unit_quaternions()
        t = torch.stack([2 * (x * y + z * w), 1 - 2 * (x ** 2 + z ** 2), 2 * (y * z - x * w)], dim=-1)
        b = torch.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x ** 2 + y ** 2)], dim=-1)
        return t / t.norm(dim=-1, keepdim=True), b / b.norm(dim=-1, keepdim=True)

    def get_att_sv(self):
        """Lobe arrays in the lobe-major layout the colour kernels read:
        axes (D, N*K, 3) unit, values (D, N*K, 3), temperatures (D, N*K)
        = the axis parameters' lengths."""
        D, K = self.args.sv_dof, self.args.num_texel_sites
        axis = self.texel_sv_axis.view(-1, K, D, 3)
        temp = axis.norm(dim=-1)
        axis = axis / axis.norm(dim=-1, keepdim=True)
        axis = axis.permute(2, 0, 1, 3).reshape(D, -1, 3).contiguous()
        rgb = self.texel_sv_rgb.view(-1, K, D, 3).permute(2, 0, 1, 3).reshape(D, -1, 3).contiguous()
        temp = temp.permute(2, 0, 1).reshape(D, -1).contiguous()
        return axis, rgb, temp

    def world_texels(self):
        """(texel_sites (N,K,3) world positions, texel_height (N,K))."""
        radii = self.get_radii()
        t, b = self.get_tangents()
        off = self.texel_sites * radii[:, None, None]
        off = off[..., 0:1] * t[:, None, :] + off[..., 1:2] * b[:, None, :]
        return self.points[:, None, :] + off, self.texel_height * radii[:, None]

    def texel_colors(self, camera, texel_sites):
        axis, rgb, temp = self.get_att_sv()
        c = self.sv.forward(texel_sites.reshape(-1, 3).detach().float(), camera, axis, rgb, temp)
        return c.view(self.points.shape[0], self.args.num_texel_sites, 3)

    # --------------------------------------------------------- renders --
    def forward(self, camera, depth_quantiles=None, ray_gt=None, return_point_err=False):
        """Differentiable raster render; Rasterizer.forward's 9-tuple."""
        sites, heights = self.world_texels()
        rgb = self.texel_colors(camera, sites)
        return self.rasterizer.forward(camera, depth_quantiles, self.points, self.get_radii(), self.get_density(),
                                       self.get_normals(), sites, rgb, heights, self.adjacency,
                                       self.adjacency_offsets, ray_gt, return_point_err)

    def render_walk(self, camera, power_adjacency, transmittance_threshold=1e-3):
        """Differentiable walk render on a power-diagram CSR
        (adjacency.build_power_adjacency).  Returns (color, alpha, depth) of
        shape (1,H,W,...)."""
        adj, off = power_adjacency
        adj, off = adj.to(self.device).int().contiguous(), off.to(self.device).int().contiguous()
        sites, heights = self.world_texels()
        rgb = self.texel_colors(camera, sites)
        radii = self.get_radii()
        faces = walk_faces(self.points.detach(), radii.detach().float(), adj, off)
        return self.raytracer.forward(camera, None, self.points, radii, self.get_density(), self.get_normals(),
                                      sites, rgb, heights, adj, off, faces, transmittance_threshold)

    # ------------------------------------------------------ checkpoints --
    def save_pt(self, pt_path):
        """The reference's checkpoint dict (scene.py:772-799)."""
        data = {k: getattr(self, k).detach().float().cpu() for k in CHECKPOINT_KEYS[:8]}
        data["adjacency"] = self.adjacency.cpu()
        data["adjacency_offsets"] = self.adjacency_offsets.cpu()
        torch.save(data, pt_path)

    def load_pt(self, pt_path, device=None):
        """Load a checkpoint written by the reference's save_pt
        (scene.py:801-821).  Tensors only: the file is read with
        weights_only=True, so nothing in it is executed.  Parameters are
        created at the checkpoint's size if the scene has none yet."""
        data = torch.load(pt_path, map_location="cpu", weights_only=True)
        missing = [k for k in CHECKPOINT_KEYS if k not in data]
        if missing:
            raise ValueError(f"checkpoint lacks {missing}")
        device = torch.device(device) if device is not None else (
            self.points.device if hasattr(self, "points") else torch.device("cuda"))
        self._set_params(data, device)
        self.adjacency = data["adjacency"].to(device).int().contiguous()
        self.adjacency_offsets = data["adjacency_offsets"].to(device).int().contiguous()
        return self
