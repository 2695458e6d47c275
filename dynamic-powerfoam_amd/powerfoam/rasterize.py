"""Tile rasterizer over power cells — the drop-in for ``powerfoam/rasterize.py``.

``Rasterizer(args, device, attr_dtype)`` keeps the reference's constructor,
``forward`` / ``benchmark`` / ``visualize`` signatures and return tuples
(rasterize.py:38, :1589-1606, :1986-2000, :2137-2150) and ``RasterGradFn``'s
autograd contract (rasterize.py:23-34, :1865-1984).  Every kernel is a
hand-written gfx950 kernel (``csrc/raster.hip``) called through the C ABI on
torch's current stream.  As in the reference, the binning learns the number
of (tile, cell) pairs with one host sync (rasterize.py:1717-1718).  Only the
pinhole path exists (``is_pinhole=False`` raises NotImplementedError).
"""
import ctypes
from dataclasses import dataclass, field

import torch

from . import _lib
from .camera import cameras_to_device

TILE_WIDTH = 8
TILE_SIZE = TILE_WIDTH * TILE_WIDTH


@dataclass
class VisOptions:
    """rasterize.py:15-20 (a Warp struct there, a dataclass here)."""
    transmittance_threshold: float = 1e-3
    max_intersections: int = 1024
    depth_quantile: float = 0.5
    bkgd_color: tuple = field(default_factory=lambda: (0.0, 0.0, 0.0))


class RasterGradFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rasterizer, *args):
        return rasterizer._forward(ctx, *args)

    @staticmethod
    def backward(ctx, *args):
        return ctx.rasterizer._backward(ctx, *args)


def _c(t, dtype=None):
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    return t.contiguous()


def _scene_struct(spheres, nsigmas, sites, rgb, height, adjacency, offsets, faces, attr_code, face_code):
    s = _lib.PfScene()
    s.spheres = _lib.ptr(spheres)
    s.nsigmas = _lib.ptr(nsigmas)
    s.texel_sites = _lib.ptr(sites)
    s.texel_rgb = _lib.ptr(rgb)
    s.texel_height = _lib.ptr(height)
    s.adjacency = _lib.ptr(adjacency)
    s.adjacency_offsets = _lib.ptr(offsets)
    s.faces = _lib.ptr(faces)
    s.num_cells = spheres.shape[0]
    s.num_edges = adjacency.numel()
    s.num_texel_sites = sites.shape[1]
    s.attr_dtype = attr_code
    s.face_dtype = face_code
    return s


def bin_tiles(cam_dev, spheres, width, height, want_mask=True, capacity=None, status=None):
    """Binning on the GPU: returns (offsets int64 (tiles+1), prims int32,
    visible mask bool (N,)).  rasterize.py:1681-1772.  With capacity (and
    an int32 status tensor) the list buffer has that fixed size and nothing
    is read back to the host; overflow sets PF_STATUS_LIST_CAP in status
    and prims keeps the full capacity (valid up to offsets[-1])."""
    lib = _lib.load()
    dev = spheres.device
    tiles = ((height + 7) // 8) * ((width + 7) // 8)
    n = spheres.shape[0]
    st = _lib.stream_ptr()
    counts = torch.zeros(tiles + 1, dtype=torch.int32, device=dev)
    _lib.check(lib.pf_raster_bin_count(_lib.ptr(spheres), n, _lib.ptr(cam_dev), _lib.ptr(counts), st),
               "pf_raster_bin_count")
    offsets = torch.empty(tiles + 1, dtype=torch.int64, device=dev)
    if capacity is not None:
        if status is None:
            raise ValueError("a bounded list needs a status tensor")
        _lib.check(lib.pf_raster_scan_bounded(_lib.ptr(counts), tiles + 1, _lib.ptr(offsets), int(capacity),
                                              _lib.ptr(status), st), "pf_raster_scan_bounded")
        n_isect = int(capacity)
    else:
        _lib.check(lib.pf_raster_scan(_lib.ptr(counts), tiles + 1, _lib.ptr(offsets), st), "pf_raster_scan")
        n_isect = int(offsets[-1].item())
    keys = torch.empty(max(n_isect, 1), dtype=torch.int64, device=dev)
    scratch = torch.empty(max(n_isect, 1), dtype=torch.int64, device=dev)
    prims = torch.empty(max(n_isect, 1), dtype=torch.int32, device=dev)
    wc = torch.zeros(tiles, dtype=torch.int32, device=dev)
    mask = torch.zeros(n, dtype=torch.uint8, device=dev) if want_mask else None
    _lib.check(lib.pf_raster_bin_write(_lib.ptr(spheres), n, _lib.ptr(cam_dev), tiles, _lib.ptr(offsets),
                                       _lib.ptr(wc), _lib.ptr(keys), _lib.ptr(scratch), _lib.ptr(prims),
                                       _lib.ptr(mask), n_isect, st), "pf_raster_bin_write")
    return offsets, (prims if capacity is not None else prims[:n_isect]), (mask.bool() if want_mask else None)


def raster_faces(spheres, adjacency, adjacency_offsets, dtype=torch.float16):
    lib = _lib.load()
    faces = torch.empty(adjacency.numel(), 4, dtype=dtype, device=spheres.device)
    _lib.check(lib.pf_raster_faces(_lib.ptr(spheres), spheres.shape[0], _lib.ptr(adjacency),
                                   _lib.ptr(adjacency_offsets), _lib.ptr(faces), _lib.dtype_code(dtype),
                                   _lib.stream_ptr()), "pf_raster_faces")
    return faces


class Rasterizer:
    def __init__(self, args, device, attr_dtype="float"):
        self.device = device
        self.args = args
        if attr_dtype == "float":
            self.tscalar = torch.float32
        elif attr_dtype == "half":
            self.tscalar = torch.float16
        else:
            raise ValueError(f"Unsupported attribute dtype: {attr_dtype}")
        self.attr_dtype = attr_dtype
        self.loss_accumulation = getattr(args, "render_objective", "volume") == "surface"
        self.num_texel_sites = getattr(args, "num_texel_sites", 8)
        if not getattr(args, "is_pinhole", True):
            raise NotImplementedError("only the pinhole camera path is implemented on this backend")
        # The coop-load flags (rasterize.py:101-102, :610-611) select how the
        # reference stages cell data; the results are identical either way
        # and the gfx950 kernels always stage through LDS.
        self.disable_coop_prim_load = getattr(args, "disable_coop_prim_load", False)
        self.disable_coop_adj_load = getattr(args, "disable_coop_adj_load", False)
        # Optional fixed size of the (tile, cell) list: binning then reads
        # nothing back to the host; overflow shows in list_status
        # (PF_STATUS_LIST_CAP) instead of resizing the buffer.
        self.list_capacity = getattr(args, "raster_list_capacity", None)
        self.list_status = None
        if self.list_capacity is not None:
            self.list_status = torch.zeros(1, dtype=torch.int32, device=device)
        _lib.load()

    # ---------------------------------------------------------- forward --
    def _forward(self, grad_ctx, camera, depth_quantiles, points, radii, density, normals, texel_sites, texel_rgb,
                 texel_height, adjacency, adjacency_offsets, ray_gt, return_point_err, transmittance_threshold=None):
        n_inputs = 14 if transmittance_threshold is None else 15  # rasterizer + positional args
        if transmittance_threshold is None:
            transmittance_threshold = getattr(self, "transmittance_threshold", 1e-3)
        if self.attr_dtype != "float":
            raise ValueError("Forward pass only supports float precision")
        if ray_gt is None and self.loss_accumulation:
            raise ValueError("Ground truth rays must be provided for surface loss accumulation")
        if ray_gt is None and return_point_err:
            raise ValueError("Ground truth rays must be provided to return point error")
        lib = _lib.load()
        dev = points.device
        H, W = camera.height, camera.width
        N = points.shape[0]
        spheres = _c(torch.cat([points.detach(), radii.detach()[:, None]], dim=-1), torch.float32)
        nsigmas = _c(torch.cat([normals.detach(), density.detach()[:, None]], dim=-1), torch.float32)
        sites = _c(texel_sites.detach(), torch.float32)
        rgb = _c(texel_rgb.detach(), torch.float32)
        height = _c(texel_height.detach(), torch.float32)
        adj = _c(adjacency, torch.int32)
        off = _c(adjacency_offsets, torch.int32)
        cam_dev = cameras_to_device([camera], None, dev)
        offsets, prims, mask = bin_tiles(cam_dev, spheres, W, H, capacity=self.list_capacity,
                                         status=self.list_status)
        faces = raster_faces(spheres, adj, off)
        scene = _scene_struct(spheres, nsigmas, sites, rgb, height, adj, off, faces, _lib.PF_F32, _lib.PF_F16)
        color = torch.zeros(H, W, 3, dtype=torch.float32, device=dev)
        log_t = torch.zeros(H, W, dtype=torch.float32, device=dev)
        ndist = torch.zeros(H, W, dtype=torch.float32, device=dev)
        normal = torch.zeros(H, W, 3, dtype=torch.float32, device=dev)
        contrib = torch.zeros(N, dtype=torch.float32, device=dev)
        dq = None
        qd = None
        if depth_quantiles is not None:
            dq = _c(depth_quantiles.detach(), torch.float32)
            qd = -1 * torch.ones(*depth_quantiles.shape, dtype=torch.float32, device=dev)
        perr = torch.zeros(N, dtype=torch.float32, device=dev) if return_point_err else None
        err = torch.zeros(H, W, dtype=torch.float32, device=dev) if self.loss_accumulation else None
        gt = _c(ray_gt.detach(), torch.float32) if ray_gt is not None else None
        # the kernels index these per cell and per pixel
        K = sites.shape[1] if sites.dim() == 3 else -1
        if tuple(sites.shape) != (N, K, 3) or tuple(rgb.shape) != (N, K, 3) or tuple(height.shape) != (N, K) or \
                off.numel() != N + 1:
            raise ValueError("texel_sites / texel_rgb must be (N, K, 3), texel_height (N, K), offsets N + 1")
        if dq is not None and (dq.dim() < 1 or dq.shape[-1] < 1 or dq.numel() != H * W * dq.shape[-1]):
            raise ValueError("depth_quantiles must be (H, W, Q)")
        if gt is not None and gt.numel() != H * W * 3:
            raise ValueError("ray_gt must hold (H, W, 3) values")
        tiles = offsets.numel() - 1
        early = torch.zeros(tiles, dtype=torch.int32, device=dev)
        out = _lib.PfRasterOut()
        out.color, out.log_t, out.normal_distance, out.normal = (_lib.ptr(color), _lib.ptr(log_t), _lib.ptr(ndist),
                                                                 _lib.ptr(normal))
        out.quantile_depths, out.err, out.contrib, out.point_err = (_lib.ptr(qd), _lib.ptr(err), _lib.ptr(contrib),
                                                                    _lib.ptr(perr))
        Q = dq.shape[-1] if dq is not None else 0
        rc = lib.pf_raster_forward(ctypes.byref(scene), _lib.ptr(cam_dev), W, H, _lib.ptr(offsets), _lib.ptr(prims),
                                   _lib.ptr(early), float(transmittance_threshold), _lib.ptr(dq), Q, _lib.ptr(gt),
                                   int(bool(return_point_err)), int(self.loss_accumulation), ctypes.byref(out),
                                   _lib.stream_ptr())
        _lib.check(rc, "pf_raster_forward")
        # the last forward's tile lists and stop positions (for byte counts)
        self.last_lists = (offsets, prims, early, off)
        grad_ctx.rasterizer = self
        grad_ctx.camera = camera
        grad_ctx.state = dict(cam_dev=cam_dev, spheres=spheres, nsigmas=nsigmas, sites=sites, rgb=rgb, height=height,
                              adj=adj, off=off, faces=faces, offsets=offsets, prims=prims, early=early, dq=dq,
                              log_t=log_t, n_inputs=n_inputs)
        grad_ctx.mark_non_differentiable(mask)
        opacity = 1.0 - torch.exp(log_t)
        return color, opacity, ndist, normal, qd, err, contrib, perr, mask

    def forward(self, *args):
        return RasterGradFn.apply(self, *args)

    # --------------------------------------------------------- backward --
    def _backward(self, grad_ctx, grad_color_in, grad_opacity_in, grad_normal_distance_in, grad_normal_in,
                  grad_quantile_depths_in, grad_err_in, grad_contrib_in, grad_point_err_in,
                  grad_prim_visible_mask_in):
        del grad_point_err_in, grad_prim_visible_mask_in
        if self.attr_dtype != "float":
            raise ValueError("Backward pass only supports float precision")
        if self.loss_accumulation:
            raise NotImplementedError()  # rasterize.py:1232-1234
        lib = _lib.load()
        st = grad_ctx.state
        H, W = grad_ctx.camera.height, grad_ctx.camera.width
        dev = st["spheres"].device
        N, K = st["spheres"].shape[0], st["sites"].shape[1]
        if K > 8:
            raise ValueError("the backward supports at most 8 texel sites (rasterize.py:65 max_sites)")

        def g(t, shape):
            if t is None:
                return torch.zeros(shape, dtype=torch.float32, device=dev)
            return _c(t.detach(), torch.float32)

        g_color = g(grad_color_in, (H, W, 3))
        g_log_t = None
        if grad_opacity_in is not None:
            g_log_t = _c(-grad_opacity_in.detach().float() * torch.exp(st["log_t"]))
        # every tensor whose pointer goes to the kernel stays referenced here
        # until the launch is enqueued (incoming gradients are often expanded
        # views, so .contiguous() makes new tensors)
        g_nd = g(grad_normal_distance_in, (H, W)) if grad_normal_distance_in is not None else None
        g_n = g(grad_normal_in, (H, W, 3)) if grad_normal_in is not None else None
        gq = g(grad_quantile_depths_in, tuple(st["dq"].shape)) if st["dq"] is not None else None
        g_ct = g(grad_contrib_in, (N,)) if grad_contrib_in is not None else None
        gin = _lib.PfRasterGradIn()
        gin.color = _lib.ptr(g_color)
        gin.log_t = _lib.ptr(g_log_t)
        gin.normal_distance = _lib.ptr(g_nd)
        gin.normal = _lib.ptr(g_n)
        gin.quantile_depths = _lib.ptr(gq)
        gin.contrib = _lib.ptr(g_ct)
        keep = [g_color, g_log_t, g_nd, g_n, gq, g_ct]
        grads = dict(spheres=torch.zeros(N, 4, device=dev), nsigmas=torch.zeros(N, 4, device=dev),
                     texel_sites=torch.zeros(N, K, 3, device=dev), texel_rgb=torch.zeros(N, K, 3, device=dev),
                     texel_height=torch.zeros(N, K, device=dev))
        gs = _lib.PfGrads()
        for k in ("spheres", "nsigmas", "texel_sites", "texel_rgb", "texel_height"):
            setattr(gs, k, _lib.ptr(grads[k]))
        scene = _scene_struct(st["spheres"], st["nsigmas"], st["sites"], st["rgb"], st["height"], st["adj"],
                              st["off"], st["faces"], _lib.PF_F32, _lib.PF_F16)
        Q = st["dq"].shape[-1] if st["dq"] is not None else 0
        rc = lib.pf_raster_backward(ctypes.byref(scene), _lib.ptr(st["cam_dev"]), W, H, _lib.ptr(st["offsets"]),
                                    _lib.ptr(st["prims"]), _lib.ptr(st["early"]), _lib.ptr(st["dq"]), Q,
                                    _lib.ptr(st["log_t"]), ctypes.byref(gin), ctypes.byref(gs), _lib.stream_ptr())
        _lib.check(rc, "pf_raster_backward")
        del keep
        for t in grads.values():
            t[~t.isfinite()] = 0  # rasterize.py:1958-1962
        n_inputs = st["n_inputs"]
        del grad_ctx.rasterizer, grad_ctx.camera, grad_ctx.state
        out = (None, None, None, grads["spheres"][:, :3], grads["spheres"][:, 3], grads["nsigmas"][:, 3],
               grads["nsigmas"][:, :3], grads["texel_sites"], grads["texel_rgb"], grads["texel_height"], None, None,
               None, None)
        return out + (None,) * (n_inputs - len(out))

    # ------------------------------------------------------- benchmark --
    def benchmark(self, camera, points, radii, density, normals, texel_sites, texel_rgb, texel_height, adjacency,
                  adjacency_offsets, adjacency_diff, transmittance_threshold=1e-3):
        """rasterize.py:1986-2135; adjacency_diff are the walk faces
        (dp, dpm) in fp16 (benchmark.py:295-302)."""
        lib = _lib.load()
        dev = points.device
        H, W = camera.height, camera.width
        spheres = _c(torch.cat([points.detach(), radii.detach()[:, None]], dim=-1), torch.float32)
        nsigmas = _c(torch.cat([normals.detach(), density.detach()[:, None]], dim=-1), self.tscalar)
        sites = _c(texel_sites.detach(), self.tscalar)
        rgb = _c(texel_rgb.detach(), self.tscalar)
        height = _c(texel_height.detach(), self.tscalar)
        adj, off = _c(adjacency, torch.int32), _c(adjacency_offsets, torch.int32)
        faces = _c(adjacency_diff, torch.float16)
        cam_dev = cameras_to_device([camera], None, dev)
        offsets, prims, _ = bin_tiles(cam_dev, spheres, W, H, want_mask=False)
        color = torch.zeros(H, W, 3, dtype=torch.float32, device=dev)
        scene = _scene_struct(spheres, nsigmas, sites, rgb, height, adj, off, faces, _lib.dtype_code(self.tscalar),
                              _lib.PF_F16)
        rc = lib.pf_raster_benchmark(ctypes.byref(scene), _lib.ptr(cam_dev), W, H, _lib.ptr(offsets),
                                     _lib.ptr(prims), float(transmittance_threshold), _lib.ptr(color),
                                     _lib.stream_ptr())
        _lib.check(rc, "pf_raster_benchmark")
        return color

    # ------------------------------------------------------- visualize --
    def visualize(self, camera, points, radii, density, normals, texel_sites, texel_rgb, texel_height, adjacency,
                  adjacency_offsets, vis_options=None):
        """rasterize.py:2137-2289 -> (color, depth, normal, alpha, intersections)."""
        if vis_options is None:
            vis_options = VisOptions(transmittance_threshold=1e-3, max_intersections=1024)
        lib = _lib.load()
        dev = points.device
        H, W = camera.height, camera.width
        spheres = _c(torch.cat([points.detach(), radii.detach()[:, None]], dim=-1), torch.float32)
        nsigmas = _c(torch.cat([normals.detach(), density.detach()[:, None]], dim=-1), self.tscalar)
        sites = _c(texel_sites.detach(), self.tscalar)
        rgb = _c(texel_rgb.detach(), self.tscalar)
        height = _c(texel_height.detach(), self.tscalar)
        adj, off = _c(adjacency, torch.int32), _c(adjacency_offsets, torch.int32)
        cam_dev = cameras_to_device([camera], None, dev)
        offsets, prims, _ = bin_tiles(cam_dev, spheres, W, H, want_mask=False)
        faces = raster_faces(spheres, adj, off)
        color = torch.zeros(H, W, 3, dtype=torch.float32, device=dev)
        depth = torch.zeros(H, W, dtype=torch.float32, device=dev)
        normal = torch.zeros(H, W, 3, dtype=torch.float32, device=dev)
        alpha = torch.zeros(H, W, dtype=torch.float32, device=dev)
        isect = torch.zeros(H, W, dtype=torch.int32, device=dev)
        opts = _lib.PfVisOptions()
        opts.transmittance_threshold = float(vis_options.transmittance_threshold)
        opts.max_intersections = int(vis_options.max_intersections)
        opts.depth_quantile = float(vis_options.depth_quantile)
        for k in range(3):
            opts.bkgd_color[k] = float(vis_options.bkgd_color[k])
        vo = _lib.PfVisOut()
        vo.color, vo.depth, vo.normal, vo.alpha, vo.intersections = (_lib.ptr(color), _lib.ptr(depth),
                                                                     _lib.ptr(normal), _lib.ptr(alpha),
                                                                     _lib.ptr(isect))
        scene = _scene_struct(spheres, nsigmas, sites, rgb, height, adj, off, faces, _lib.dtype_code(self.tscalar),
                              _lib.PF_F16)
        rc = lib.pf_raster_visualize(ctypes.byref(scene), _lib.ptr(cam_dev), W, H, _lib.ptr(offsets),
                                     _lib.ptr(prims), ctypes.byref(opts), ctypes.byref(vo), _lib.stream_ptr())
        _lib.check(rc, "pf_raster_visualize")
        return color, depth, normal, alpha, isect
