"""Power-cell walk — the drop-in for ``powerfoam/raytrace.py``'s ``RayTracer``.

``RayTracer.benchmark`` keeps the reference's signature and return value
(``powerfoam/raytrace.py:247-309``) and runs the hand-written gfx950 walk
kernel (``csrc/walk.hip``) through the C ABI on torch's current stream.
Extensions, not changes: ``forward`` renders a batch of cameras, returns
colour, alpha and depth (plus optional visited-cell sequences) and is
differentiable through ``WalkGradFn``, whose backward replays each ray's walk
from a compact per-ray checkpoint.
"""
import ctypes

import torch

from . import _lib
from .camera import cameras_to_device

TILE_WIDTH = 8
TILE_SIZE = TILE_WIDTH * TILE_WIDTH
DEFAULT_MAX_STEPS = 1 << 16
# contributing segments per ray the differentiable forward records for its
# backward, and the most device memory one record and its grouped-backward
# workspace may take (96 B per segment slot; 16 cameras of 800x800 keep a cap
# of 64 under it, where a cap of 24 leaves 2% of the rays to replay)
DEFAULT_RECORD_CAP = 128
DEFAULT_RECORD_BUDGET = 64 << 30


def _contig(t, dtype=None):
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    return t.contiguous()


class WalkScene:
    """Device arrays of one scene in the kernels' layout, plus the C struct.

    spheres = [points, radii] (f32) and nsigmas = [normals, density]
    (attribute dtype), as packed at raytrace.py:274-280.
    """

    def __init__(self, points, radii, density, normals, texel_sites, texel_rgb, texel_height, adjacency,
                 adjacency_offsets, faces, attr_dtype=torch.float32, pack_faces=False):
        dev = points.device
        if dev.type != "cuda":
            raise ValueError("Only CUDA tensors are supported as inputs")
        n = points.shape[0]
        self.num_cells = n
        self.spheres = _contig(torch.cat([points.detach(), radii.detach()[:, None]], dim=-1), torch.float32)
        self.nsigmas = _contig(torch.cat([normals.detach(), density.detach()[:, None]], dim=-1), attr_dtype)
        self.texel_sites = _contig(texel_sites.detach(), attr_dtype)
        self.texel_rgb = _contig(texel_rgb.detach(), attr_dtype)
        self.texel_height = _contig(texel_height.detach(), attr_dtype)
        if self.texel_sites.dim() != 3 or self.texel_sites.shape[0] != n:
            raise ValueError("texel_sites must be (N, K, 3)")
        self.K = self.texel_sites.shape[1]
        # texel_rgb (N, K, 3), or (B, N, K, 3): one set of colours per camera
        # (the view-dependent colour pre-pass of each camera)
        self.rgb_cams = 0
        if self.texel_rgb.dim() == 4:
            self.rgb_cams = self.texel_rgb.shape[0]
        if tuple(self.texel_rgb.shape[-3:]) != (n, self.K, 3) or self.texel_rgb.dim() not in (3, 4) or \
                tuple(self.texel_height.shape) != (n, self.K):
            raise ValueError("texel_rgb must be (N, K, 3) or (B, N, K, 3) and texel_height (N, K)")
        if adjacency.dtype not in (torch.int32, torch.int64) or adjacency_offsets.dtype not in (torch.int32,
                                                                                                torch.int64):
            raise ValueError("adjacency / adjacency_offsets must be integer tensors")
        self.adjacency = _contig(adjacency, torch.int32)
        self.offsets = _contig(adjacency_offsets, torch.int32)
        if self.offsets.numel() != n + 1:
            raise ValueError("adjacency_offsets must have N+1 entries")
        self.faces = _contig(faces)
        E = self.adjacency.numel()
        if self.faces.dtype == torch.int32 and self.faces.shape == (E, 3):
            face_code = _lib.PF_F16_PACKED  # already packed records
        elif self.faces.dtype in (torch.float16, torch.float32) and self.faces.shape == (E, 4):
            face_code = _lib.dtype_code(self.faces.dtype)
            if pack_faces and face_code == _lib.PF_F16:
                # same values, one 12-byte record per face with its
                # neighbour (the unpacked fp16 array is read two faces per
                # load, which measured faster in the walk)
                self.faces = pack_walk_faces(self.faces, self.adjacency)
                face_code = _lib.PF_F16_PACKED
        else:
            raise ValueError("adjacency_diff must be (E, 4) float16 or float32, or (E, 3) int32 packed records")
        s = _lib.PfScene()
        s.spheres = _lib.ptr(self.spheres)
        s.nsigmas = _lib.ptr(self.nsigmas)
        s.texel_sites = _lib.ptr(self.texel_sites)
        s.texel_rgb = _lib.ptr(self.texel_rgb)
        s.texel_height = _lib.ptr(self.texel_height)
        s.adjacency = _lib.ptr(self.adjacency)
        s.adjacency_offsets = _lib.ptr(self.offsets)
        s.faces = _lib.ptr(self.faces)
        s.num_cells = n
        s.num_edges = self.adjacency.numel()
        s.num_texel_sites = self.K
        s.attr_dtype = _lib.dtype_code(attr_dtype)
        s.face_dtype = face_code
        s.texel_rgb_cam_stride = n * self.K * 3 if self.rgb_cams else 0
        self.c = s




def _check_rgb_cams(scene, B):
    """Per-camera texel colours must cover exactly the launch's cameras."""
    n = getattr(scene, "rgb_cams", 0)
    if n and n != B:
        raise ValueError(f"texel_rgb holds colours for {n} cameras, the launch has {B}")


def start_cells_device(points, radii, cams_dev):
    """Write each camera's start cell (benchmark.py:320-324) into the device
    pf_camera array on the GPU, without a host sync."""
    lib = _lib.load()
    B = cams_dev.shape[0]
    ws = torch.empty(B, dtype=torch.int64, device=cams_dev.device)
    pts = _contig(points.detach(), torch.float32)
    rad = _contig(radii.detach(), torch.float32)
    rc = lib.pf_start_cells(_lib.ptr(pts), _lib.ptr(rad), pts.shape[0], _lib.ptr(cams_dev), B, _lib.ptr(ws),
                            _lib.stream_ptr())
    _lib.check(rc, "pf_start_cells")
    return cams_dev


def pack_walk_faces(faces_f16, adjacency):
    """(E, 4) f16 walk faces + adjacency -> (E, 3) int32 packed records
    (PF_F16_PACKED): the same fp16 values with the neighbour alongside."""
    lib = _lib.load()
    f = _contig(faces_f16)
    adj = _contig(adjacency, torch.int32)
    if f.dtype != torch.float16 or f.shape != (adj.numel(), 4):
        raise ValueError("faces must be (E, 4) float16")
    out = torch.empty(adj.numel(), 3, dtype=torch.int32, device=f.device)
    _lib.check(lib.pf_walk_pack_faces(_lib.ptr(f), _lib.ptr(adj), adj.numel(), _lib.ptr(out), _lib.stream_ptr()),
               "pf_walk_pack_faces")
    return out


def walk_faces(points, radii, adjacency, adjacency_offsets, dtype=torch.float16):
    """Walk faces (benchmark.py:295-302) computed on the GPU: (E, 4)
    [p_j - p_i, pm_j - pm_i], pm = 0.5(|p|^2 - r^2)."""
    lib = _lib.load()
    pts = _contig(points.detach(), torch.float32)
    rad = _contig(radii.detach(), torch.float32)
    adj = _contig(adjacency, torch.int32)
    off = _contig(adjacency_offsets, torch.int32)
    faces = torch.empty(adj.numel(), 4, dtype=dtype, device=pts.device)
    rc = lib.pf_walk_faces(_lib.ptr(pts), _lib.ptr(rad), pts.shape[0], _lib.ptr(adj), _lib.ptr(off), adj.numel(),
                           _lib.ptr(faces), _lib.dtype_code(dtype), _lib.stream_ptr())
    _lib.check(rc, "pf_walk_faces")
    return faces


def record_slots(num_cams, width, height):
    """Slots of a walk record: one per lane of the 8x8 tiles."""
    return num_cams * ((height + 7) // 8) * ((width + 7) // 8) * 64


def record_bytes(num_cams, width, height, cap, partials=True):
    """Device bytes of a record: 16 per segment (32 with partials) and 4 per
    slot, plus its 32-byte stamp."""
    S = record_slots(num_cams, width, height)
    return (32 if partials else 16) * cap * S + 4 * S + 32


RECORD_CHUNK = 8  # segments of one ray per 128-byte line; caps are multiples


def new_record(num_cams, width, height, cap, device, partials=True):
    """Device buffers of a walk record holding up to cap contributing
    segments per ray (16 bytes each; cap rounded up to a multiple of 8),
    every ray's count, the stamp the forward writes and, with partials, each
    segment's transmittance before it and the ray's colour after it (16
    bytes more, which the grouped backward needs).  Segment k of slot s is
    row ((k // 8) * S + s) * 8 + k % 8 of "segments" viewed as (cap * S, 4)."""
    if cap <= 0:
        raise ValueError("record cap must be positive")
    cap = -(-int(cap) // RECORD_CHUNK) * RECORD_CHUNK
    S = record_slots(num_cams, width, height)
    return dict(segments=torch.empty(cap * S, 4, dtype=torch.int32, device=device),
                counts=torch.empty(S, dtype=torch.int32, device=device), cap=cap,
                partials=torch.empty(cap * S, 4, dtype=torch.float32, device=device) if partials else None,
                header=torch.zeros(8, dtype=torch.int32, device=device))


def record_segment(record, k, slot):
    """Row of segment k of slot `slot` in record["segments"]."""
    S = record["counts"].numel()
    return ((k // RECORD_CHUNK) * S + slot) * RECORD_CHUNK + k % RECORD_CHUNK


def record_counts(record, num_cams, width, height):
    """The record's per-ray segment counts as a (B, H, W) image.  Slot
    tile * 64 + lane holds pixel (tile_i * 8 + lane % 8, tile_j * 8 +
    lane // 8), tiles column-major within a camera."""
    th, tw = (height + 7) // 8, (width + 7) // 8
    c = record["counts"].view(num_cams, tw, th, 8, 8)  # (b, tile_j, tile_i, lane // 8, lane % 8)
    return c.permute(0, 2, 4, 1, 3).reshape(num_cams, th * 8, tw * 8)[:, :height, :width]


def _record_struct(record):
    r = _lib.PfWalkRecord()
    r.segments = _lib.ptr(record["segments"])
    r.counts = _lib.ptr(record["counts"])
    r.cap = record["cap"]
    r.partials = _lib.ptr(record.get("partials"))
    r.header = _lib.ptr(record.get("header"))
    return r


def grouped_workspace(num_cells, num_cams, width, height, cap, device):
    """Device workspace of the grouped backward (bucket offsets, cursors,
    the cell-sorted segment ids)."""
    lib = _lib.load()
    n = lib.pf_walk_grouped_workspace_bytes(int(num_cells), int(num_cams), int(width), int(height), int(cap))
    if n <= 0:
        raise ValueError("sizes out of range for the grouped backward")
    return torch.empty(n, dtype=torch.uint8, device=device)


def _expect(t, shape, name, dtype=torch.float32, device=None):
    """The kernels index these buffers by the launch's sizes: refuse any
    other shape, dtype, device or layout instead of reading past them."""
    if t is None:
        return
    if tuple(t.shape) != tuple(shape) or t.dtype != dtype or not t.is_contiguous() or \
            (device is not None and t.device != device):
        raise ValueError(f"{name} must be a contiguous {dtype} tensor of shape {tuple(shape)} on {device}, "
                         f"got {t.dtype} {tuple(t.shape)} on {t.device}")


def launch_walk_forward(scene, cams_dev, width, height, transmittance_threshold=1e-3, depth_quantile=0.5,
                        max_steps=DEFAULT_MAX_STEPS, want_log_t=True, want_depth=True, want_sequences=False,
                        seq_cap=0, want_bytes=False, record_cap=0, want_ray_status=False):
    """One walk launch over a batch of cameras.  record_cap > 0 also writes
    the record of contributing segments that launch_walk_backward(record=)
    differentiates without replaying the walk (out["record"])."""
    lib = _lib.load()
    B = cams_dev.shape[0]
    dev = cams_dev.device
    _expect(cams_dev, (B, 12), "cams_dev", dtype=torch.int32)
    _check_rgb_cams(scene, B)
    color = torch.empty(B, height, width, 3, dtype=torch.float32, device=dev)
    log_t = torch.empty(B, height, width, dtype=torch.float32, device=dev) if want_log_t else None
    depth = torch.empty(B, height, width, dtype=torch.float32, device=dev) if want_depth else None
    status = torch.zeros(1, dtype=torch.int32, device=dev)
    visits = seq_hash = seq = None
    if want_sequences:
        visits = torch.empty(B, height, width, dtype=torch.int32, device=dev)
        seq_hash = torch.empty(B, height, width, dtype=torch.int64, device=dev)
        if seq_cap > 0:
            seq = torch.full((B, height, width, seq_cap), -1, dtype=torch.int32, device=dev)
    out = _lib.PfWalkOut()
    out.color = _lib.ptr(color)
    out.log_t = _lib.ptr(log_t)
    out.depth = _lib.ptr(depth)
    out.visits = _lib.ptr(visits)
    out.seq_hash = _lib.ptr(seq_hash)
    out.seq = _lib.ptr(seq)
    out.seq_cap = seq_cap if seq is not None else 0
    out.status = _lib.ptr(status)
    alg_bytes = torch.zeros(B, height, width, dtype=torch.int64, device=dev) if want_bytes else None
    out.alg_bytes = _lib.ptr(alg_bytes)
    ray_status = torch.empty(B, height, width, dtype=torch.int32, device=dev) if want_ray_status else None
    out.ray_status = _lib.ptr(ray_status)
    record = rec = None
    if record_cap > 0:
        record = new_record(B, width, height, record_cap, dev)
        rec = _record_struct(record)
        out.record = ctypes.pointer(rec)
    rc = lib.pf_walk_forward(ctypes.byref(scene.c), _lib.ptr(cams_dev), B, width, height,
                             float(transmittance_threshold), float(depth_quantile), int(max_steps),
                             ctypes.byref(out), _lib.stream_ptr())
    _lib.check(rc, "pf_walk_forward")
    return dict(color=color, log_t=log_t, depth=depth, visits=visits, seq_hash=seq_hash, seq=seq, status=status,
                alg_bytes=alg_bytes, record=record, ray_status=ray_status)


# Tiles per band of the banded replay backward: 8 192 tiles of 8x8 rays
# hold 524 288 record slots, about 6.4 GB of workspace at cap 128.  On the
# 1M-cell 1080p frame (32 400 tiles) its backward took 21.9 ms, against
# 20.4 ms in one band and 30.3 ms in bands of 4 096 tiles, whose recording
# forward launches no longer fill the chip.
REPLAY_BAND_TILES = 8192


def launch_walk_backward(scene, cams_dev, width, height, color, log_t, grad_color, grad_log_t,
                         transmittance_threshold=1e-3, max_steps=DEFAULT_MAX_STEPS, grads=None, exact_texel=False,
                         texel_rgb_cam_stride=0, want_bytes=False, record=None, grouped=None, workspace=None,
                         band_tiles=REPLAY_BAND_TILES, replay_cap=128):
    """Gradients of the walk.  record: the forward's record (its "record"
    output), read instead of replaying every ray; None replays.  grouped:
    differentiate the record's segments sorted by cell
    (pf_walk_backward_grouped); None: whenever the record has partials and
    K = 8.  grouped=True with no record: checkpoint and replay through the
    grouped pipeline (pf_walk_backward_replay_grouped), band_tiles 8x8
    tiles at a time, each band's segments (cap replay_cap) held in the
    workspace.  workspace: the device workspace (allocated if None)."""
    lib = _lib.load()
    n, K = scene.num_cells, scene.K
    dev = cams_dev.device
    B = cams_dev.shape[0]
    _expect(cams_dev, (B, 12), "cams_dev", dtype=torch.int32)
    _check_rgb_cams(scene, B)
    _expect(color, (B, height, width, 3), "color", device=dev)
    _expect(log_t, (B, height, width), "log_t", device=dev)
    _expect(grad_color, (B, height, width, 3), "grad_color", device=dev)
    _expect(grad_log_t, (B, height, width), "grad_log_t", device=dev)
    if grads is None:
        per_cam = getattr(scene, "rgb_cams", 0) > 0
        grads = dict(
            spheres=torch.zeros(n, 4, dtype=torch.float32, device=dev),
            nsigmas=torch.zeros(n, 4, dtype=torch.float32, device=dev),
            texel_sites=torch.zeros(n, K, 3, dtype=torch.float32, device=dev),
            texel_rgb=torch.zeros((B, n, K, 3) if per_cam else (n, K, 3), dtype=torch.float32, device=dev),
            texel_height=torch.zeros(n, K, dtype=torch.float32, device=dev),
        )
        if per_cam:
            texel_rgb_cam_stride = n * K * 3
    elif getattr(scene, "rgb_cams", 0) and not texel_rgb_cam_stride:
        if tuple(grads["texel_rgb"].shape) != (B, n, K, 3):
            raise ValueError("per-camera texel colours need per-camera gradients: grads['texel_rgb'] (B, N, K, 3)")
        texel_rgb_cam_stride = n * K * 3
    for k, shape in (("spheres", (n, 4)), ("nsigmas", (n, 4)), ("texel_sites", (n, K, 3)),
                     ("texel_rgb", (n, K, 3)), ("texel_height", (n, K))):
        t = grads[k]
        if k == "texel_rgb" and texel_rgb_cam_stride:
            # camera b's rows start texel_rgb_cam_stride floats after camera b - 1's
            need = (B - 1) * texel_rgb_cam_stride + 3 * K * n
            if t.dtype != torch.float32 or not t.is_contiguous() or t.device != dev or t.numel() < need:
                raise ValueError(f"grads['texel_rgb'] must hold {B} cameras' rows at stride {texel_rgb_cam_stride}")
        else:
            _expect(t, shape, f"grads[{k!r}]", device=dev)
    g = _lib.PfGrads()
    for k in ("spheres", "nsigmas", "texel_sites", "texel_rgb", "texel_height"):
        setattr(g, k, _lib.ptr(grads[k]))
    g.texel_rgb_cam_stride = texel_rgb_cam_stride
    status = torch.zeros(1, dtype=torch.int32, device=dev)
    alg_bytes = torch.zeros(color.shape[:-1], dtype=torch.int64, device=dev) if want_bytes else None
    flags = _lib.PF_BWD_EXACT_TEXEL if exact_texel else 0
    if record is not None and grouped is None:
        grouped = record.get("partials") is not None and K == 8
    if record is not None and grouped:
        if record["counts"].numel() != record_slots(B, width, height):
            raise ValueError("the record belongs to a walk of another size")
        if record.get("partials") is None or K != 8:
            raise ValueError("the grouped backward needs a record with partials and K = 8 texels")
        if workspace is None:
            workspace = grouped_workspace(n, B, width, height, record["cap"], dev)
        rec = _record_struct(record)
        rc = lib.pf_walk_backward_grouped(ctypes.byref(scene.c), _lib.ptr(cams_dev), B, width, height,
                                          float(transmittance_threshold), int(max_steps), _lib.ptr(color),
                                          _lib.ptr(log_t), _lib.ptr(grad_color), _lib.ptr(grad_log_t),
                                          ctypes.byref(rec), ctypes.byref(g), flags, _lib.ptr(workspace),
                                          workspace.numel(), _lib.ptr(status), _lib.ptr(alg_bytes),
                                          _lib.stream_ptr())
        _lib.check(rc, "pf_walk_backward_grouped")
    elif record is None and grouped:
        if K != 8:
            raise ValueError("the grouped backward needs K = 8 texels")
        if want_bytes:
            raise ValueError("the banded replay does not count algorithmic bytes")
        nbytes = lib.pf_walk_replay_grouped_workspace_bytes(int(n), int(B), int(width), int(height), int(replay_cap),
                                                            int(band_tiles))
        if nbytes <= 0:
            raise ValueError("sizes out of range for the banded replay backward")
        if workspace is None:
            workspace = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        rc = lib.pf_walk_backward_replay_grouped(ctypes.byref(scene.c), _lib.ptr(cams_dev), B, width, height,
                                                 float(transmittance_threshold), int(max_steps), _lib.ptr(color),
                                                 _lib.ptr(log_t), _lib.ptr(grad_color), _lib.ptr(grad_log_t),
                                                 ctypes.byref(g), flags, int(replay_cap), int(band_tiles),
                                                 _lib.ptr(workspace), workspace.numel(), _lib.ptr(status),
                                                 _lib.stream_ptr())
        _lib.check(rc, "pf_walk_backward_replay_grouped")
    elif record is not None:
        if record["counts"].numel() != record_slots(cams_dev.shape[0], width, height):
            raise ValueError("the record belongs to a walk of another size")
        rec = _record_struct(record)
        rc = lib.pf_walk_backward_recorded(ctypes.byref(scene.c), _lib.ptr(cams_dev), cams_dev.shape[0], width,
                                           height, float(transmittance_threshold), int(max_steps), _lib.ptr(color),
                                           _lib.ptr(log_t), _lib.ptr(grad_color), _lib.ptr(grad_log_t),
                                           ctypes.byref(rec), ctypes.byref(g), flags, _lib.ptr(status),
                                           _lib.ptr(alg_bytes), _lib.stream_ptr())
        _lib.check(rc, "pf_walk_backward_recorded")
    else:
        rc = lib.pf_walk_backward(ctypes.byref(scene.c), _lib.ptr(cams_dev), cams_dev.shape[0], width, height,
                                  float(transmittance_threshold), int(max_steps), _lib.ptr(color), _lib.ptr(log_t),
                                  _lib.ptr(grad_color), _lib.ptr(grad_log_t), ctypes.byref(g), flags,
                                  _lib.ptr(status), _lib.ptr(alg_bytes), _lib.stream_ptr())
        _lib.check(rc, "pf_walk_backward")
    grads["status"] = status
    grads["alg_bytes"] = alg_bytes
    return grads


class WalkGradFn(torch.autograd.Function):
    """Autograd glue of the walk, shaped like RasterGradFn
    (rasterize.py:23-34, _backward :1865-1984): outputs (color, alpha, depth),
    gradients for (points, radii, density, normals, texel_sites, texel_rgb,
    texel_height); non-finite gradients are zeroed (rasterize.py:1958-1962)."""

    @staticmethod
    def forward(ctx, tracer, cams_dev, width, height, transmittance_threshold, points, radii, density, normals,
                texel_sites, texel_rgb, texel_height, adjacency, adjacency_offsets, adjacency_diff):
        scene = WalkScene(points, radii, density, normals, texel_sites, texel_rgb, texel_height, adjacency,
                          adjacency_offsets, adjacency_diff, attr_dtype=tracer.tscalar)
        cap = tracer.record_cap_for(cams_dev.shape[0], width, height) if any(ctx.needs_input_grad) else 0
        out = launch_walk_forward(scene, cams_dev, width, height, transmittance_threshold,
                                  max_steps=tracer.max_steps, record_cap=cap)
        ctx.scene = scene
        ctx.record = out["record"]
        ctx.cams_dev = cams_dev
        ctx.dims = (width, height, transmittance_threshold, tracer.max_steps)
        ctx.exact_texel = tracer.exact_texel
        ctx.save_for_backward(out["color"], out["log_t"])
        ctx.mark_non_differentiable(out["depth"])
        alpha = 1.0 - torch.exp(out["log_t"])
        return out["color"], alpha, out["depth"]

    @staticmethod
    def backward(ctx, grad_color, grad_alpha, grad_depth):
        color, log_t = ctx.saved_tensors
        width, height, thr, max_steps = ctx.dims
        g_log_t = None
        if grad_alpha is not None:
            g_log_t = (-grad_alpha * torch.exp(log_t)).contiguous()
        gc = grad_color.contiguous() if grad_color is not None else None
        grads = launch_walk_backward(ctx.scene, ctx.cams_dev, width, height, color, log_t, gc, g_log_t, thr,
                                     max_steps, exact_texel=ctx.exact_texel, record=ctx.record)
        for k in ("spheres", "nsigmas", "texel_sites", "texel_rgb", "texel_height"):
            t = grads[k]
            t[~t.isfinite()] = 0
        sp, ns = grads["spheres"], grads["nsigmas"]
        del ctx.scene, ctx.record
        return (None, None, None, None, None, sp[:, :3], sp[:, 3], ns[:, 3], ns[:, :3], grads["texel_sites"],
                grads["texel_rgb"], grads["texel_height"], None, None, None)


class RayTracer:
    """Drop-in for powerfoam.raytrace.RayTracer (raytrace.py:15-309)."""

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
        self.max_steps = getattr(args, "max_walk_steps", DEFAULT_MAX_STEPS) if args is not None else DEFAULT_MAX_STEPS
        # False: the reference's texel-gradient approximations (texture.py:118-119)
        self.exact_texel = bool(getattr(args, "exact_texel_grad", False)) if args is not None else False
        # the forward records each ray's contributing segments (up to this
        # many) so that the backward need not replay the walk; 0 replays
        self.record_cap = int(getattr(args, "walk_record_cap", DEFAULT_RECORD_CAP)) if args is not None \
            else DEFAULT_RECORD_CAP
        self.record_budget = int(getattr(args, "walk_record_budget", DEFAULT_RECORD_BUDGET)) if args is not None \
            else DEFAULT_RECORD_BUDGET
        _lib.load()

    def record_cap_for(self, num_cams, width, height):
        """Segments per ray the differentiable forward records: record_cap,
        lowered to fit record_budget bytes (96 per segment slot: the
        segment and its partials, 32, and its sorted entry in the grouped
        backward's workspace, 64); 0 (replay) below 8."""
        if self.record_cap <= 0:
            return 0
        fit = (self.record_budget - 8 * record_slots(num_cams, width, height)) // \
            (96 * record_slots(num_cams, width, height))
        cap = min(self.record_cap, int(fit)) // RECORD_CHUNK * RECORD_CHUNK
        return cap if cap >= 8 else 0

    def _cams(self, cameras, start_point_idx):
        if not isinstance(cameras, (list, tuple)):
            cameras = [cameras]
        w, h = cameras[0].width, cameras[0].height
        if any(c.width != w or c.height != h for c in cameras):
            raise ValueError("all cameras of one call must share width and height")
        if start_point_idx is None:
            return cameras, None
        if isinstance(start_point_idx, int):
            start_point_idx = [start_point_idx]
        return cameras, list(start_point_idx)

    def benchmark(self, camera, start_point_idx, points, radii, density, normals, texel_sites, texel_rgb,
                  texel_height, adjacency, adjacency_offsets, adjacency_diff, transmittance_threshold=1e-3):
        """raytrace.py:247-309: returns color_out (H, W, 3) f32."""
        scene = WalkScene(points, radii, density, normals, texel_sites, texel_rgb, texel_height, adjacency,
                          adjacency_offsets, adjacency_diff, attr_dtype=self.tscalar)
        cams_dev = cameras_to_device([camera], [int(start_point_idx)], points.device)
        out = launch_walk_forward(scene, cams_dev, camera.width, camera.height, transmittance_threshold,
                                  max_steps=self.max_steps, want_log_t=False, want_depth=False)
        return out["color"][0]

    def forward(self, cameras, start_point_idx, points, radii, density, normals, texel_sites, texel_rgb,
                texel_height, adjacency, adjacency_offsets, adjacency_diff, transmittance_threshold=1e-3):
        """Differentiable batch render.  start_point_idx: int, list of ints,
        or None (start cells found on the GPU).  Returns (color (B,H,W,3),
        alpha (B,H,W), depth (B,H,W))."""
        cameras, starts = self._cams(cameras, start_point_idx)
        cams_dev = cameras_to_device(cameras, starts, points.device)
        if starts is None:
            start_cells_device(points, radii, cams_dev)
        return WalkGradFn.apply(self, cams_dev, cameras[0].width, cameras[0].height, transmittance_threshold,
                                points, radii, density, normals, texel_sites, texel_rgb, texel_height, adjacency,
                                adjacency_offsets, adjacency_diff)

    def trace_sequences(self, cameras, start_point_idx, points, radii, density, normals, texel_sites, texel_rgb,
                        texel_height, adjacency, adjacency_offsets, adjacency_diff, transmittance_threshold=1e-3,
                        seq_cap=0, depth_quantile=0.5):
        """Forward render that also returns each ray's visit count, sequence
        hash and (if seq_cap > 0) its first seq_cap visited cells."""
        cameras, starts = self._cams(cameras, start_point_idx)
        cams_dev = cameras_to_device(cameras, starts, points.device)
        if starts is None:
            start_cells_device(points, radii, cams_dev)
        scene = WalkScene(points, radii, density, normals, texel_sites, texel_rgb, texel_height, adjacency,
                          adjacency_offsets, adjacency_diff, attr_dtype=self.tscalar)
        out = launch_walk_forward(scene, cams_dev, cameras[0].width, cameras[0].height, transmittance_threshold,
                                  depth_quantile=depth_quantile, max_steps=self.max_steps, want_sequences=True,
                                  seq_cap=seq_cap)
        out["cams"] = cams_dev
        return out
