"""ctypes access to the CPU restatement ``oracle/build/libpf_oracle.so``.

TEST INFRASTRUCTURE: only tests/, the smoke check and bench.py's CPU
baseline use this module, as the oracle.  The product path never imports
it.  Arrays are numpy (host); f16 attribute arrays are widened to f32 by the
caller (an exact conversion), f16 faces are passed as uint16 bits.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "build", "libpf_oracle.so")

c_void_p, c_int32, c_int64, c_float, c_double = (ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_float,
                                                 ctypes.c_double)


class SceneIn(ctypes.Structure):
    _fields_ = [
        ("spheres", c_void_p), ("nsigmas", c_void_p), ("texel_sites", c_void_p), ("texel_rgb", c_void_p),
        ("texel_height", c_void_p), ("adjacency", c_void_p), ("offsets", c_void_p), ("faces", c_void_p),
        ("num_cells", c_int64), ("num_edges", c_int64), ("K", c_int32), ("face_dtype", c_int32),
    ]


_lib = None


def build():
    r = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    return LIB


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        lib.orc_num_threads.restype = c_int32
        lib.orc_start_cell.restype = c_int32
        lib.orc_walk_forward.restype = c_int32
        lib.orc_walk_backward.restype = c_int32
        lib.orc_walk_loss_f64.restype = c_double
        lib.orc_walk_loss_f64.argtypes = [c_void_p] * 6 + [c_int32, c_int64, c_void_p, c_void_p, c_int64,
                                                          c_void_p, c_int32, c_double, c_int32, c_void_p, c_int64,
                                                          c_void_p, c_double, c_int32, c_int64]
        lib.orc_walk_forward.argtypes = [c_void_p, c_void_p, c_int32, c_float, c_float, c_int32, c_void_p, c_int64,
                                         c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int32,
                                         c_int32, c_void_p]
        lib.orc_walk_backward.argtypes = [c_void_p, c_void_p, c_int32, c_float, c_int32] + [c_void_p] * 8 + [c_int32, c_int32]
        lib.orc_start_cell.argtypes = [c_void_p, c_void_p, c_int64, c_void_p]
        lib.orc_walk_faces.argtypes = [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_int32]
        lib.orc_ray_dirs.argtypes = [c_void_p, c_void_p]
        for f in ("orc_f32_to_f16", "orc_f16_to_f32", "orc_expf_vec", "orc_logf_vec"):
            getattr(lib, f).argtypes = [c_void_p, c_int64, c_void_p]
        _lib = lib
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def num_threads():
    return load().orc_num_threads()


def f32_to_f16_bits(x):
    x = _c(x, np.float32)
    out = np.empty(x.shape, np.uint16)
    load().orc_f32_to_f16(_p(x), x.size, _p(out))
    return out


def f16_bits_to_f32(x):
    x = _c(x, np.uint16)
    out = np.empty(x.shape, np.float32)
    load().orc_f16_to_f32(_p(x), x.size, _p(out))
    return out


def expf(x):
    x = _c(x, np.float32)
    out = np.empty_like(x)
    load().orc_expf_vec(_p(x), x.size, _p(out))
    return out


def logf(x):
    x = _c(x, np.float32)
    out = np.empty_like(x)
    load().orc_logf_vec(_p(x), x.size, _p(out))
    return out


def ray_dirs(cam_words):
    cam = _c(cam_words, np.int32)
    H, W = int(cam[11]), int(cam[10])
    out = np.empty((H, W, 3), np.float32)
    load().orc_ray_dirs(_p(cam), _p(out))
    return out


def walk_faces(points, radii, adjacency, offsets, f16=True):
    points, radii = _c(points, np.float32), _c(radii, np.float32)
    adjacency, offsets = _c(adjacency, np.int32), _c(offsets, np.int32)
    out = np.empty((adjacency.size, 4), np.uint16 if f16 else np.float32)
    load().orc_walk_faces(_p(points), _p(radii), points.shape[0], _p(adjacency), _p(offsets), _p(out),
                          1 if f16 else 0)
    return out


def start_cell(points, radii, eye):
    points, radii, eye = _c(points, np.float32), _c(radii, np.float32), _c(eye, np.float32)
    return int(load().orc_start_cell(_p(points), _p(radii), points.shape[0], _p(eye)))


class Scene:
    """Host scene in the kernels' layout; keeps the numpy arrays alive."""

    def __init__(self, points, radii, density, normals, texel_sites, texel_rgb, texel_height, adjacency, offsets,
                 faces):
        self.spheres = _c(np.concatenate([points, radii[:, None]], 1), np.float32)
        self.nsigmas = _c(np.concatenate([normals, density[:, None]], 1), np.float32)
        self.texel_sites = _c(texel_sites, np.float32)
        self.texel_rgb = _c(texel_rgb, np.float32)
        self.texel_height = _c(texel_height, np.float32)
        self.adjacency = _c(adjacency, np.int32)
        self.offsets = _c(offsets, np.int32)
        self.faces = np.ascontiguousarray(faces)
        f16 = self.faces.dtype == np.uint16
        if not f16:
            self.faces = _c(self.faces, np.float32)
        s = SceneIn()
        s.spheres, s.nsigmas = _p(self.spheres), _p(self.nsigmas)
        s.texel_sites, s.texel_rgb, s.texel_height = _p(self.texel_sites), _p(self.texel_rgb), _p(self.texel_height)
        s.adjacency, s.offsets, s.faces = _p(self.adjacency), _p(self.offsets), _p(self.faces)
        s.num_cells = self.spheres.shape[0]
        s.num_edges = self.adjacency.size
        s.K = self.texel_sites.shape[1]
        s.face_dtype = 1 if f16 else 0
        self.c = s
        self.K = s.K
        self.N = s.num_cells


def walk_forward(scene, cams, thr=1e-3, depth_q=0.5, max_steps=1 << 16, rays=None, seq_cap=0, nthreads=None):
    """cams: (B, 12) int32 pf_camera words.  Returns dict of per-ray arrays
    shaped (B, H, W, ...) or (n_rays, ...) when rays is given."""
    lib = load()
    cams = _c(cams, np.int32).reshape(-1, 12)
    B, W, H = cams.shape[0], int(cams[0, 10]), int(cams[0, 11])
    if rays is not None:
        rays = _c(rays, np.int64)
        shape = (rays.size,)
    else:
        shape = (B, H, W)
    color = np.empty(shape + (3,), np.float32)
    log_t = np.empty(shape, np.float32)
    depth = np.empty(shape, np.float32)
    visits = np.empty(shape, np.int32)
    hsh = np.empty(shape, np.uint64)
    seq = np.full(shape + (seq_cap,), -1, np.int32) if seq_cap > 0 else None
    ray_status = np.empty(shape, np.int32)
    st = lib.orc_walk_forward(ctypes.byref(scene.c), _p(cams), B, thr, depth_q, max_steps, _p(rays),
                              0 if rays is None else rays.size, _p(color), _p(log_t), _p(depth), _p(visits), _p(hsh),
                              _p(seq), seq_cap, nthreads or num_threads(), _p(ray_status))
    return dict(color=color, log_t=log_t, depth=depth, visits=visits, seq_hash=hsh, seq=seq, status=st,
                ray_status=ray_status)


def walk_backward(scene, cams, color, grad_color, grad_log_t=None, thr=1e-3, max_steps=1 << 16, exact=False,
                  nthreads=1):
    """nthreads 1: fixed summation order; > 1: atomic adds (timing only)."""
    lib = load()
    cams = _c(cams, np.int32).reshape(-1, 12)
    n, K = scene.N, scene.K
    g = dict(spheres=np.zeros((n, 4), np.float32), nsigmas=np.zeros((n, 4), np.float32),
             texel_sites=np.zeros((n, K, 3), np.float32), texel_rgb=np.zeros((n, K, 3), np.float32),
             texel_height=np.zeros((n, K), np.float32))
    color = _c(color, np.float32)
    gc = None if grad_color is None else _c(grad_color, np.float32)
    gl = None if grad_log_t is None else _c(grad_log_t, np.float32)
    st = lib.orc_walk_backward(ctypes.byref(scene.c), _p(cams), cams.shape[0], thr, max_steps, _p(color), _p(gc),
                               _p(gl), _p(g["spheres"]), _p(g["nsigmas"]), _p(g["texel_sites"]),
                               _p(g["texel_rgb"]), _p(g["texel_height"]), 1 if exact else 0,
                               nthreads)
    g["status"] = st
    return g


def walk_loss_f64(params, adjacency, offsets, cams, target, thr=1e-3, w_alpha=0.0, rays=None, max_steps=1 << 16,
                  nthreads=None):
    """FP64 forward loss (exact faces) for finite differences.  params: dict
    of float64 arrays points, radii, nsigmas (N,4), texel_sites, texel_rgb
    (N,K,3), or (B,N,K,3) with each camera's own colours, texel_height."""
    lib = load()
    P = {k: _c(v, np.float64) for k, v in params.items()}
    adjacency, offsets = _c(adjacency, np.int32), _c(offsets, np.int32)
    cams = _c(cams, np.int32).reshape(-1, 12)
    target = _c(target, np.float32)
    K = P["texel_sites"].shape[1]
    N = P["points"].shape[0]
    r = None if rays is None else _c(rays, np.int64)
    stride = 0
    if P["texel_rgb"].ndim == 4:
        if P["texel_rgb"].shape[0] != cams.shape[0]:
            raise ValueError("per-camera texel_rgb needs one set of colours per camera")
        stride = N * K * 3
    return lib.orc_walk_loss_f64(_p(P["points"]), _p(P["radii"]), _p(P["nsigmas"]), _p(P["texel_sites"]),
                                 _p(P["texel_rgb"]), _p(P["texel_height"]), K, N, _p(adjacency), _p(offsets),
                                 adjacency.size, _p(cams), cams.shape[0], thr, max_steps, _p(r),
                                 0 if r is None else r.size, _p(target), w_alpha, nthreads or num_threads(), stride)


# ------------------------------------------------------------ rasterizer --
class RScene(ctypes.Structure):
    _fields_ = [("spheres", c_void_p), ("nsigmas", c_void_p), ("sites", c_void_p), ("rgb", c_void_p),
                ("height", c_void_p), ("adj", c_void_p), ("off", c_void_p), ("faces", c_void_p), ("N", c_int64),
                ("K", c_int32), ("face_f16", c_int32)]


def _rlib():
    lib = load()
    if not getattr(lib, "_raster_sigs", False):
        lib.orc_raster_faces.argtypes = [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_int32]
        lib.orc_raster_bin.restype = c_int64
        lib.orc_raster_bin.argtypes = [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_int64]
        lib.orc_raster_forward.argtypes = ([c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_void_p, c_int32,
                                            c_void_p, c_int32, c_int32] + [c_void_p] * 9 + [c_int32])
        lib.orc_raster_backward.argtypes = ([c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int32]
                                            + [c_void_p] * 12 + [c_int32, c_int32])
        lib.orc_raster_loss_f64.restype = c_double
        lib.orc_raster_loss_f64.argtypes = ([c_void_p] * 6 + [c_int32, c_int64, c_void_p, c_void_p, c_int64,
                                             c_void_p, c_double, c_void_p, c_double, c_void_p, c_int32])
        lib.orc_raster_plain.argtypes = ([c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_int32, c_int32,
                                          c_float, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                          c_int32])
        lib._raster_sigs = True
    return lib


class RasterScene:
    def __init__(self, points, radii, density, normals, texel_sites, texel_rgb, texel_height, adjacency, offsets,
                 faces):
        self.spheres = _c(np.concatenate([points, radii[:, None]], 1), np.float32)
        self.nsigmas = _c(np.concatenate([normals, density[:, None]], 1), np.float32)
        self.sites, self.rgb = _c(texel_sites, np.float32), _c(texel_rgb, np.float32)
        self.height = _c(texel_height, np.float32)
        self.adj, self.off = _c(adjacency, np.int32), _c(offsets, np.int32)
        self.faces = np.ascontiguousarray(faces)
        f16 = self.faces.dtype == np.uint16
        if not f16:
            self.faces = _c(self.faces, np.float32)
        s = RScene()
        s.spheres, s.nsigmas, s.sites, s.rgb, s.height = (_p(self.spheres), _p(self.nsigmas), _p(self.sites),
                                                          _p(self.rgb), _p(self.height))
        s.adj, s.off, s.faces = _p(self.adj), _p(self.off), _p(self.faces)
        s.N = self.spheres.shape[0]
        s.K = self.sites.shape[1]
        s.face_f16 = 1 if f16 else 0
        self.c = s
        self.N, self.K = s.N, s.K


def raster_faces(spheres, adjacency, offsets, f16=True):
    spheres = _c(spheres, np.float32)
    adjacency, offsets = _c(adjacency, np.int32), _c(offsets, np.int32)
    out = np.empty((adjacency.size, 4), np.uint16 if f16 else np.float32)
    _rlib().orc_raster_faces(_p(spheres), spheres.shape[0], _p(adjacency), _p(offsets), _p(out), 1 if f16 else 0)
    return out


def raster_bin(spheres, cam_words):
    lib = _rlib()
    spheres = _c(spheres, np.float32)
    cam = _c(cam_words, np.int32).reshape(12)
    W, H = int(cam[10]), int(cam[11])
    tiles = ((H + 7) // 8) * ((W + 7) // 8)
    offsets = np.zeros(tiles + 1, np.int64)
    n = lib.orc_raster_bin(_p(spheres), spheres.shape[0], _p(cam), _p(offsets), None, 0)
    prims = np.zeros(max(n, 1), np.int32)
    lib.orc_raster_bin(_p(spheres), spheres.shape[0], _p(cam), _p(offsets), _p(prims), n)
    return offsets, prims[:n]


def raster_forward(scene, cam_words, offsets, prims, thr=1e-3, depth_q=None, ray_gt=None, point_err=False,
                   surface=False, nthreads=None):
    lib = _rlib()
    cam = _c(cam_words, np.int32).reshape(12)
    W, H = int(cam[10]), int(cam[11])
    tiles = offsets.size - 1
    Q = 0 if depth_q is None else depth_q.shape[-1]
    dq = None if depth_q is None else _c(depth_q, np.float32)
    gt = None if ray_gt is None else _c(ray_gt, np.float32)
    out = dict(color=np.zeros((H, W, 3), np.float32), log_t=np.zeros((H, W), np.float32),
               ndist=np.zeros((H, W), np.float32), normal=np.zeros((H, W, 3), np.float32),
               qdepth=None if dq is None else -np.ones(dq.shape, np.float32),
               err=np.zeros((H, W), np.float32) if surface else None, contrib=np.zeros(scene.N, np.float32),
               perr=np.zeros(scene.N, np.float32) if point_err else None, early=np.zeros(tiles, np.int32))
    offsets, prims = _c(offsets, np.int64), _c(prims, np.int32)
    lib.orc_raster_forward(ctypes.byref(scene.c), _p(cam), _p(offsets), _p(prims), thr, _p(dq), Q, _p(gt),
                           int(point_err), int(surface), _p(out["color"]), _p(out["log_t"]), _p(out["ndist"]),
                           _p(out["normal"]), _p(out["qdepth"]), _p(out["err"]), _p(out["contrib"]),
                           _p(out["perr"]), _p(out["early"]), nthreads or num_threads())
    return out


def raster_backward(scene, cam_words, offsets, prims, early, log_t, g_color, g_log_t=None, g_ndist=None,
                    g_normal=None, depth_q=None, g_qdepth=None, g_contrib=None, exact=False, nthreads=None):
    """exact: keep every texel and the radius term of the texel weights (the
    reference drops both, texture.py:118-119, :158-159)."""
    lib = _rlib()
    cam = _c(cam_words, np.int32).reshape(12)
    n, K = scene.N, scene.K
    g = dict(spheres=np.zeros((n, 4), np.float32), nsigmas=np.zeros((n, 4), np.float32),
             texel_sites=np.zeros((n, K, 3), np.float32), texel_rgb=np.zeros((n, K, 3), np.float32),
             texel_height=np.zeros((n, K), np.float32))
    cv = lambda a: None if a is None else _c(a, np.float32)
    args = [cv(depth_q), cv(log_t), cv(g_color), cv(g_log_t), cv(g_ndist), cv(g_normal), cv(g_qdepth), cv(g_contrib)]
    Q = 0 if depth_q is None else depth_q.shape[-1]
    lib.orc_raster_backward(ctypes.byref(scene.c), _p(cam), _p(_c(offsets, np.int64)), _p(_c(prims, np.int32)),
                            _p(_c(early, np.int32)), _p(args[0]), Q, *[_p(a) for a in args[1:]],
                            _p(g["spheres"]), _p(g["nsigmas"]), _p(g["texel_sites"]), _p(g["texel_rgb"]),
                            _p(g["texel_height"]), 1 if exact else 0, nthreads or num_threads())
    return g


class _RasterWeights(ctypes.Structure):
    _fields_ = [("g_nd", c_void_p), ("nd_pow", c_int32), ("g_normal", c_void_p), ("quantiles", c_void_p),
                ("g_q", c_void_p), ("Q", c_int32), ("g_contrib", c_void_p)]


def raster_loss_f64(params, adjacency, offsets, cam_words, target, w_alpha=0.0, g_ndist=None, nd_pow=2,
                    g_normal=None, depth_q=None, g_qdepth=None, g_contrib=None, thr=1e-7, nthreads=None):
    """FP64 value of sum_px [|C - target|^2 + w_alpha (1 - T) + g_ndist nd
    + g_normal . normal + sum_q g_qdepth_q depth_q] + sum_cells g_contrib
    contrib, the raster's outputs evaluated along the exact-face walk of
    the power CSR (one camera).  nd is sum_{n.d>0} (n.d)^nd_pow alpha T."""
    lib = _rlib()
    P = {k: _c(v, np.float64) for k, v in params.items()}
    adjacency, offsets = _c(adjacency, np.int32), _c(offsets, np.int32)
    cam = _c(cam_words, np.int32).reshape(12)
    d64 = lambda a: None if a is None else _c(a, np.float64)
    keep = [d64(g_ndist), d64(g_normal), None if depth_q is None else _c(depth_q, np.float32), d64(g_qdepth),
            d64(g_contrib)]
    rw = _RasterWeights(_p(keep[0]), nd_pow, _p(keep[1]), _p(keep[2]), _p(keep[3]),
                        0 if depth_q is None else depth_q.shape[-1], _p(keep[4]))
    K = P["texel_sites"].shape[1]
    N = P["points"].shape[0]
    return lib.orc_raster_loss_f64(_p(P["points"]), _p(P["radii"]), _p(P["nsigmas"]), _p(P["texel_sites"]),
                                   _p(P["texel_rgb"]), _p(P["texel_height"]), K, N, _p(adjacency), _p(offsets),
                                   adjacency.size, _p(cam), thr, _p(_c(target, np.float32)), w_alpha,
                                   ctypes.byref(rw), nthreads or num_threads())


def raster_plain(scene, cam_words, offsets, prims, thr=1e-3, vis=False, max_isect=1024, depth_qv=0.5,
                 bkgd=(0.0, 0.0, 0.0), nthreads=None):
    lib = _rlib()
    cam = _c(cam_words, np.int32).reshape(12)
    W, H = int(cam[10]), int(cam[11])
    out = dict(color=np.zeros((H, W, 3), np.float32), depth=np.zeros((H, W), np.float32),
               normal=np.zeros((H, W, 3), np.float32), alpha=np.zeros((H, W), np.float32),
               isect=np.zeros((H, W), np.int32))
    bk = np.asarray(bkgd, np.float32)
    lib.orc_raster_plain(ctypes.byref(scene.c), _p(cam), _p(_c(offsets, np.int64)), _p(_c(prims, np.int32)), thr,
                         int(vis), max_isect, depth_qv, _p(bk), _p(out["color"]), _p(out["depth"]), _p(out["normal"]),
                         _p(out["alpha"]), _p(out["isect"]), nthreads or num_threads())
    return out


def cech(centers, radii):
    """Brute-force Čech CSR (rows ascending): (adjacency int32, offsets int32)."""
    lib = load()
    lib.orc_cech.restype = c_int64
    lib.orc_cech.argtypes = [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_int64]
    centers, radii = _c(centers, np.float32), _c(radii, np.float32)
    n = centers.shape[0]
    off = np.zeros(n + 1, np.int32)
    total = lib.orc_cech(_p(centers), _p(radii), n, _p(off), None, 0)
    adj = np.zeros(max(total, 1), np.int32)
    lib.orc_cech(_p(centers), _p(radii), n, _p(off), _p(adj), total)
    return adj[:total], off


def _sv_args(sites, eye, fwd, axes, vals, temps):
    sites = _c(sites, np.float32).reshape(-1, 3)
    axes, vals, temps = _c(axes, np.float32), _c(vals, np.float32), _c(temps, np.float32)
    dof, n = temps.shape
    if axes.shape != (dof, n, 3) or vals.shape != (dof, n, 3) or sites.shape[0] != n:
        raise ValueError("lobe arrays must be (dof, n, 3) / (dof, n) with n texel sites")
    return sites, _c(eye, np.float32), _c(fwd, np.float32), axes, vals, temps, dof, n


def sv_forward(sites, eye, fwd, cutoff, axes, vals, temps, f64=False):
    """Spherical-Voronoi colour per texel site -> (value (n,3), wsum (n,))."""
    lib = load()
    sites, eye, fwd, axes, vals, temps, dof, n = _sv_args(sites, eye, fwd, axes, vals, temps)
    dt = np.float64 if f64 else np.float32
    fn = lib.orc_sv_forward_f64 if f64 else lib.orc_sv_forward
    fn.restype = None
    fn.argtypes = [c_void_p, c_int64, c_void_p, c_void_p, ctypes.c_float] + [c_void_p] * 3 + [ctypes.c_int32] + \
        [c_void_p] * 2
    value, wsum = np.zeros((n, 3), dt), np.zeros(n, dt)
    fn(_p(sites), n, _p(eye), _p(fwd), cutoff, _p(axes), _p(vals), _p(temps), dof, _p(value), _p(wsum))
    return value, wsum


def sv_backward(sites, eye, fwd, cutoff, axes, vals, temps, wsum, value, grad, f64=False):
    """Gradients (axes, vals, temps) of sum(grad * value); unset lobes stay 0."""
    lib = load()
    sites, eye, fwd, axes, vals, temps, dof, n = _sv_args(sites, eye, fwd, axes, vals, temps)
    dt = np.float64 if f64 else np.float32
    fn = lib.orc_sv_backward_f64 if f64 else lib.orc_sv_backward
    fn.restype = None
    fn.argtypes = [c_void_p, c_int64, c_void_p, c_void_p, ctypes.c_float] + [c_void_p] * 3 + [ctypes.c_int32] + \
        [c_void_p] * 6
    ga, gv, gt = np.zeros_like(axes, dt), np.zeros_like(vals, dt), np.zeros_like(temps, dt)
    fn(_p(sites), n, _p(eye), _p(fwd), cutoff, _p(axes), _p(vals), _p(temps), dof, _p(_c(wsum, dt)),
       _p(_c(value, dt)), _p(_c(grad, np.float32)), _p(ga), _p(gv), _p(gt))
    return ga, gv, gt


def morton_codes(points):
    """Morton codes as morton_code_kernel (geometry.py:18-47) computes them:
    f32 normalisation over the bounding box, 2^21 steps per axis clamped and
    truncated, bits interleaved with x highest (expand_bits, geometry.py:7-15).
    int64 (N,)."""
    p = np.asarray(points, dtype=np.float32)
    if p.shape[0] == 0:
        return np.zeros(0, np.int64)
    lo, hi = p.min(0), p.max(0)
    ext = np.maximum(hi - lo, np.float32(1e-6))
    res = np.float32(2097152.0)
    q = np.clip((p - lo) / ext * res, np.float32(0.0), res - np.float32(1.0)).astype(np.uint64)

    def spread(v):
        v = v & np.uint64(0x1FFFFF)
        for shift, mask in ((32, 0x1F00000000FFFF), (16, 0x1F0000FF0000FF), (8, 0x100F00F00F00F00F),
                            (4, 0x10C30C30C30C30C3), (2, 0x1249249249249249)):
            v = (v | (v << np.uint64(shift))) & np.uint64(mask)
        return v

    return ((spread(q[:, 0]) << np.uint64(2)) | (spread(q[:, 1]) << np.uint64(1)) | spread(q[:, 2])).astype(np.int64)
