"""FP64 brute-force oracle for the power-cell walk.  TEST INFRASTRUCTURE ONLY.

Independent of the CSR adjacency and of fp16 faces: along a ray
x(t) = o + t d (|d| = 1) the power of site i is
    |x - p_i|^2 - r_i^2 = t^2 + s_i t + c_i,  s_i = 2 d.(o - p_i),
    c_i = |o - p_i|^2 - r_i^2,
so the cell owning x(t) is the lower envelope of the lines s_i t + c_i.  The
visited-cell sequence of the walk (powerfoam/raytrace.py:163-236) is the
sequence of envelope owners from t = 0 (the start cell of
benchmark.py:320-324 is the owner at t = 0), and each owner's interval is
its power cell's chord.  Compositing follows the walk: sphere chord
(rendering_math.py:4-27) with the |c - o| < 4r cull (raytrace.py:173-175),
sigma < 1e-3 skip, texel plane clip and colour (raytrace.py:42-113),
transmittance stop (raytrace.py:164-166), depth at a transmittance quantile
(rasterize.py:871-877).  This is the pattern of the reference's FP64
all-faces oracle (experiments/shared_affine_render.py:29-90) with the
envelope replacing the explicit all-faces clip.
"""
import numpy as np


def camera_rays(cam_words):
    """FP64 pinhole rays of a pf_camera (camera.py:19-24)."""
    w = np.asarray(cam_words, np.int32)
    f = w.view(np.float32).astype(np.float64)
    eye, right, up = f[0:3], f[3:6], f[6:9]
    W, H = int(w[10]), int(w[11])
    fwd = np.cross(up, right)
    fwd /= np.linalg.norm(fwd)
    i, j = np.meshgrid(np.arange(H, dtype=np.float64), np.arange(W, dtype=np.float64), indexing="ij")
    x = 2.0 * j / (W - 1.0) - 1.0
    y = 1.0 - 2.0 * i / (H - 1.0)
    d = x[..., None] * right + y[..., None] * up + fwd
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    return eye, d.reshape(-1, 3)


def owner_sequence(o, d, points, radii, max_steps=100000):
    """Envelope owners along one ray: list of (cell, t_in, t_out) and the
    smallest decision margin (relative gap between the two earliest
    candidate crossings, or between the two lowest lines at t = 0)."""
    p = points.astype(np.float64)
    r = radii.astype(np.float64)
    oc = o[None, :] - p
    s = 2.0 * (oc @ d)
    c = np.einsum("ij,ij->i", oc, oc) - r * r
    order = np.argsort(c, kind="stable")
    cur = int(order[0])
    margin = (c[order[1]] - c[order[0]]) / max(1.0, abs(c[order[0]])) if len(c) > 1 else np.inf
    t = 0.0
    seq = []
    for _ in range(max_steps):
        cand = s < s[cur]
        tj = np.full(len(s), np.inf)
        tj[cand] = (c[cand] - c[cur]) / (s[cur] - s[cand])
        tj[tj <= t] = np.inf
        if not np.isfinite(tj).any():
            seq.append((cur, t, np.inf))
            break
        k = np.argsort(tj, kind="stable")[:2]
        nxt = int(k[0])
        t_out = tj[nxt]
        if len(k) > 1 and np.isfinite(tj[k[1]]):
            margin = min(margin, (tj[k[1]] - t_out) / max(1.0, abs(t_out)))
        seq.append((cur, t, t_out))
        cur, t = nxt, t_out
    return seq, margin


def sphere_chord(o, d, c, r):
    oc = o - c
    qa = d @ d
    b = (oc @ d) / qa
    perp = oc - b * d
    h = (r * r - perp @ perp) / qa
    if h < 0:
        return False, 0.0, 0.0
    sq = np.sqrt(h)
    tf, tn = -b + sq, -b - sq
    if tn < 0 and tf < 0:
        return False, 0.0, 0.0
    if tn < 0:
        return True, 0.0, tf
    return True, tn, tf


def texel(o, d, t_near, c, n, r, sites, rgbs, heights):
    dp = n @ d
    t0 = ((c - o) @ n) / dp
    tq0 = t_near if dp >= 0 else max(t_near, t0)
    q0 = o + tq0 * d
    w = np.exp(-10.0 * np.sum((q0 - sites) ** 2, -1) / (r * r))
    h = (w @ heights) / max(w.sum(), 1e-20)
    ts = ((c - o) @ n + h) / dp
    tq = t_near if dp >= 0 else max(t_near, ts)
    q = o + tq * d
    w = np.exp(-10.0 * np.sum((q - sites) ** 2, -1) / (r * r))
    rgb = (w @ rgbs) / max(w.sum(), 1e-20)
    return ts, dp, rgb


def walk_ray(o, d, scene, thr=1e-3, depth_q=0.5):
    """FP64 composite of one ray; returns (rgb, log_t, depth, seq, margin)."""
    seq, margin = owner_sequence(o, d, scene["points"], scene["radii"])
    rgb = np.zeros(3)
    log_t = 0.0
    depth = 0.0
    found = False
    visited = []
    for cell, t_in, t_out in seq:
        trans = np.exp(log_t)
        if trans < thr:
            break
        visited.append(cell)
        c = scene["points"][cell].astype(np.float64)
        r = float(scene["radii"][cell])
        hit, tn, tf = sphere_chord(o, d, c, r)
        if np.linalg.norm(c - o) < 4.0 * r:
            hit = False
        tn, tf = max(tn, t_in), min(tf, t_out)
        sigma = float(scene["density"][cell])
        if not hit or tn > tf or sigma < 1e-3:
            continue
        n = scene["normals"][cell].astype(np.float64)
        ts, dp, col = texel(o, d, tn, c, n, r, scene["texel_sites"][cell].astype(np.float64),
                            scene["texel_rgb"][cell].astype(np.float64),
                            scene["texel_height"][cell].astype(np.float64))
        if dp >= 0:
            tf = min(ts, tf)
        else:
            tn = max(ts, tn)
        dt = tf - tn
        if dt > 0:
            dlt = -sigma * dt
            alpha = -np.expm1(dlt)
            rgb += col * alpha * trans
            log_t += dlt
            if not found and np.exp(log_t) < depth_q:
                depth = tn + np.log(trans / depth_q) / sigma
                found = True
    return rgb, log_t, depth, visited, margin


def raster_ray(o, d, scene, quantiles=(), gt=None, thr=1e-7):
    """FP64 side outputs of the tile rasterizer's forward for one ray
    (rasterize.py:1088-1129): colour, log T, normal distance
    sum_{n.d > 0} (n.d)^2 alpha T, normal sum alpha T n, the depths where T
    first drops below each quantile (quantiles descending; -1 where never
    reached), and per cell the terms alpha T (contrib before the 1/(H W)
    scale) and alpha T |c - gt|_1 (point_err before the scale).  Cells come
    in envelope order, which is the power-distance order of the raster's
    tile lists (along a ray the envelope owners' power distances from the
    eye increase); a tiny threshold makes the raster's per-tile stop
    irrelevant.  No sigma skip: the raster composites every hit."""
    seq, margin = owner_sequence(o, d, scene["points"], scene["radii"])
    rgb = np.zeros(3)
    normal = np.zeros(3)
    nd = 0.0
    log_t = 0.0
    qd = np.full(len(quantiles), -1.0)
    qi = 0
    cells, contrib, perr = [], [], []
    for cell, t_in, t_out in seq:
        trans = np.exp(log_t)
        if trans < thr:
            break
        c = scene["points"][cell].astype(np.float64)
        r = float(scene["radii"][cell])
        hit, tn, tf = sphere_chord(o, d, c, r)
        if np.linalg.norm(c - o) < 4.0 * r:
            hit = False
        tn, tf = max(tn, t_in), min(tf, t_out)
        if not hit or tn > tf:
            continue
        sigma = float(scene["density"][cell])
        n = scene["normals"][cell].astype(np.float64)
        ts, dp, col = texel(o, d, tn, c, n, r, scene["texel_sites"][cell].astype(np.float64),
                            scene["texel_rgb"][cell].astype(np.float64),
                            scene["texel_height"][cell].astype(np.float64))
        if dp >= 0:
            tf = min(ts, tf)
        else:
            tn = max(ts, tn)
        dt = tf - tn
        if dt <= 0:
            continue
        alpha = -np.expm1(-sigma * dt)
        w = alpha * trans
        rgb += col * w
        log_t -= sigma * dt
        ndv = float(n @ d)
        if ndv > 0:
            nd += ndv * ndv * w
        normal += w * n
        cells.append(cell)
        contrib.append(w)
        perr.append(w * (np.abs(col - gt).sum() if gt is not None else 0.0))
        nxt = np.exp(log_t)
        while qi < len(quantiles) and nxt < quantiles[qi]:
            qd[qi] = tn + np.log(trans / quantiles[qi]) / sigma
            qi += 1
    return dict(color=rgb, log_t=log_t, normal_distance=nd, normal=normal, qdepth=qd, cells=cells,
                contrib=np.array(contrib), point_err=np.array(perr), margin=margin)


def raster_render(scene, cam_words, quantiles=(), ray_gt=None, thr=1e-7, pixels=None):
    """FP64 raster side outputs over the given flat pixel ids (default all);
    contrib and point_err are per cell and carry the 1/(H W) scale."""
    eye, dirs = camera_rays(cam_words)
    w = np.asarray(cam_words, np.int32)
    H, W = int(w[11]), int(w[10])
    ids = np.arange(dirs.shape[0]) if pixels is None else np.asarray(pixels)
    n = len(scene["radii"])
    out = dict(color=np.zeros((ids.size, 3)), log_t=np.zeros(ids.size), normal_distance=np.zeros(ids.size),
               normal=np.zeros((ids.size, 3)), qdepth=np.zeros((ids.size, len(quantiles))),
               contrib=np.zeros(n), point_err=np.zeros(n), margin=np.zeros(ids.size))
    for k, pid in enumerate(ids):
        gt = None if ray_gt is None else np.asarray(ray_gt, np.float64).reshape(-1, 3)[pid]
        r = raster_ray(eye, dirs[pid], scene, quantiles, gt, thr)
        for key in ("color", "log_t", "normal_distance", "normal", "qdepth", "margin"):
            out[key][k] = r[key]
        np.add.at(out["contrib"], r["cells"], r["contrib"] / (H * W))
        np.add.at(out["point_err"], r["cells"], r["point_err"] / (H * W))
    return out


def render(scene, cam_words, thr=1e-3, depth_q=0.5, pixels=None):
    """FP64 render of the given flat pixel ids (default: all)."""
    eye, dirs = camera_rays(cam_words)
    ids = np.arange(dirs.shape[0]) if pixels is None else np.asarray(pixels)
    out = dict(color=np.zeros((ids.size, 3)), log_t=np.zeros(ids.size), depth=np.zeros(ids.size), seq=[],
               margin=np.zeros(ids.size))
    for k, pid in enumerate(ids):
        rgb, lt, dep, vis, m = walk_ray(eye, dirs[pid], scene, thr, depth_q)
        out["color"][k] = rgb
        out["log_t"][k] = lt
        out["depth"][k] = dep
        out["seq"].append(vis)
        out["margin"][k] = m
    return out


# Real spherical harmonics, degree <= 3, in FP64 with the constants and
# signs of powerfoam/sh_utils.py:4-74 (the 3DGS ordering).
SH_C0 = 0.28209479177387814
SH_C1 = 0.4886025119029199
SH_C2 = (1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792, 0.5462742152960396)
SH_C3 = (-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154, -0.4570457994644658,
         1.445305721320277, -0.5900435899266435)


def sh_basis_fp64(v, degree):
    """Basis values at the directions v / |v|: (..., (degree+1)^2)."""
    v = np.asarray(v, np.float64)
    v = v / np.linalg.norm(v, axis=-1, keepdims=True)
    x, y, z = v[..., 0], v[..., 1], v[..., 2]
    Y = [np.full(x.shape, SH_C0)]
    if degree >= 1:
        Y += [-SH_C1 * y, SH_C1 * z, -SH_C1 * x]
    if degree >= 2:
        xx, yy, zz = x * x, y * y, z * z
        Y += [SH_C2[0] * x * y, SH_C2[1] * y * z, SH_C2[2] * (2 * zz - xx - yy), SH_C2[3] * x * z,
              SH_C2[4] * (xx - yy)]
    if degree >= 3:
        Y += [SH_C3[0] * y * (3 * xx - yy), SH_C3[1] * x * y * z, SH_C3[2] * y * (4 * zz - xx - yy),
              SH_C3[3] * z * (2 * zz - 3 * xx - 3 * yy), SH_C3[4] * x * (4 * zz - xx - yy), SH_C3[5] * z * (xx - yy),
              SH_C3[6] * x * (xx - 3 * yy)]
    return np.stack(Y, -1)
