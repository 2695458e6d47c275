"""Raster binning without a host read-back: with a fixed list capacity the
whole raster forward (bin count, bounded scan, bin write + per-tile sort,
raster) runs from device memory, is captured in a HIP graph, and every
replay (with eager work in between) equals the eager result and the
unbounded path; a capacity below the need raises PF_STATUS_LIST_CAP and
keeps every access inside the buffers.  The forward, the L2 loss's incoming
gradient and the backward (rasterize.py:1150-1516) captured in one graph
give, on every replay, the eager pass's image bit for bit and its
gradients, which equal autograd through Rasterizer.forward (gradient sums
are atomics: 1e-5 relative L2)."""
import ctypes
import types

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from powerfoam import _lib, workloads
from powerfoam.camera import cameras_to_device
from powerfoam.rasterize import Rasterizer, _scene_struct, bin_tiles, raster_faces

pytestmark = pytest.mark.gpu
DEV = "cuda"
ARGS = types.SimpleNamespace(num_texel_sites=8)


def t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


class Pass:
    """Preallocated raster forward over one camera with a list capacity."""

    def __init__(self, sc, adj, off, cam, capacity):
        self.W, self.H = cam.width, cam.height
        self.spheres = t(np.concatenate([sc["points"], sc["radii"][:, None]], 1).astype(np.float32))
        self.nsig = t(np.concatenate([sc["normals"], sc["density"][:, None]], 1).astype(np.float32))
        self.sites, self.rgb, self.h = t(sc["texel_sites"]), t(sc["texel_rgb"]), t(sc["texel_height"])
        self.adj, self.off = t(adj), t(off)
        self.faces = raster_faces(self.spheres, self.adj, self.off)
        self.scene = _scene_struct(self.spheres, self.nsig, self.sites, self.rgb, self.h, self.adj, self.off,
                                   self.faces, _lib.PF_F32, _lib.PF_F16)
        self.cam = cameras_to_device([cam], None, DEV)
        tiles = ((self.H + 7) // 8) * ((self.W + 7) // 8)
        n = self.spheres.shape[0]
        self.cap = capacity
        self.counts = torch.zeros(tiles + 1, dtype=torch.int32, device=DEV)
        self.offsets = torch.zeros(tiles + 1, dtype=torch.int64, device=DEV)
        self.keys = torch.zeros(capacity, dtype=torch.int64, device=DEV)
        self.scratch = torch.zeros(capacity, dtype=torch.int64, device=DEV)
        self.prims = torch.zeros(capacity, dtype=torch.int32, device=DEV)
        self.wc = torch.zeros(tiles, dtype=torch.int32, device=DEV)
        self.mask = torch.zeros(n, dtype=torch.uint8, device=DEV)
        self.early = torch.zeros(tiles, dtype=torch.int32, device=DEV)
        self.status = torch.zeros(1, dtype=torch.int32, device=DEV)
        self.color = torch.zeros(self.H, self.W, 3, device=DEV)
        self.log_t = torch.zeros(self.H, self.W, device=DEV)
        self.ndist = torch.zeros(self.H, self.W, device=DEV)
        self.normal = torch.zeros(self.H, self.W, 3, device=DEV)
        self.contrib = torch.zeros(n, device=DEV)
        self.tiles = tiles

    def run(self):
        lib, st, P = _lib.load(), _lib.stream_ptr(), _lib.ptr
        for x in (self.counts, self.wc, self.mask, self.early, self.color, self.log_t, self.ndist, self.normal,
                  self.contrib):
            x.zero_()  # fill kernels, not memsets, inside a graph
        n = self.spheres.shape[0]
        _lib.check(lib.pf_raster_bin_count(P(self.spheres), n, P(self.cam), P(self.counts), st), "count")
        _lib.check(lib.pf_raster_scan_bounded(P(self.counts), self.tiles + 1, P(self.offsets), self.cap,
                                              P(self.status), st), "scan")
        _lib.check(lib.pf_raster_bin_write(P(self.spheres), n, P(self.cam), self.tiles, P(self.offsets), P(self.wc),
                                           P(self.keys), P(self.scratch), P(self.prims), P(self.mask), self.cap, st),
                   "write")
        out = _lib.PfRasterOut()
        out.color, out.log_t, out.normal_distance, out.normal = P(self.color), P(self.log_t), P(self.ndist), \
            P(self.normal)
        out.contrib = P(self.contrib)
        _lib.check(lib.pf_raster_forward(ctypes.byref(self.scene), P(self.cam), self.W, self.H, P(self.offsets),
                                         P(self.prims), P(self.early), 1e-3, None, 0, None, 0, 0,
                                         ctypes.byref(out), st), "forward")

    def backward_buffers(self, target):
        n, K = self.spheres.shape[0], self.sites.shape[1]
        self.target = target.contiguous()
        self.g_color = torch.zeros(self.H, self.W, 3, device=DEV)
        self.grads = dict(spheres=torch.zeros(n, 4, device=DEV), nsigmas=torch.zeros(n, 4, device=DEV),
                          texel_sites=torch.zeros(n, K, 3, device=DEV), texel_rgb=torch.zeros(n, K, 3, device=DEV),
                          texel_height=torch.zeros(n, K, device=DEV))

    def run_backward(self):
        """dL/dC = 2 (C - target), then the raster backward into zeroed
        gradient arrays; no allocation, no host read-back."""
        lib, st, P = _lib.load(), _lib.stream_ptr(), _lib.ptr
        torch.sub(self.color, self.target, out=self.g_color)
        self.g_color.mul_(2.0)
        for x in self.grads.values():
            x.zero_()
        gin = _lib.PfRasterGradIn()
        gin.color = P(self.g_color)
        gs = _lib.PfGrads()
        for k, v in self.grads.items():
            setattr(gs, k, P(v))
        _lib.check(lib.pf_raster_backward(ctypes.byref(self.scene), P(self.cam), self.W, self.H, P(self.offsets),
                                          P(self.prims), P(self.early), None, 0, P(self.log_t), ctypes.byref(gin),
                                          ctypes.byref(gs), st), "backward")


@pytest.fixture(scope="module")
def setup(scene4k):
    sc = scene4k
    adj, off = orc.cech(sc["points"], sc["radii"])
    return sc, adj, off, workloads.front_camera(80, 64)


def test_bounded_pass_equals_unbounded_and_replays(setup):
    sc, adj, off, cam = setup
    rz = Rasterizer(ARGS, DEV)
    with torch.no_grad():
        ref = rz.forward(cam, None, t(sc["points"]), t(sc["radii"]), t(sc["density"]), t(sc["normals"]),
                         t(sc["texel_sites"]), t(sc["texel_rgb"]), t(sc["texel_height"]), t(adj), t(off), None,
                         False)[0].clone()
    p = Pass(sc, adj, off, cam, capacity=1 << 20)
    p.run()
    torch.cuda.synchronize()
    assert int(p.status.item()) == 0
    eager = p.color.clone()
    assert torch.equal(eager, ref)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        p.run()  # warm-up on the capture stream
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        p.run()
    for k in range(3):
        p.color.fill_(-1.0)  # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(p.color, eager), k
        assert int(p.status.item()) == 0


def test_overflow_is_flagged_and_bounded(setup):
    sc, adj, off, cam = setup
    full = Pass(sc, adj, off, cam, capacity=1 << 20)
    full.run()
    torch.cuda.synchronize()
    need = int(full.offsets[-1].item())
    p = Pass(sc, adj, off, cam, capacity=need // 2)
    p.run()
    torch.cuda.synchronize()
    assert int(p.status.item()) & _lib.PF_STATUS_LIST_CAP
    assert int(p.offsets[-1].item()) == need // 2
    assert torch.isfinite(p.color).all()


def test_rasterizer_list_capacity_option(setup):
    sc, adj, off, cam = setup
    args = types.SimpleNamespace(num_texel_sites=8, raster_list_capacity=1 << 20)
    rz = Rasterizer(args, DEV)
    ps = {k: t(sc[k]).requires_grad_(True) for k in ("points", "radii", "density", "normals", "texel_sites",
                                                     "texel_rgb", "texel_height")}
    out = rz.forward(cam, None, ps["points"], ps["radii"], ps["density"], ps["normals"], ps["texel_sites"],
                     ps["texel_rgb"], ps["texel_height"], t(adj), t(off), None, False)
    out[0].sum().backward()
    assert int(rz.list_status.item()) == 0
    ref = Rasterizer(ARGS, DEV).forward(cam, None, *(t(sc[k]) for k in ("points", "radii", "density", "normals",
                                                                          "texel_sites", "texel_rgb",
                                                                          "texel_height")), t(adj), t(off), None,
                                        False)
    assert torch.equal(out[0].detach(), ref[0])
    assert torch.isfinite(ps["points"].grad).all()


def _rel(a, b):
    return float(torch.linalg.norm((a - b).flatten()) / max(float(torch.linalg.norm(b.flatten())), 1e-30))


def test_forward_and_backward_captured_in_one_graph(setup):
    sc, adj, off, cam = setup
    target = torch.from_numpy(np.random.default_rng(4).uniform(0, 1, (cam.height, cam.width, 3))
                              .astype(np.float32)).to(DEV)
    p = Pass(sc, adj, off, cam, capacity=1 << 20)
    p.backward_buffers(target)
    p.run()
    p.run_backward()
    torch.cuda.synchronize()
    assert int(p.status.item()) == 0
    eager_color = p.color.clone()
    eager = {k: v.clone() for k, v in p.grads.items()}
    # the eager pass against autograd through the drop-in rasterizer
    ps = {k: t(sc[k]).requires_grad_(True) for k in ("points", "radii", "density", "normals", "texel_sites",
                                                     "texel_rgb", "texel_height")}
    out = Rasterizer(ARGS, DEV).forward(cam, None, ps["points"], ps["radii"], ps["density"], ps["normals"],
                                        ps["texel_sites"], ps["texel_rgb"], ps["texel_height"], t(adj), t(off),
                                        None, False)
    ((out[0] - target) ** 2).sum().backward()
    assert torch.equal(out[0].detach(), eager_color)
    fin = lambda x: torch.where(torch.isfinite(x), x, torch.zeros_like(x))
    want = dict(points=eager["spheres"][:, :3], radii=eager["spheres"][:, 3], density=eager["nsigmas"][:, 3],
                normals=eager["nsigmas"][:, :3], texel_sites=eager["texel_sites"], texel_rgb=eager["texel_rgb"],
                texel_height=eager["texel_height"])
    for k, v in want.items():
        assert float(ps[k].grad.abs().max()) > 0, k
        assert _rel(fin(v), ps[k].grad) <= 1e-5, k
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        p.run()  # warm-up on the capture stream
        p.run_backward()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        p.run()
        p.run_backward()
    for k in range(3):
        p.color.fill_(-1.0)  # eager work between replays
        for v in p.grads.values():
            v.fill_(7.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(p.color, eager_color), k
        assert int(p.status.item()) == 0
        for name, v in p.grads.items():
            assert _rel(v, eager[name]) <= 1e-5, (k, name)
