"""The no-allocation entry point pf_walk_backward_grouped_noalloc: the grouped
walk backward with the library's static scratch array as its workspace.

Bar: through the C ABI, its gradients equal pf_walk_backward_grouped's with
a caller workspace (same segments, same kernels: within 1e-5 relative L2,
the order of float additions aside) and the restatement's (1e-3), on two
calls in a row, the second with a smaller input (the array keeps the first
call's contents, which the kernels must clear); sizes past the array are
refused with PF_ERR_OVERFLOW before any device work (no GPU needed)."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from powerfoam import _lib
from powerfoam import raytrace as rt
from powerfoam import workloads

KEYS = ("spheres", "nsigmas", "texel_sites", "texel_rgb", "texel_height")


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_noalloc_refuses_sizes_past_its_scratch():
    lib = _lib.load()
    fake = 256
    scene = _lib.PfScene(spheres=fake, nsigmas=fake, texel_sites=fake, texel_rgb=fake, texel_height=fake,
                         adjacency=fake, adjacency_offsets=fake, faces=fake, num_cells=1_000_000,
                         num_edges=15_000_000, num_texel_sites=8, attr_dtype=_lib.PF_F32,
                         face_dtype=_lib.PF_F16_PACKED)
    grads = _lib.PfGrads(spheres=fake, nsigmas=fake, texel_sites=fake, texel_rgb=fake, texel_height=fake)
    rec = _lib.PfWalkRecord(segments=fake, counts=fake, cap=128, partials=fake)
    limit = lib.pf_walk_grouped_scratch_bytes()
    assert limit == 64 << 20
    # a 1920x1080 record at cap 128 needs over 1 GB of workspace
    assert lib.pf_walk_grouped_workspace_bytes(1_000_000, 1, 1920, 1080, 128) > limit
    rc = lib.pf_walk_backward_grouped_noalloc(ctypes.byref(scene), fake, 1, 1920, 1080, 1e-3, 64, fake, fake, fake,
                                              None, ctypes.byref(rec), ctypes.byref(grads), 0, None, None, None)
    assert rc == _lib.PF_ERR_OVERFLOW


def _grads(n, K, dev):
    return dict(spheres=torch.zeros(n, 4, device=dev), nsigmas=torch.zeros(n, 4, device=dev),
                texel_sites=torch.zeros(n, K, 3, device=dev), texel_rgb=torch.zeros(n, K, 3, device=dev),
                texel_height=torch.zeros(n, K, device=dev))


def _pfgrads(g):
    out = _lib.PfGrads()
    for k in KEYS:
        setattr(out, k, _lib.ptr(g[k]))
    return out


@pytest.mark.gpu
def test_noalloc_equals_grouped_twice_second_smaller(scene_textured):
    from test_walk_gpu import _cams, _cpu_scene, _gpu_scene

    dev = "cuda"
    lib = _lib.load()
    sc = scene_textured
    gs, faces = _gpu_scene(sc)
    cs = _cpu_scene(sc, faces)
    for W, H, cap in ((64, 56, 64), (24, 16, 16)):  # the second call is smaller
        cams = [workloads.front_camera(W, H, hfov_deg=40.0)]
        words = _cams(sc, cams)
        cams_dev = torch.from_numpy(words).to(dev)
        fwd = rt.launch_walk_forward(gs, cams_dev, W, H, record_cap=cap)
        rng = np.random.default_rng(W)
        target = torch.from_numpy(rng.uniform(0, 1, (1, H, W, 3)).astype(np.float32)).to(dev)
        gc = (2.0 * (fwd["color"] - target)).contiguous()
        gl = (-0.3 * torch.exp(fwd["log_t"])).contiguous()
        rec = rt._record_struct(fwd["record"])
        assert lib.pf_walk_grouped_workspace_bytes(gs.num_cells, 1, W, H, fwd["record"]["cap"]) <= \
            lib.pf_walk_grouped_scratch_bytes()
        # the entry point with a caller workspace
        ref = rt.launch_walk_backward(gs, cams_dev, W, H, fwd["color"], fwd["log_t"], gc, gl, record=fwd["record"],
                                      grouped=True)
        # the no-allocation one, straight through the C ABI
        g = _grads(gs.num_cells, gs.K, dev)
        status = torch.zeros(1, dtype=torch.int32, device=dev)
        torch.cuda.synchronize()
        rc = lib.pf_walk_backward_grouped_noalloc(ctypes.byref(gs.c), _lib.ptr(cams_dev), 1, W, H, 1e-3, 1 << 16,
                                                  _lib.ptr(fwd["color"]), _lib.ptr(fwd["log_t"]), _lib.ptr(gc),
                                                  _lib.ptr(gl), ctypes.byref(rec), ctypes.byref(_pfgrads(g)), 0,
                                                  _lib.ptr(status), None, _lib.stream_ptr())
        _lib.check(rc, "pf_walk_backward_grouped_noalloc")
        torch.cuda.synchronize()
        assert int(status.item()) == 0
        c = orc.walk_forward(cs, words)
        cg = orc.walk_backward(cs, words, c["color"], gc.cpu().numpy(), gl.cpu().numpy())
        for k in KEYS:
            a = g[k].cpu().numpy()
            assert float(np.abs(a).sum()) > 0, k
            assert _rel(a, ref[k].cpu().numpy()) < 1e-5, (W, k)
            assert _rel(a, cg[k]) < 1e-3, (W, k)
