"""Host side of the walk record: slot count, allocation, the chunked
segment index and the slot -> pixel map (no GPU)."""
import numpy as np
import pytest
import torch

from powerfoam import raytrace as rt


def test_slots_cover_whole_tiles():
    assert rt.record_slots(1, 1920, 1080) == 240 * 135 * 64
    assert rt.record_slots(3, 37, 29) == 3 * 5 * 4 * 64


def test_cap_rounds_up_to_the_chunk():
    r = rt.new_record(1, 16, 16, 13, "cpu")
    assert r["cap"] == 16 and r["segments"].shape == (16 * 4 * 64, 4)
    with pytest.raises(ValueError):
        rt.new_record(1, 16, 16, 0, "cpu")


def test_segment_index_is_a_bijection_with_chunks_of_eight():
    r = rt.new_record(2, 24, 16, 24, "cpu")
    S = r["counts"].numel()
    rows = np.array([[rt.record_segment(r, k, s) for s in range(S)] for k in range(24)])
    assert sorted(rows.ravel().tolist()) == list(range(24 * S))
    # eight consecutive segments of a slot are one 128-byte line
    assert np.array_equal(rows[8:16, 5], rows[8, 5] + np.arange(8)) and rows[8, 5] % 8 == 0


def test_counts_image_follows_the_tile_lane_map():
    B, W, H = 2, 20, 13
    r = rt.new_record(B, W, H, 8, "cpu")
    S = r["counts"].numel()
    r["counts"].copy_(torch.arange(S, dtype=torch.int32))
    img = rt.record_counts(r, B, W, H).numpy()
    th, tw = (H + 7) // 8, (W + 7) // 8
    for b in range(B):
        for i in range(H):
            for j in range(W):
                tile = b * th * tw + (j // 8) * th + i // 8
                lane = (i % 8) + 8 * (j % 8)
                assert img[b, i, j] == tile * 64 + lane


def test_record_cap_fits_budget():
    import types

    tracer = rt.RayTracer(types.SimpleNamespace(walk_record_cap=64, walk_record_budget=2 << 30), "cpu")
    cap = tracer.record_cap_for(1, 1920, 1080)
    assert 8 <= cap < 64 and cap % 8 == 0 and rt.record_bytes(1, 1920, 1080, cap) <= 2 << 30
    assert tracer.record_cap_for(64, 1920, 1080) == 0
    assert rt.RayTracer(types.SimpleNamespace(walk_record_cap=0), "cpu").record_cap_for(1, 64, 64) == 0
    # the default: 128 segments per ray, 8.5 GB of record and 17 GB of
    # grouped workspace for one 1080p frame
    assert rt.RayTracer(None, "cpu").record_cap_for(1, 1920, 1080) == rt.DEFAULT_RECORD_CAP == 128


def test_walk_launch_buffers_are_checked_before_any_launch():
    """Wrong shapes, dtypes or layouts of the camera words, checkpoint,
    incoming gradients or gradient buffers raise ValueError on the host
    (CPU tensors: the check comes before any pointer reaches the library)."""
    import types

    import pytest
    import torch

    from powerfoam import raytrace as rt

    scene = types.SimpleNamespace(num_cells=10, K=8)
    B, W, H = 2, 16, 8
    cams = torch.zeros(B, 12, dtype=torch.int32)
    color = torch.zeros(B, H, W, 3)
    log_t = torch.zeros(B, H, W)
    bwd = lambda **kw: rt.launch_walk_backward(scene, kw.pop("cams", cams), W, H, kw.pop("color", color), log_t,
                                               kw.pop("gc", color), None, **kw)
    with pytest.raises(ValueError, match="cams_dev"):
        bwd(cams=torch.zeros(B, 12))
    with pytest.raises(ValueError, match="color"):
        bwd(color=torch.zeros(B, W, H, 3))
    with pytest.raises(ValueError, match="grad_color"):
        bwd(gc=torch.zeros(B, H, W, 3).transpose(1, 2).contiguous().transpose(1, 2))
    good = dict(spheres=torch.zeros(10, 4), nsigmas=torch.zeros(10, 4), texel_sites=torch.zeros(10, 8, 3),
                texel_rgb=torch.zeros(10, 8, 3), texel_height=torch.zeros(10, 8))
    for k in good:
        bad = dict(good)
        bad[k] = torch.zeros(9, *good[k].shape[1:])
        with pytest.raises(ValueError, match=k):
            bwd(grads=bad)
    with pytest.raises(ValueError, match="stride"):
        bwd(grads=good, texel_rgb_cam_stride=240)  # two cameras' rows need 480 floats
    with pytest.raises(ValueError, match="cams_dev"):
        rt.launch_walk_forward(scene, torch.zeros(B, 12), W, H)
    # per-camera texel colours must match the launch's cameras, and their
    # gradients must be per camera too
    per_cam = types.SimpleNamespace(num_cells=10, K=8, rgb_cams=3)
    with pytest.raises(ValueError, match="3 cameras"):
        rt.launch_walk_forward(per_cam, cams, W, H)
    with pytest.raises(ValueError, match="3 cameras"):
        rt.launch_walk_backward(per_cam, cams, W, H, color, log_t, color, None)
    per_cam2 = types.SimpleNamespace(num_cells=10, K=8, rgb_cams=2)
    with pytest.raises(ValueError, match="per-camera"):
        rt.launch_walk_backward(per_cam2, cams, W, H, color, log_t, color, None, grads=good)


def test_sh_prepass_buffers_are_checked_before_any_launch():
    import pytest
    import torch

    from powerfoam import shading

    coeffs, sites = torch.zeros(10, 16, 3), torch.zeros(10, 8, 3)
    cams = torch.zeros(2, 12, dtype=torch.int32)
    with pytest.raises(ValueError, match="coeffs"):
        shading.sh_forward(torch.zeros(10, 9, 3), 3, sites, cams)
    with pytest.raises(ValueError, match="texel_sites"):
        shading.sh_forward(coeffs, 3, torch.zeros(10, 8, 2), cams)
    with pytest.raises(ValueError, match="cams_dev"):
        shading.sh_forward(coeffs, 3, sites, torch.zeros(2, 12))
    with pytest.raises(ValueError, match="out"):
        shading.sh_forward(coeffs, 3, sites, cams, out=torch.zeros(1, 10, 8, 3))
    with pytest.raises(ValueError, match="grad_rgb"):
        shading.sh_backward(coeffs, 3, sites, cams, torch.zeros(1, 10, 8, 3))
    with pytest.raises(ValueError, match="grad_coeffs"):
        shading.sh_backward(coeffs, 3, sites, cams, torch.zeros(2, 10, 8, 3), grad_coeffs=torch.zeros(10, 9, 3))
