"""The C-ABI library loads and exports every symbol include/pf_hip.h declares
(no GPU needed: nothing here launches a kernel)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT


def declared_symbols():
    names = []
    for f in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if f.endswith(".h"):
            text = open(os.path.join(ROOT, "include", f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(pf_\w+)\s*\(", text, flags=re.M)
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "pf_walk_forward" in names and "pf_walk_backward" in names
    assert len(names) == len(set(names))


def test_library_exports_every_declared_symbol():
    from powerfoam import _lib

    lib = _lib.load()
    raw = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(raw, name), f"{name} declared in include/ but not exported"
    assert set(declared_symbols()) == set(_lib.exported_symbols())
    assert lib.pf_status_string(0) == b"ok"


def test_struct_layouts_match_header():
    from powerfoam import _lib

    # the same sizes are static_asserts in csrc/capi.hip, so the library
    # and this binding agree or the library does not build
    sizes = dict(PfCamera=48, PfScene=104, PfWalkOut=88, PfWalkRecord=40, PfGrads=48, PfServeScene=120,
                 PfServeRequest=168, PfCechGrid=32)
    for name, size in sizes.items():
        assert ctypes.sizeof(getattr(_lib, name)) == size, name
    src = open(os.path.join(ROOT, "dynamic-powerfoam_amd", "csrc", "capi.hip")).read()
    for cname, size in (("pf_camera", 48), ("pf_scene", 104), ("pf_walk_out", 88), ("pf_walk_record", 40),
                        ("pf_grads", 48), ("pf_serve_scene", 120), ("pf_serve_request", 168),
                        ("pf_cech_grid", 32)):
        assert f"sizeof({cname}) == {size}" in src, cname


def test_cpu_tensors_rejected_like_the_reference():
    import torch
    from powerfoam import _lib

    with pytest.raises(ValueError):
        _lib.ptr(torch.zeros(3))


def test_walk_arguments_rejected_before_any_launch():
    """Bad sizes come back as PF_ERR_ARG from the host-side checks, before
    any HIP call: the pointers here are placeholders that are never read."""
    from powerfoam import _lib

    lib = _lib.load()
    fake = 256
    scene = _lib.PfScene(spheres=fake, nsigmas=fake, texel_sites=fake, texel_rgb=fake, texel_height=fake,
                         adjacency=fake, adjacency_offsets=fake, faces=fake, num_cells=10, num_edges=40,
                         num_texel_sites=8, attr_dtype=_lib.PF_F32, face_dtype=_lib.PF_F16_PACKED)
    out = _lib.PfWalkOut(color=fake)
    fwd = lambda B, W, H, o=out, s=scene: lib.pf_walk_forward(ctypes.byref(s), fake, B, W, H, 1e-4, 0.5, 64,
                                                             ctypes.byref(o), None)
    assert fwd(1, 0, 8) == _lib.PF_ERR_ARG
    assert fwd(1, 8, -1) == _lib.PF_ERR_ARG
    assert fwd(0, 8, 8) == _lib.PF_ERR_ARG
    # fewer than 2^25 tiles per grid (128 work-items each, under 2^32): 32
    # cameras of 8192 x 8192 rays are exactly 2^25 tiles, 4096 far more
    assert fwd(32, 8192, 8192) == _lib.PF_ERR_ARG
    assert fwd(4096, 8192, 8192) == _lib.PF_ERR_ARG
    bad_dtype = _lib.PfScene.from_buffer_copy(scene)
    bad_dtype.face_dtype = 7
    assert fwd(1, 8, 8, s=bad_dtype) == _lib.PF_ERR_DTYPE
    no_cells = _lib.PfScene.from_buffer_copy(scene)
    no_cells.num_cells = 0
    assert fwd(1, 8, 8, s=no_cells) == _lib.PF_ERR_ARG
    for cap in (0, 12, -8):
        rec = _lib.PfWalkRecord(segments=fake, counts=fake, cap=cap)
        o = _lib.PfWalkOut(color=fake, record=ctypes.pointer(rec))
        assert fwd(1, 8, 8, o=o) == _lib.PF_ERR_ARG, cap
    grads = _lib.PfGrads(spheres=fake, nsigmas=fake, texel_sites=fake, texel_rgb=fake, texel_height=fake)
    rec = _lib.PfWalkRecord(segments=fake, counts=fake, cap=20)
    assert lib.pf_walk_backward_recorded(ctypes.byref(scene), fake, 1, 8, 8, 1e-4, 64, fake, fake, fake, None,
                                         ctypes.byref(rec), ctypes.byref(grads), 0, None, None,
                                         None) == _lib.PF_ERR_ARG
    for B in (32, 4096):
        assert lib.pf_walk_backward(ctypes.byref(scene), fake, B, 8192, 8192, 1e-4, 64, fake, fake, fake, None,
                                    ctypes.byref(grads), 0, None, None, None) == _lib.PF_ERR_ARG
        assert lib.pf_walk_backward_recorded(ctypes.byref(scene), fake, B, 8192, 8192, 1e-4, 64, fake, fake, fake,
                                             None, ctypes.byref(_lib.PfWalkRecord(segments=fake, counts=fake,
                                                                                  cap=16)),
                                             ctypes.byref(grads), 0, None, None, None) == _lib.PF_ERR_ARG
    assert lib.pf_walk_record_slots(2, 9, 17) == 2 * 3 * 2 * 64
    assert lib.pf_walk_record_slots(0, 9, 17) == 0
    # the grouped backward: partials required, K = 8, a large enough workspace
    ws = lib.pf_walk_grouped_workspace_bytes(10, 1, 8, 8, 16)
    assert ws >= 4 * 64 * 16 and lib.pf_walk_grouped_workspace_bytes(10, 1, 8, 8, 0) == 0
    # entry ids are int32: 2^31 record entries and more are refused
    assert lib.pf_walk_grouped_workspace_bytes(10, 1, 8192, 8192, 2048) == 0
    grouped = lambda r, w=fake, nb=ws, s=scene: lib.pf_walk_backward_grouped(
        ctypes.byref(s), fake, 1, 8, 8, 1e-4, 64, fake, fake, fake, None, ctypes.byref(r), ctypes.byref(grads), 0,
        w, nb, None, None, None)
    no_part = _lib.PfWalkRecord(segments=fake, counts=fake, cap=16)
    assert grouped(no_part) == _lib.PF_ERR_ARG
    rec16 = _lib.PfWalkRecord(segments=fake, counts=fake, cap=16, partials=fake)
    assert grouped(rec16, nb=ws - 1) == _lib.PF_ERR_OVERFLOW
    assert grouped(rec16, w=None) == _lib.PF_ERR_OVERFLOW
    k4 = _lib.PfScene.from_buffer_copy(scene)
    k4.num_texel_sites = 4
    assert grouped(rec16, s=k4) == _lib.PF_ERR_ARG
    assert grouped(_lib.PfWalkRecord(segments=fake, counts=fake, cap=12, partials=fake)) == _lib.PF_ERR_ARG


def test_raster_and_serving_sizes_rejected_before_any_launch():
    from powerfoam import _lib

    lib = _lib.load()
    fake = 256
    scene = _lib.PfScene(spheres=fake, nsigmas=fake, texel_sites=fake, texel_rgb=fake, texel_height=fake,
                         adjacency=fake, adjacency_offsets=fake, faces=fake, num_cells=10, num_edges=40,
                         num_texel_sites=8, attr_dtype=_lib.PF_F32, face_dtype=_lib.PF_F16)
    out = _lib.PfRasterOut(color=fake, log_t=fake, normal_distance=fake, normal=fake, contrib=fake)
    raster = lambda W, H: lib.pf_raster_forward(ctypes.byref(scene), fake, W, H, fake, fake, fake, 1e-4, None, 0,
                                                None, 0, 0, ctypes.byref(out), None)
    # fewer than 2^24 tiles of 8x8 per raster image: 65536 x 16384 is exactly 2^24
    assert raster(65536, 16384) == _lib.PF_ERR_ARG
    assert raster(65536, 16385) == _lib.PF_ERR_ARG
    assert raster(0, 8) == _lib.PF_ERR_ARG
    for tiles in (1 << 24, (1 << 24) + 1):
        assert lib.pf_raster_bin_write(fake, 10, fake, tiles, fake, fake, fake, fake, fake, None, 10,
                                       None) == _lib.PF_ERR_ARG
    step = lambda cap, flags=0, rec=None: lib.pf_serve_step(fake, fake, fake, cap, 1e-4, 64, flags, None, rec, None)
    assert step(0) == _lib.PF_ERR_ARG
    assert step(1 << 26) == _lib.PF_ERR_ARG  # 64 work-items per tile: fewer than 2^26 tiles
    assert step((1 << 26) + 1) == _lib.PF_ERR_ARG
    assert step(64, flags=0x10) == _lib.PF_ERR_ARG
    rec = _lib.PfWalkRecord(segments=fake, counts=fake, cap=100)
    assert step(64, rec=ctypes.pointer(rec)) == _lib.PF_ERR_ARG
    assert lib.pf_serve_record_slots(5) == 8 * 64
