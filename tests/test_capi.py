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
    sizes = dict(PfCamera=48, PfScene=104, PfWalkOut=80, PfWalkRecord=24, PfGrads=48, PfServeScene=120,
                 PfServeRequest=160, PfCechGrid=32)
    for name, size in sizes.items():
        assert ctypes.sizeof(getattr(_lib, name)) == size, name
    src = open(os.path.join(ROOT, "dynamic-powerfoam_amd", "csrc", "capi.hip")).read()
    for cname, size in (("pf_camera", 48), ("pf_scene", 104), ("pf_walk_out", 80), ("pf_walk_record", 24),
                        ("pf_grads", 48), ("pf_serve_scene", 120), ("pf_serve_request", 160),
                        ("pf_cech_grid", 32)):
        assert f"sizeof({cname}) == {size}" in src, cname


def test_cpu_tensors_rejected_like_the_reference():
    import torch
    from powerfoam import _lib

    with pytest.raises(ValueError):
        _lib.ptr(torch.zeros(3))
