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

    assert ctypes.sizeof(_lib.PfCamera) == 48
    assert ctypes.sizeof(_lib.PfScene) == 8 * 8 + 2 * 8 + 4 * 4 + 8
    assert ctypes.sizeof(_lib.PfGrads) == 5 * 8 + 8


def test_cpu_tensors_rejected_like_the_reference():
    import torch
    from powerfoam import _lib

    with pytest.raises(ValueError):
        _lib.ptr(torch.zeros(3))
