"""ctypes binding of the gfx950 C-ABI library ``libpf_hip.so``.

The library has a plain C interface (``include/pf_hip.h``): device pointers,
sizes and a ``hipStream_t``.  This module loads it, mirrors its structs and
maps status codes onto the reference's exception types
(ValueError for bad inputs, as ``powerfoam/warp_interop.py:14-15`` and
``powerfoam/rasterize.py:1620-1637`` raise; RuntimeError for device errors).
There is no CPU fallback: if the library is missing the import fails.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpf_hip.so")

PF_OK, PF_ERR_ARG, PF_ERR_DTYPE, PF_ERR_HIP, PF_ERR_OVERFLOW = 0, 1, 2, 3, 4
PF_STATUS_STEP_CAP, PF_STATUS_LIST_CAP, PF_STATUS_CYCLE, PF_STATUS_BAD_RECORD = 1, 2, 4, 8
PF_F32, PF_F16, PF_F16_PACKED = 0, 1, 2
PF_BWD_EXACT_TEXEL = 1
PF_SERVE_SH_IN_WALK = 1

c_void_p = ctypes.c_void_p
c_int32 = ctypes.c_int32
c_int64 = ctypes.c_int64
c_float = ctypes.c_float


class PfCamera(ctypes.Structure):
    _fields_ = [
        ("eye", c_float * 3),
        ("right", c_float * 3),
        ("up", c_float * 3),
        ("start_cell", c_int32),
        ("width", c_int32),
        ("height", c_int32),
    ]


class PfScene(ctypes.Structure):
    _fields_ = [
        ("spheres", c_void_p),
        ("nsigmas", c_void_p),
        ("texel_sites", c_void_p),
        ("texel_rgb", c_void_p),
        ("texel_height", c_void_p),
        ("adjacency", c_void_p),
        ("adjacency_offsets", c_void_p),
        ("faces", c_void_p),
        ("num_cells", c_int64),
        ("num_edges", c_int64),
        ("num_texel_sites", c_int32),
        ("attr_dtype", c_int32),
        ("face_dtype", c_int32),
        ("reserved", c_int32),
        ("texel_rgb_cam_stride", c_int64),
    ]


class PfWalkRecord(ctypes.Structure):
    _fields_ = [
        ("segments", c_void_p),
        ("counts", c_void_p),
        ("cap", c_int32),
        ("reserved", c_int32),
        ("partials", c_void_p),
        ("header", c_void_p),
    ]


class PfWalkOut(ctypes.Structure):
    _fields_ = [
        ("color", c_void_p),
        ("log_t", c_void_p),
        ("depth", c_void_p),
        ("visits", c_void_p),
        ("seq_hash", c_void_p),
        ("seq", c_void_p),
        ("seq_cap", c_int32),
        ("reserved", c_int32),
        ("status", c_void_p),
        ("alg_bytes", c_void_p),
        ("record", ctypes.POINTER(PfWalkRecord)),
        ("ray_status", c_void_p),
    ]


class PfGrads(ctypes.Structure):
    _fields_ = [
        ("spheres", c_void_p),
        ("nsigmas", c_void_p),
        ("texel_sites", c_void_p),
        ("texel_rgb", c_void_p),
        ("texel_height", c_void_p),
        ("texel_rgb_cam_stride", c_int64),
    ]


class PfRasterOut(ctypes.Structure):
    _fields_ = [(k, c_void_p) for k in ("color", "log_t", "normal_distance", "normal", "quantile_depths", "err",
                                         "contrib", "point_err")]


class PfRasterGradIn(ctypes.Structure):
    _fields_ = [(k, c_void_p) for k in ("color", "log_t", "normal_distance", "normal", "quantile_depths",
                                         "contrib")]


class PfVisOptions(ctypes.Structure):
    _fields_ = [("transmittance_threshold", c_float), ("max_intersections", c_int32), ("depth_quantile", c_float),
                ("bkgd_color", c_float * 3)]


class PfVisOut(ctypes.Structure):
    _fields_ = [(k, c_void_p) for k in ("color", "depth", "normal", "alpha", "intersections")]


class PfCechGrid(ctypes.Structure):
    _fields_ = [("origin", c_float * 3), ("cell", c_float), ("dims", c_int32 * 3), ("levels", c_int32)]


class PfServeScene(ctypes.Structure):
    _fields_ = [("scene", PfScene), ("sh", c_void_p), ("sh_degree", c_int32), ("reserved", c_int32)]


class PfServeRequest(ctypes.Structure):
    _fields_ = [("cam", PfCamera), ("scene", c_int32), ("mode", c_int32), ("tile_begin", c_int32),
                ("tile_count", c_int32), ("step_tile0", c_int64), ("start_cell", c_void_p), ("color", c_void_p),
                ("log_t", c_void_p), ("target", c_void_p), ("grads", PfGrads), ("texel_rgb", c_void_p),
                ("bucket0", c_int64)]


# name -> (restype, argtypes); every symbol declared in include/pf_hip.h
_SIGS = {
    "pf_device_count": (c_int32, [ctypes.POINTER(c_int32)]),
    "pf_status_string": (ctypes.c_char_p, [c_int32]),
    "pf_walk_forward": (
        c_int32,
        [ctypes.POINTER(PfScene), c_void_p, c_int32, c_int32, c_int32, c_float, c_float, c_int32,
         ctypes.POINTER(PfWalkOut), c_void_p],
    ),
    "pf_walk_backward": (
        c_int32,
        [ctypes.POINTER(PfScene), c_void_p, c_int32, c_int32, c_int32, c_float, c_int32, c_void_p, c_void_p,
         c_void_p, c_void_p, ctypes.POINTER(PfGrads), c_int32, c_void_p, c_void_p, c_void_p],
    ),
    "pf_walk_record_slots": (c_int64, [c_int32, c_int32, c_int32]),
    "pf_walk_backward_recorded": (
        c_int32,
        [ctypes.POINTER(PfScene), c_void_p, c_int32, c_int32, c_int32, c_float, c_int32, c_void_p, c_void_p,
         c_void_p, c_void_p, ctypes.POINTER(PfWalkRecord), ctypes.POINTER(PfGrads), c_int32, c_void_p, c_void_p,
         c_void_p],
    ),
    "pf_walk_grouped_workspace_bytes": (c_int64, [c_int64, c_int32, c_int32, c_int32, c_int32]),
    "pf_walk_backward_grouped": (
        c_int32,
        [ctypes.POINTER(PfScene), c_void_p, c_int32, c_int32, c_int32, c_float, c_int32, c_void_p, c_void_p,
         c_void_p, c_void_p, ctypes.POINTER(PfWalkRecord), ctypes.POINTER(PfGrads), c_int32, c_void_p, c_int64,
         c_void_p, c_void_p, c_void_p],
    ),
    "pf_walk_backward_grouped_noalloc": (
        c_int32,
        [ctypes.POINTER(PfScene), c_void_p, c_int32, c_int32, c_int32, c_float, c_int32, c_void_p, c_void_p,
         c_void_p, c_void_p, ctypes.POINTER(PfWalkRecord), ctypes.POINTER(PfGrads), c_int32, c_void_p, c_void_p,
         c_void_p],
    ),
    "pf_walk_grouped_scratch_bytes": (c_int64, []),
    "pf_walk_backward_replay_grouped": (
        c_int32,
        [ctypes.POINTER(PfScene), c_void_p, c_int32, c_int32, c_int32, c_float, c_int32, c_void_p, c_void_p,
         c_void_p, c_void_p, ctypes.POINTER(PfGrads), c_int32, c_int32, c_int32, c_void_p, c_int64, c_void_p,
         c_void_p],
    ),
    "pf_walk_replay_grouped_workspace_bytes": (c_int64, [c_int64, c_int32, c_int32, c_int32, c_int32, c_int32]),
    "pf_walk_faces": (
        c_int32,
        [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_int64, c_void_p, c_int32, c_void_p],
    ),
    "pf_start_cells": (c_int32, [c_void_p, c_void_p, c_int64, c_void_p, c_int32, c_void_p, c_void_p]),
    "pf_walk_pack_faces": (c_int32, [c_void_p, c_void_p, c_int64, c_void_p, c_void_p]),
    "pf_raster_faces": (c_int32, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_int32, c_void_p]),
    "pf_raster_bin_count": (c_int32, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p]),
    "pf_raster_scan": (c_int32, [c_void_p, c_int64, c_void_p, c_void_p]),
    "pf_raster_scan_bounded": (c_int32, [c_void_p, c_int64, c_void_p, c_int64, c_void_p, c_void_p]),
    "pf_raster_bin_write": (
        c_int32,
        [c_void_p, c_int64, c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64,
         c_void_p],
    ),
    "pf_raster_forward": (
        c_int32,
        [ctypes.POINTER(PfScene), c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_float, c_void_p,
         c_int32, c_void_p, c_int32, c_int32, ctypes.POINTER(PfRasterOut), c_void_p],
    ),
    "pf_raster_backward": (
        c_int32,
        [ctypes.POINTER(PfScene), c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_int32,
         c_void_p, ctypes.POINTER(PfRasterGradIn), ctypes.POINTER(PfGrads), c_void_p],
    ),
    "pf_raster_benchmark": (
        c_int32,
        [ctypes.POINTER(PfScene), c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_float, c_void_p, c_void_p],
    ),
    "pf_raster_visualize": (
        c_int32,
        [ctypes.POINTER(PfScene), c_void_p, c_int32, c_int32, c_void_p, c_void_p, ctypes.POINTER(PfVisOptions),
         ctypes.POINTER(PfVisOut), c_void_p],
    ),
    "pf_cech_bounds": (c_int32, [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p]),
    "pf_cech_count": (
        c_int32,
        [c_void_p, c_void_p, c_int64, ctypes.POINTER(PfCechGrid), c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
         c_void_p, c_void_p, c_void_p],
    ),
    "pf_cech_grid_size": (c_int64, [ctypes.POINTER(PfCechGrid)]),
    "pf_cech_write_workspace_bytes": (c_int64, [c_int64, c_int64]),
    "pf_cech_write": (
        c_int32,
        [c_void_p, c_void_p, c_int64, ctypes.POINTER(PfCechGrid), c_void_p, c_void_p, c_void_p, c_int64, c_void_p,
         c_void_p, c_void_p, c_int64, c_void_p],
    ),
    "pf_power_workspace_bytes": (c_int64, [ctypes.POINTER(PfCechGrid)]),
    "pf_morton_codes": (c_int32, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p]),
    "pf_power_cells": (
        c_int32,
        [c_void_p, c_void_p, c_int64, ctypes.POINTER(PfCechGrid), c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
         c_void_p, c_int64, c_void_p, c_float, c_float, c_int32, c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "pf_power_hull": (
        c_int32,
        [c_void_p, c_void_p, c_void_p, c_int32, c_void_p, c_float, c_int32, c_void_p, c_void_p, c_void_p, c_void_p],
    ),
    "pf_serve_record_slots": (c_int64, [c_int64]),
    "pf_serve_grouped_workspace_bytes": (c_int64, [c_int64, c_int32, c_int64]),
    "pf_serve_step_grouped": (
        c_int32,
        [c_void_p, c_void_p, c_void_p, c_int64, c_int64, c_float, c_int32, c_void_p, ctypes.POINTER(PfWalkRecord),
         c_int64, c_int32, c_void_p, c_int64, c_void_p],
    ),
    "pf_serve_step": (
        c_int32,
        [c_void_p, c_void_p, c_void_p, c_int64, c_float, c_int32, c_int32, c_void_p, ctypes.POINTER(PfWalkRecord),
         c_void_p],
    ),
    "pf_sh_forward": (
        c_int32,
        [c_void_p, c_int32, c_void_p, c_int64, c_int32, c_void_p, c_int32, c_void_p, c_int32, c_int64, c_void_p],
    ),
    "pf_sh_backward": (
        c_int32,
        [c_void_p, c_int32, c_void_p, c_int64, c_int32, c_void_p, c_int32, c_void_p, c_int64, c_void_p, c_void_p],
    ),
    "pf_sv_forward": (
        c_int32,
        [c_void_p, c_int64, c_void_p, c_void_p, c_float, c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_void_p,
         c_void_p, c_void_p],
    ),
    "pf_sv_backward": (
        c_int32,
        [c_void_p, c_int64, c_void_p, c_void_p, c_float, c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_void_p,
         c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    ),
}

_lib = None


def load():
    """Load libpf_hip.so (once).  Raises ImportError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run dynamic-powerfoam_amd/build_lib.py (hipcc, gfx950)"
        )
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    return list(_SIGS)


def check(rc, what):
    if rc == PF_OK:
        return
    msg = load().pf_status_string(rc).decode()
    if rc in (PF_ERR_ARG, PF_ERR_DTYPE, PF_ERR_OVERFLOW):
        raise ValueError(f"{what}: {msg}")
    raise RuntimeError(f"{what}: {msg}")


def ptr(t):
    """Device pointer of a tensor (None -> NULL).  CPU tensors are rejected
    as the reference rejects them (warp_interop.py:14-15)."""
    if t is None:
        return None
    if not isinstance(t, torch.Tensor):
        raise ValueError("expected a torch.Tensor")
    if t.device.type != "cuda":
        raise ValueError("Only CUDA tensors are supported as inputs")
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return t.data_ptr()


def stream_ptr(device=None):
    return torch.cuda.current_stream(device).cuda_stream


def dtype_code(dt):
    if dt == torch.float32:
        return PF_F32
    if dt == torch.float16:
        return PF_F16
    raise ValueError(f"Unsupported attribute dtype: {dt}")
