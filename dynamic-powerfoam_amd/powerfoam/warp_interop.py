"""``launch_kernel_from_torch`` (powerfoam/warp_interop.py:5-24) for the C-ABI
library: the reference's one launcher that takes torch tensors, rejects
host tensors with ValueError and enqueues on torch's current stream.

Here ``kernel`` names an entry point of ``libpf_hip.so`` (or is the ctypes
function itself); tensors are passed as device pointers, Python ints and
floats as the C scalars the entry point declares, and the stream argument
(the last parameter of every entry point that takes one) is torch's
current stream.  ``dim`` and ``block_dim`` are accepted for signature
compatibility: each entry point sizes its own grid for gfx950.  The
reference makes tensors contiguous but discards the result (:13); here a
non-contiguous tensor is an error, since its pointer would be wrong.
"""
import ctypes

import torch

from . import _lib


def launch_kernel_from_torch(device, kernel, dim, inputs, block_dim=256):
    del dim, block_dim
    lib = _lib.load()
    name = kernel if isinstance(kernel, str) else getattr(kernel, "__name__", "kernel")
    if name not in _lib._SIGS:
        raise ValueError(f"{name} is not an entry point of libpf_hip.so")
    fn = getattr(lib, name)
    argtypes = _lib._SIGS[name][1]
    args = []
    for x in inputs:
        if isinstance(x, torch.Tensor):
            if x.device.type != "cuda":
                raise ValueError("Only CUDA tensors are supported as inputs")
            args.append(_lib.ptr(x))
        else:
            args.append(x)
    if len(args) == len(argtypes) - 1 and argtypes[-1] is ctypes.c_void_p:
        args.append(_lib.stream_ptr())
    if len(args) != len(argtypes):
        raise ValueError(f"{name} takes {len(argtypes)} arguments, got {len(args)}")
    with torch.cuda.device(torch.device(device)):
        rc = fn(*args)
    _lib.check(rc, name)
    return rc
