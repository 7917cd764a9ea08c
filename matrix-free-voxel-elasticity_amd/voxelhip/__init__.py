"""voxelhip: matrix-free Hex8 voxel elasticity on one AMD Instinct MI355X (gfx950).

Device arrays are torch tensors on the HIP device; every numerical operation on them runs in
the hand-written HIP kernels of libvoxelhip.so through its C-ABI (include/voxelhip.h).

`voxelhip.xp` is the array namespace handed to the drop-in voxel_mg functions in place of the
reference's numpy / cupy module (src/voxel_mg.py passes `xp` everywhere).
"""
from __future__ import annotations

import types

import numpy as np
import torch

from . import _lib, ops  # noqa: F401

__all__ = ["xp", "device", "to_device", "ops"]


def device():
    if not torch.cuda.is_available():
        raise RuntimeError("voxelhip needs a HIP device (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def to_device(a, dtype=None):
    """numpy / torch -> contiguous device tensor (float64 unless `dtype`)."""
    if isinstance(a, torch.Tensor):
        t = a
    else:
        t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    elif t.dtype not in (torch.float64, torch.float32, torch.uint8, torch.bfloat16):
        t = t.to(torch.float64)
    return t.to(device()).contiguous()


def _norm(v):
    partial = torch.empty(2 * ops.reduce_parts(v.numel()), dtype=torch.float64, device=v.device)
    scal = torch.zeros(8, dtype=torch.float64, device=v.device)
    ops.reduce(_lib.RED_NORM2, v, partial=partial)
    ops.finish(partial, ops.reduce_parts(v.numel()), 1, scal, 7)
    return torch.sqrt(scal[7])


xp = types.SimpleNamespace(
    asarray=to_device,
    zeros_like=torch.zeros_like,
    ascontiguousarray=lambda t: t.contiguous(),
    linalg=types.SimpleNamespace(norm=_norm),
    float64=torch.float64,
    float32=torch.float32,
    __name__="voxelhip.xp",
)
