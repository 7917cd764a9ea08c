"""Drop-in for the reference's src/voxel_kernels.py on an MI355X.

Same names and argument meaning (src/voxel_kernels.py:28-110):

  PROTOTYPE_SOURCE, make_source(real)   self-contained HIP source of the parity product: 'double'
                                        returns PROTOTYPE_SOURCE (the FP64 instantiation),
                                        'float' the FP32 one (other values: ValueError, as
                                        src/voxel_kernels.py:34-43)
  Modal, quadrature, XYZ                element algebra (host setup; Modal uses the C-ABI's
                                        vh_modal_blocks and raises ValueError outside the
                                        parity-block model, code/hex_modal/modal.py:64-66)
  element_nodes(shape)                  global node of every (element, binary lane)
  emulate(path, shape, ke, E, u, dtype) runs the GPU kernel of `path` ("modal8": the parity
                                        product, "dense8": the dense-ke node product) at dtype
                                        and returns y on the host, where the reference ran a
                                        NumPy replica of the CUDA lane logic
  assembled_reference(shape, ke, E, u)  global K u in FP64 through the dense-ke kernel, an
                                        implementation independent of the parity algebra

Inputs may be numpy arrays or torch tensors; numpy in -> numpy out.
"""
from __future__ import annotations

import re
from pathlib import Path

import numpy as np
import torch

import voxelhip
from voxelhip import ops

_CSRC = Path(__file__).resolve().parent / "csrc"


def _translation_unit(name: str) -> str:
    """The text of csrc/<name> with its project includes inlined once each (system headers
    stay as #include lines): one self-contained source that hipcc compiles on its own."""
    seen = set()

    def expand(fname):
        out = []
        for line in (_CSRC / fname).read_text().splitlines(keepends=True):
            m = re.match(r'\s*#include "([^"]+)"', line)
            if m and (_CSRC / m.group(1)).exists():
                if m.group(1) not in seen:
                    seen.add(m.group(1))
                    out.append(expand(m.group(1)))
            elif not line.startswith("#pragma once"):
                out.append(line)
        return "".join(out)

    return expand(name)


# the FP64 instantiation of the parity product (kernel template, epilogues and launcher)
PROTOTYPE_SOURCE = _translation_unit("matvec_f64.hip")
_FLOAT_SOURCE = None

XYZ = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])


def make_source(real: str) -> str:
    """HIP source for `real` in {'double', 'float'}; 'double' returns PROTOTYPE_SOURCE unchanged,
    'float' the FP32 instantiation of the same template (FP32 arithmetic, FP32 or BF16 storage)."""
    global _FLOAT_SOURCE
    if real == "double":
        return PROTOTYPE_SOURCE
    if real != "float":
        raise ValueError(real)
    if _FLOAT_SOURCE is None:
        _FLOAT_SOURCE = _translation_unit("matvec_f32.hip")
    return _FLOAT_SOURCE


def quadrature(lengths, nu):
    """Hex8 stiffness of an axis-aligned brick, unit Young's modulus, 2x2x2 Gauss points,
    engineering-strain Voigt order, nodes in the native order XYZ (code/hex_modal/modal.py:19-40)."""
    h = np.asarray(lengths, dtype=np.float64)
    lam = nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
    mu = 1.0 / (2.0 * (1.0 + nu))
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C[[0, 1, 2], [0, 1, 2]] += 2.0 * mu
    C[[3, 4, 5], [3, 4, 5]] = mu
    sgn = 2.0 * XYZ - 1.0
    # strain rows (xx, yy, zz, xy, yz, zx) -> displacement-gradient pairs (row i, axis a)
    voigt = [((0, 0),), ((1, 1),), ((2, 2),), ((0, 1), (1, 0)), ((1, 2), (2, 1)), ((0, 2), (2, 0))]
    ke = np.zeros((24, 24))
    for xi in sgn / np.sqrt(3.0):
        factors = 1.0 + sgn * xi[None, :]                  # (8 nodes, 3 axes)
        grad = np.empty((8, 3))
        for a in range(3):
            grad[:, a] = sgn[:, a] / (4.0 * h[a]) * np.prod(np.delete(factors, a, axis=1), axis=1)
        B = np.zeros((6, 8, 3))
        for row, pairs in enumerate(voigt):
            for comp, axis in pairs:
                B[row, :, comp] = grad[:, axis]
        B = B.reshape(6, 24)
        ke += B.T @ C @ B * (np.prod(h) / 8.0)
    return ke


class Modal:
    """Parity-block factorisation (code/hex_modal/modal.py:55-77): `blocks` (8, 3, 3) and
    `discarded_relative`; ValueError when the element is outside the model."""

    def __init__(self, ke, tolerance=2e-14):
        self.ke = np.asarray(ke, dtype=np.float64).reshape(24, 24)
        self.blocks, self.discarded_relative = ops.modal_blocks(self.ke, tolerance)


def element_nodes(shape):
    """(n_el, 8) global node of every (element, binary lane), as the kernels index them."""
    nx, ny, nz = (int(s) for s in shape)
    ex, ey, ez = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    lane = np.arange(8)
    return (((ex.reshape(-1, 1) + (lane & 1)) * (ny + 1) + ey.reshape(-1, 1) + ((lane >> 1) & 1))
            * (nz + 1) + ez.reshape(-1, 1) + (lane >> 2))


_TORCH = {np.float64: torch.float64, np.float32: torch.float32}


def _tdtype(dtype):
    return _TORCH[np.dtype(dtype).type]


def emulate(path, shape, ke, E, u, dtype=np.float64):
    """Runs the product kernel of `path` on the GPU at `dtype`; returns y (host if inputs are)."""
    host = not isinstance(u, torch.Tensor)
    td = _tdtype(dtype)
    Ed = voxelhip.to_device(E, td)
    ud = voxelhip.to_device(u, td)
    y = torch.empty_like(ud)
    if path == "modal8":
        ops.matvec(Modal(ke).blocks, Ed, ud, y, shape)
    elif path == "dense8":
        ops.dense_product(ke, Ed, ud, y, shape)
    else:
        raise ValueError(path)
    return y.cpu().numpy() if host else y


def assembled_reference(shape, ke, E, u):
    """Global K u in FP64 with the dense element matrix (node-owned gather kernel)."""
    host = not isinstance(u, torch.Tensor)
    Ed = voxelhip.to_device(E, torch.float64)
    ud = voxelhip.to_device(u, torch.float64)
    y = torch.empty_like(ud)
    ops.dense_product(ke, Ed, ud, y, shape)
    return y.cpu().numpy() if host else y
