"""Tensor-level wrappers over the C-ABI: argument checks, stream plumbing, dtype dispatch.

Every function takes torch tensors on the current HIP device and launches asynchronously on
torch's current stream.  There is no CPU path: a tensor that is not on the GPU is an error.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib as L

_DT = {torch.float64: L.F64, torch.float32: L.F32, torch.bfloat16: L.BF16}


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev(t, name):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch tensor on the GPU, got {type(t).__name__}")
    if not t.is_cuda:
        raise ValueError(f"{name}: tensor is on {t.device}; voxelhip runs on the GPU only")
    if not t.is_contiguous():
        raise ValueError(f"{name}: tensor must be contiguous")
    return t


def dtype_code(t):
    try:
        return _DT[t.dtype]
    except KeyError:
        raise TypeError(f"unsupported dtype {t.dtype}") from None


def n_nodes(shape):
    nx, ny, nz = (int(s) for s in shape)
    return (nx + 1) * (ny + 1) * (nz + 1)


def n_elements(shape):
    nx, ny, nz = (int(s) for s in shape)
    return nx * ny * nz


# --------------------------------------------------------------------------- host setup
def modal_blocks(ke, tolerance=2e-14):
    """(blocks[8,3,3], discarded_relative) of C = T ke T^T / 64; ValueError outside the model."""
    ke = np.ascontiguousarray(np.asarray(ke, dtype=np.float64).reshape(24, 24))
    blocks = np.zeros(72)
    disc = ctypes.c_double(0.0)
    lib = L.load()
    st = lib.vh_modal_blocks(ke.ctypes.data_as(ctypes.c_void_p), float(tolerance),
                             blocks.ctypes.data_as(ctypes.c_void_p), ctypes.byref(disc))
    if st == -3:
        raise ValueError(lib.vh_last_error().decode())
    L.check(st, "vh_modal_blocks")
    return blocks.reshape(8, 3, 3), float(disc.value)


class Mask:
    """Per-DOF constraint mask on the device: kind NONE, X0 (every DOF of node plane x = 0 fixed;
    no array) or ARRAY (uint8, 1 = free)."""

    def __init__(self, kind, array=None):
        self.kind, self.array = kind, array

    @staticmethod
    def none():
        return Mask(L.MASK_NONE)

    @staticmethod
    def from_values(mask, shape, device=None):
        """Classify a 0/1 mask (host or device); the x = 0 clamp needs no array."""
        m = mask.detach().cpu().numpy() if isinstance(mask, torch.Tensor) else np.asarray(mask)
        m = (m.reshape(-1) != 0)
        nx, ny, nz = (int(s) for s in shape)
        if m.size != 3 * n_nodes(shape):
            raise ValueError(f"mask has {m.size} entries, grid {shape} has {3 * n_nodes(shape)} DOFs")
        g = m.reshape(nx + 1, (ny + 1) * (nz + 1) * 3)
        if m.all():
            return Mask(L.MASK_NONE)
        if (not g[0].any()) and g[1:].all():
            return Mask(L.MASK_X0)
        dev = device or torch.device("cuda", torch.cuda.current_device())
        return Mask(L.MASK_ARRAY, torch.from_numpy(m.astype(np.uint8)).to(dev))

    def values(self, shape, dtype=torch.float64, device=None):
        """The mask as a 0/1 float tensor (for callers that want the reference's float masks)."""
        dev = device or torch.device("cuda", torch.cuda.current_device())
        if self.kind == L.MASK_ARRAY:
            return self.array.to(dtype)
        v = torch.ones(3 * n_nodes(shape), dtype=dtype, device=dev)
        if self.kind == L.MASK_X0:
            nx, ny, nz = (int(s) for s in shape)
            v[: 3 * (ny + 1) * (nz + 1)] = 0
        return v

    def coarsen(self, shape):
        """Injection onto the 2x coarser grid (src/voxel_mg.py:114-116)."""
        if self.kind != L.MASK_ARRAY:
            return Mask(self.kind)
        nx, ny, nz = (int(s) for s in shape)
        a = self.array.view(nx + 1, ny + 1, nz + 1, 3)[::2, ::2, ::2, :].contiguous().view(-1)
        return Mask(L.MASK_ARRAY, a)

    @property
    def ptr(self):
        return _p(self.array) if self.kind == L.MASK_ARRAY else None


# --------------------------------------------------------------------------- operator
def matvec(blocks, E, u, y, shape, mask=None, mode=L.MODE_RAW, b=None, dinv=None, d=None,
           partial=None, omega=0.0, cheb_a=0.0, cheb_c=0.0, fixed_zero=False, lx=0, storage=None, tile=0,
           kdiag=None, prev_zero=False, w=None, sdev=None, sidx=0, u_out=None, y32=None, load_scale=1.0,
           prev=None, zold=None):
    """One pass of the parity product with a fused epilogue (see include/voxelhip.h).  An FP32
    u with FP64 moduli selects the FP32-input DOT pass of the mixed-precision Krylov loop; an
    FP32 w with FP64 moduli, the RESID pass that adds an FP32 search direction to the FP64 x."""
    _dev(E, "E"); _dev(u, "u")
    if y is None and not (mode == L.MODE_DOT or (mode == L.MODE_RESID and y32 is not None)):
        raise ValueError("only a DOT pass, or a RESID pass writing y32, may omit y")
    if y is not None:
        _dev(y, "y")
    nx, ny, nz = (int(s) for s in shape)
    dt = dtype_code(E)
    st = dt if storage is None else storage
    if storage is None and u.dtype == torch.float32 and E.dtype == torch.float64:
        st = L.F32                      # FP32 input of an FP64 pass
    if y is not None and E.dtype != y.dtype:
        raise TypeError("E and y must share the compute dtype")
    if st == L.BF16:
        if u.dtype not in (torch.bfloat16, torch.int16, torch.uint16):
            raise TypeError("bf16 storage expects a bfloat16 u")
    elif u.dtype != E.dtype and not (st == L.F32 and dt == L.F64):
        raise TypeError("u and y must share dtype")
    nd = 3 * n_nodes(shape)
    if u.numel() != nd or (y is not None and y.numel() != nd) or E.numel() != nx * ny * nz:
        raise ValueError(f"sizes do not match grid {shape}")
    m = mask or Mask.none()
    hb = np.ascontiguousarray(np.asarray(blocks, dtype=np.float64).reshape(72))
    kd = (ctypes.c_double * 3)(*([0.0] * 3 if kdiag is None else [float(v) for v in kdiag]))
    if mode in (L.MODE_JACOBI, L.MODE_CHEB, L.MODE_CHEB3, L.MODE_CHEB3D) and dinv is None and kdiag is None:
        raise ValueError("smoothing passes need dinv or kdiag")
    flags = L.PREV_ZERO if prev_zero else 0
    if w is not None and w.dtype == torch.float32 and E.dtype == torch.float64:
        flags |= L.W_F32                # FP32 search direction read by the FP64 RESID pass
    ep = L.Epilogue(mode, int(bool(fixed_zero)), _p(b), _p(dinv), _p(d), _p(partial),
                    float(omega), float(cheb_a), float(cheb_c), int(lx), int(tile), kd, flags,
                    _p(w), _p(sdev), int(sidx), _p(u_out), _p(y32), float(load_scale), _p(prev), _p(zold))
    L.call("vh_matvec_fused", dt, st, hb.ctypes.data_as(ctypes.c_void_p), _p(E), _p(u), _p(y),
           nx, ny, nz, m.kind, m.ptr, ctypes.byref(ep), _stream())
    return y


def matvec_tiles(shape, lx=0, tile=0, dtype=torch.float64, blocks=None, storage=None):
    """(n_tiles, n_workgroups) of the product launch: n_tiles is the DOT partial count."""
    nt, nw = ctypes.c_int(0), ctypes.c_int(0)
    dt = _DT[dtype]
    hb = None if blocks is None else np.ascontiguousarray(np.asarray(blocks, dtype=np.float64).reshape(72))
    L.call("vh_matvec_tiles", dt, dt if storage is None else storage,
           None if hb is None else hb.ctypes.data_as(ctypes.c_void_p), *[int(s) for s in shape], int(lx),
           int(tile), ctypes.byref(nt), ctypes.byref(nw))
    return nt.value, nw.value


def dense_product(ke, E, u, y, shape):
    _dev(E, "E"); _dev(u, "u"); _dev(y, "y")
    k = np.ascontiguousarray(np.asarray(ke, dtype=np.float64).reshape(576))
    L.call("vh_dense_product", dtype_code(y), k.ctypes.data_as(ctypes.c_void_p), _p(E), _p(u), _p(y),
           *[int(s) for s in shape], _stream())
    return y


def lanes24_product(blocks, E, u, y, shape):
    """y = K(E) u with the element-per-24-lanes layout control (see include/voxelhip.h)."""
    _dev(E, "E"); _dev(u, "u"); _dev(y, "y")
    hb = np.ascontiguousarray(np.asarray(blocks, dtype=np.float64).reshape(72))
    if u.dtype != E.dtype or y.dtype != E.dtype or u.numel() != y.numel() or u.numel() != 3 * n_nodes(shape):
        raise ValueError("lanes24_product: dtypes or sizes do not match")
    L.call("vh_lanes24_product", dtype_code(y), hb.ctypes.data_as(ctypes.c_void_p), _p(E), _p(u), _p(y),
           *[int(s) for s in shape], _stream())
    return y


def diagonal(ke, E, out, shape, mask=None, invert=False):
    _dev(E, "E"); _dev(out, "out")
    k = np.ascontiguousarray(np.asarray(ke, dtype=np.float64).reshape(576))
    m = mask or Mask.none()
    L.call("vh_diagonal", dtype_code(out), k.ctypes.data_as(ctypes.c_void_p), _p(E), _p(out),
           *[int(s) for s in shape], m.kind, m.ptr, int(bool(invert)), _stream())
    return out


def element_energy(blocks, u, ce, shape, mask=None, rho=None, dc=None, penal=3.0, dE=1.0):
    _dev(u, "u"); _dev(ce, "ce")
    hb = np.ascontiguousarray(np.asarray(blocks, dtype=np.float64).reshape(72))
    m = mask or Mask.none()
    L.call("vh_element_energy", dtype_code(ce), hb.ctypes.data_as(ctypes.c_void_p), _p(u), _p(ce),
           *[int(s) for s in shape], m.kind, m.ptr, _p(rho), _p(dc), float(penal), float(dE), _stream())
    return ce


def restrict(rf, rc, fine_shape, fmask=None, cmask=None):
    _dev(rf, "rf"); _dev(rc, "rc")
    fm, cm = fmask or Mask.none(), cmask or Mask.none()
    L.call("vh_restrict", dtype_code(rc), _p(rf), _p(rc), *[int(s) for s in fine_shape],
           fm.kind, fm.ptr, cm.kind, cm.ptr, _stream())
    return rc


def prolong(ec, xf, fine_shape, mask=None, assign=False):
    _dev(ec, "ec"); _dev(xf, "xf")
    m = mask or Mask.none()
    L.call("vh_prolong", dtype_code(xf), _p(ec), _p(xf), *[int(s) for s in fine_shape], m.kind, m.ptr,
           int(bool(assign)), _stream())
    return xf


def coarsen_modulus(Ef, Ec, fine_shape):
    _dev(Ef, "Ef"); _dev(Ec, "Ec")
    L.call("vh_coarsen_modulus", dtype_code(Ec), _p(Ef), _p(Ec), *[int(s) for s in fine_shape], _stream())
    return Ec


def dense_mv(A, r, y, mask=None, n=None):
    """y = mask * (A[:n, :n] r): A is a 2-D device tensor whose row stride is its leading
    dimension (a multiple of 4; see padded_matrix), FP64 or FP32; r, y FP32 or of A's type."""
    _dev(A, "A"); _dev(r, "r"); _dev(y, "y")
    n = int(r.numel()) if n is None else int(n)
    if A.dim() != 2 or A.stride(1) != 1 or A.shape[0] < n or A.shape[1] < n:
        raise ValueError("dense_mv: A must be a row-major 2-D tensor of at least n x n")
    m = mask.array if (mask is not None and mask.kind == L.MASK_ARRAY) else None
    if mask is not None and mask.kind == L.MASK_X0:
        raise ValueError("dense_mv takes an explicit mask array")
    L.call("vh_dense_mv", dtype_code(A), dtype_code(y), n, int(A.stride(0)), _p(A), _p(r), _p(y), _p(m), _stream())
    return y


def assemble_dense(ke, E, shape, mask=None, A=None):
    """Dense FP64 constrained operator M K M + (I - M) of a small grid, assembled on the device
    (ke: 24 x 24 FP64 device tensor in the native node order; E: FP64 element moduli)."""
    _dev(ke, "ke"); _dev(E, "E")
    n = 3 * n_nodes(shape)
    if ke.dtype != torch.float64 or E.dtype != torch.float64 or ke.numel() != 576:
        raise TypeError("assemble_dense: ke (576 values) and E must be float64")
    if A is None:
        A = torch.empty((n, n), dtype=torch.float64, device=E.device)
    _dev(A, "A")
    if A.dim() != 2 or A.shape[0] != n or A.shape[1] < n or A.dtype != torch.float64:
        raise ValueError("assemble_dense: A must be a float64 n x (>= n) matrix")
    m = mask or Mask.none()
    L.call("vh_assemble_dense", _p(ke), _p(E), *[int(s) for s in shape], m.kind, m.ptr, _p(A), int(A.shape[1]),
           _stream())
    return A


def galerkin_scatter(G, A, coarse_shape, mask=None):
    """Write M_c (P^T A_f P) M_c + (I - M_c) into the zeroed matrix A from the 81 probe rows G
    (see include/voxelhip.h, vh_galerkin_scatter)."""
    _dev(G, "G"); _dev(A, "A")
    n = 3 * n_nodes(coarse_shape)
    if G.shape != (81, n) or G.dtype != torch.float64 or A.dtype != torch.float64 or A.shape[0] != n \
            or A.shape[1] < n:
        raise ValueError("galerkin_scatter: G must be 81 x n and A n x (>= n), float64")
    m = mask or Mask.none()
    L.call("vh_galerkin_scatter", _p(G), *[int(s) for s in coarse_shape], m.kind, m.ptr, _p(A), int(A.shape[1]),
           _stream())
    return A


def stencil_scatter(G, U, coarse_shape):
    """Symmetric 27-point stencil U (14 x 9 x nn, float64 or float32) of the unmasked Galerkin
    operator P^T A_f P from the 81 probe rows G (see include/voxelhip.h, vh_stencil_scatter)."""
    _dev(G, "G"); _dev(U, "U")
    nn = n_nodes(coarse_shape)
    if G.shape != (81, 3 * nn) or G.dtype != torch.float64 or U.numel() != 126 * nn or not U.is_contiguous():
        raise ValueError("stencil_scatter: G must be 81 x 3nn float64 and U contiguous with 126 nn entries")
    L.call("vh_stencil_scatter", dtype_code(U), _p(G), _p(U), *[int(s) for s in coarse_shape], _stream())
    return U


def stencil_apply(U, u, y, shape, mask=None, mode=L.MODE_RAW, b=None, dinv=None, omega=0.0, cheb_a=0.0,
                  cheb_c=0.0, fixed_zero=False, prev=None, prev_zero=False, y32=None):
    """One pass of the stencil operator M K M + I - M with the RAW, RESID, JACOBI or CHEB3
    epilogue (see include/voxelhip.h, vh_stencil_apply)."""
    _dev(U, "U"); _dev(u, "u")
    if y is None and not (mode == L.MODE_RESID and y32 is not None):
        raise ValueError("only a RESID pass writing y32 may omit y")
    nd = 3 * n_nodes(shape)
    if u.numel() != nd or (y is not None and (y.numel() != nd or y.dtype != u.dtype)) or U.dtype != u.dtype \
            or U.numel() != 42 * nd:
        raise ValueError(f"stencil_apply: sizes or dtypes do not match grid {shape}")
    m = mask or Mask.none()
    ep = L.Epilogue(mode, int(bool(fixed_zero)), _p(b), _p(dinv), None, None, float(omega), float(cheb_a),
                    float(cheb_c), 0, 0, (ctypes.c_double * 3)(0.0, 0.0, 0.0), int(bool(prev_zero)),
                    None, None, 0, None, _p(y32), 1.0, _p(prev), None)
    L.call("vh_stencil_apply", dtype_code(u), _p(U), _p(u), _p(y), *[int(s) for s in shape], m.kind, m.ptr,
           ctypes.byref(ep), _stream())
    return y


def padded_matrix(A, dtype=None):
    """Copy of the square matrix A with its rows padded to a multiple of 4 entries (16-byte
    aligned rows for the vector loads of dense_mv); the padding is zero."""
    n = A.shape[0]
    ld = (n + 3) // 4 * 4
    out = torch.zeros((n, ld), dtype=dtype or A.dtype, device=A.device)
    out[:, :n] = A
    return out


# --------------------------------------------------------------------------- BLAS-1
def reduce_parts(n):
    return int(L.load().vh_reduce_parts(int(n)))


def reduce(kind, x, y=None, w=None, partial=None):
    _dev(x, "x")
    L.call("vh_reduce", dtype_code(x), kind, int(x.numel()), _p(x), _p(y), _p(w), _p(partial), _stream())
    return partial


def finish(partial, nparts, count, scal, dst, op=L.OP_NONE):
    L.call("vh_finish", _p(partial), int(nparts), int(count), _p(scal), int(dst), int(op), _stream())


def cg_update(x, p, scal, r=None, q=None, dinv=None, z=None, partial=None):
    L.call("vh_cg_update", dtype_code(x), int(x.numel()), _p(x), _p(p), _p(r), _p(q), _p(dinv), _p(z),
           _p(scal), int(r is not None), _p(partial), _stream())


def jacobi_dot(r, dinv, z, partial):
    L.call("vh_jacobi_dot", dtype_code(r), int(r.numel()), _p(r), _p(dinv), _p(z), _p(partial), _stream())


def cvt_dotdiff(zin, r, zold, znew, partial):
    _dev(zin, "zin"); _dev(r, "r"); _dev(zold, "zold"); _dev(znew, "znew")
    L.call("vh_cvt_dotdiff", dtype_code(zin), dtype_code(znew), int(r.numel()), _p(zin), _p(r), _p(zold), _p(znew),
           _p(partial), _stream())


def xpby(z, p, scal, zcopy=None):
    L.call("vh_xpby", dtype_code(p), int(p.numel()), _p(z), _p(p), _p(scal), _p(zcopy), _stream())


def axpby(a, x, b, y, w=None):
    _dev(x, "x"); _dev(y, "y")
    L.call("vh_axpby", dtype_code(y), int(y.numel()), float(a), _p(x), _p(w), float(b), _p(y), _stream())
    return y


def fill_splitmix(y, seed):
    """y[i] = 2 u_i - 1, u_i from the (i+1)-th SplitMix64 output of `seed` (vh_fill_splitmix)."""
    L.call("vh_fill_splitmix", _DT[y.dtype], y.numel(), int(seed) & (2 ** 64 - 1), y.data_ptr(), _stream())
    return y


def convert(x, y, scale=1.0, dscale=None, inv_scale=False, accumulate=False):
    _dev(x, "x"); _dev(y, "y")
    L.call("vh_convert", dtype_code(x), dtype_code(y), int(x.numel()), _p(x), _p(y), float(scale),
           _p(dscale), int(bool(inv_scale)), int(bool(accumulate)), _stream())
    return y


# --------------------------------------------------------------------------- SIMP
def cone_filter(x, y, shape, rmin, pre_div=None, post_div=None):
    _dev(x, "x"); _dev(y, "y")
    L.call("vh_cone_filter", dtype_code(y), _p(x), _p(pre_div), _p(post_div), _p(y),
           *[int(s) for s in shape], float(rmin), _stream())
    return y


def heaviside(rho_bar, beta, eta=0.5, penal=3.0, E0=1.0, Emin=1e-9, rho=None, drho=None, E=None, partial=None):
    """Projected density, its derivative, the SIMP modulus and per-block sums of the projected
    density (FP64; see include/voxelhip.h, vh_heaviside).  Returns the number of partial blocks."""
    _dev(rho_bar, "rho_bar")
    for t in (rho_bar, rho, drho, E):
        if t is not None and (t.dtype != torch.float64 or t.numel() != rho_bar.numel()):
            raise TypeError("heaviside: float64 vectors of one length")
    nparts = int(partial.numel()) if partial is not None else 0
    L.call("vh_heaviside", int(rho_bar.numel()), _p(rho_bar), float(beta), float(eta), float(penal), float(E0),
           float(Emin), _p(rho), _p(drho), _p(E), _p(partial), nparts, _stream())
    return nparts


def oc_design(rho, dc, dv, lmid, move, rho_lb, rho_new):
    for t in (rho, dc, dv, rho_new):
        _dev(t, "oc_design")
        if t.dtype != torch.float64 or t.numel() != rho.numel():
            raise TypeError("oc_design: float64 vectors of one length")
    L.call("vh_oc_design", int(rho.numel()), _p(rho), _p(dc), _p(dv), float(lmid), float(move), float(rho_lb),
           _p(rho_new), _stream())
    return rho_new


def oc_probe(x, dc, dv, w, lmid, move, xnew, partial, nparts):
    L.call("vh_oc_probe", dtype_code(x), int(x.numel()), _p(x), _p(dc), _p(dv), _p(w), float(lmid),
           float(move), _p(xnew), _p(partial), int(nparts), _stream())


OC_MAX_PROBES = 16


def oc_probe_multi(x, dc, dv, w, lmids, move, partial, nparts):
    """Volume partials of len(lmids) <= OC_MAX_PROBES bisection probes in one pass:
    partial[k * nparts + b] as oc_probe's partial[b] for lmids[k] (no xnew)."""
    lm = (ctypes.c_double * len(lmids))(*[float(v) for v in lmids])
    if partial.numel() < len(lmids) * nparts:
        raise ValueError("partial holds fewer than len(lmids) * nparts values")
    L.call("vh_oc_probe_multi", dtype_code(x), int(x.numel()), _p(x), _p(dc), _p(dv), _p(w), lm, len(lmids),
           float(move), _p(partial), int(nparts), _stream())
