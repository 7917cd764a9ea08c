"""Jacobi-PCG and mixed-precision iterative refinement on the device.

Same signatures, stopping rules and return values as the donor solvers the reference harness
calls (experiments/session_a/donor/src/gpu_fem/simp_r2.py:187-385):

  pcg(matvec, b, M_inv, x0=None, tol=1e-5, maxiter=20000, mask=None, check_every=50,
      residual_history=None) -> (x, iters, converged, rel)
  pcg_ir(matvec64, b64, matvec32, M_inv32, tol=1e-5, mask32=None, inner_tol=1e-1, max_outer=40,
         inner_maxiter=4000, check_every=50, trace=None, x0=None)
      -> (x64, outer, inner_total, converged, rel)

Differences in mechanism, not in the algorithm: alpha, beta and r.z live in FP64 device scalars
(dots accumulate in FP64 also for FP32 vectors), so an iteration never synchronises with the
host; the true residual is tested, and the recursion replaced, every `check_every` iterations
exactly as the donor does.  The donor's per-iteration `float(pAp) == 0` breakdown exit becomes a
device flag that zeroes alpha and is read at the next check.  When `matvec` is the bound
`apply` of a voxel_mg.Level built on a HipProduct, A p, the mask and p.Ap come from one fused
kernel pass.
"""
from __future__ import annotations

import torch

from . import _lib as L
from . import ops


def _level_of(matvec):
    lvl = getattr(matvec, "__self__", None)
    if lvl is not None and getattr(lvl, "fused", False) and matvec.__func__.__name__ == "apply":
        return lvl
    return None


class _Ctx:
    def __init__(self, n, dev, tiles=0):
        self.parts = ops.reduce_parts(n)
        self.partial = torch.zeros(2 * max(self.parts, tiles, 1), dtype=torch.float64, device=dev)
        self.scal = torch.zeros(16, dtype=torch.float64, device=dev)

    def norm(self, v, slot=8):
        ops.reduce(L.RED_NORM2, v, partial=self.partial)
        ops.finish(self.partial, self.parts, 1, self.scal, slot)
        return float(self.scal[slot].sqrt())


def pcg(matvec, b, M_inv, x0=None, tol=1e-5, maxiter=20000, mask=None, check_every=50,
        residual_history=None):
    """Jacobi-PCG whose stopping test is the true residual ||b - A x|| / ||b||."""
    dev, n, dt = b.device, b.numel(), b.dtype
    lvl = _level_of(matvec)
    tiles = lvl.tiles() if lvl is not None else 0
    cx = _Ctx(n, dev, tiles)
    x = torch.zeros_like(b) if x0 is None else x0.to(dt).clone()
    if mask is not None:
        ops.axpby(1.0, x, 0.0, x, w=mask)
    r = torch.empty_like(b)
    q = torch.empty_like(b)
    z = torch.empty_like(b)

    def residual(into):
        if lvl is not None:
            lvl.matvec(x, into, mode=L.MODE_RESID, b=b, fixed_zero=mask is not None)
        else:
            into.copy_(b - matvec(x))
            if mask is not None:
                ops.axpby(1.0, into, 0.0, into, w=mask)

    residual(r)
    ops.jacobi_dot(r, M_inv, z, cx.partial)
    ops.finish(cx.partial, cx.parts, 1, cx.scal, L.S_RZ)
    p = z.clone()
    b_norm = cx.norm(b, 9)
    if b_norm <= 0:
        return x, 0, True, 0.0
    tol_abs = tol * b_norm
    r_norm = cx.norm(r)
    if residual_history is not None:
        residual_history.append(r_norm / b_norm)
    if r_norm <= tol_abs:
        return x, 0, True, r_norm / b_norm
    k = max(1, int(check_every))
    converged = False
    iters = 0
    for _ in range(maxiter):
        if lvl is not None:
            lvl.matvec(p, q, mode=L.MODE_DOT, partial=cx.partial, fixed_zero=mask is not None)
            ops.finish(cx.partial, tiles, 1, cx.scal, L.S_PQ, L.OP_ALPHA)
        else:
            q.copy_(matvec(p))
            if mask is not None:
                ops.axpby(1.0, q, 0.0, q, w=mask)
            ops.reduce(L.RED_DOT, p, q, partial=cx.partial)
            ops.finish(cx.partial, cx.parts, 1, cx.scal, L.S_PQ, L.OP_ALPHA)
        iters += 1
        if iters % k == 0 or iters >= maxiter:
            ops.cg_update(x, p, cx.scal)                        # x += alpha p
            if float(cx.scal[L.S_FLAG]) != 0.0:                  # pAp == 0 seen: stop
                break
            residual(r)
            r_norm = cx.norm(r)
            if residual_history is not None:
                residual_history.append(r_norm / b_norm)
            if r_norm <= tol_abs:
                converged = True
                break
            ops.jacobi_dot(r, M_inv, z, cx.partial)
        else:
            ops.cg_update(x, p, cx.scal, r=r, q=q, dinv=M_inv, z=z, partial=cx.partial)
        ops.finish(cx.partial, cx.parts, 1, cx.scal, 5, L.OP_BETA_FR)
        ops.xpby(z, p, cx.scal)
    if not converged:
        residual(r)
        r_norm = cx.norm(r)
    return x, iters, converged, r_norm / b_norm


def pcg_ir(matvec64, b64, matvec32, M_inv32, tol=1e-5, mask32=None, inner_tol=1e-1, max_outer=40,
           inner_maxiter=4000, check_every=50, trace=None, x0=None):
    """FP64 residual, FP32 correction, FP64 update, with the donor's divergence guard."""
    dev, n = b64.device, b64.numel()
    cx = _Ctx(n, dev)
    nb = cx.norm(b64, 9)
    if nb <= 0:
        return torch.zeros_like(b64), 0, 0, True, 0.0
    x = torch.zeros_like(b64) if x0 is None else x0.to(torch.float64).clone()
    lvl = _level_of(matvec64)
    r = torch.empty_like(b64)
    r32 = torch.empty(n, dtype=torch.float32, device=dev)

    def residual():
        if lvl is not None:
            lvl.matvec(x, r, mode=L.MODE_RESID, b=b64)
        else:
            r.copy_(b64 - matvec64(x))

    inner_total, outer = 0, 0
    tol_in, rel_prev, d_last, scale_last = inner_tol, None, None, 0.0
    for outer in range(1, max_outer + 1):
        residual()
        rnorm = cx.norm(r)
        rel = rnorm / nb
        if trace is not None:
            trace.append(rel)
        if rel <= tol:
            return x, outer - 1, inner_total, True, rel
        if rel_prev is not None and rel > rel_prev and d_last is not None:
            ops.convert(d_last, x, scale=-scale_last, accumulate=True)     # undo the step
            d_last = None
            tol_in *= 0.1
            if tol_in < 1e-5:
                return x, outer, inner_total, False, rel_prev
            continue
        rel_prev = rel
        scale = rnorm
        ops.convert(r, r32, scale=1.0 / scale)
        d32, it32, _ok, _res = pcg(matvec32, r32, M_inv32, tol=tol_in, maxiter=inner_maxiter,
                                   mask=mask32, check_every=check_every)
        inner_total += it32
        d_last, scale_last = d32, scale
        ops.convert(d32, x, scale=scale, accumulate=True)                  # x += d32 * scale
    residual()
    rel = cx.norm(r) / nb
    if trace is not None:
        trace.append(rel)
    return x, outer, inner_total, bool(rel <= tol), rel
