"""CPU oracle for the Jacobi-PCG and mixed-precision iterative refinement solvers.

TEST INFRASTRUCTURE ONLY (tests/, __graft_entry__.smoke(), bench.py cpu_baseline).

Restates, in NumPy, the donor solvers the reference harness calls:
  pcg     experiments/session_a/donor/src/gpu_fem/simp_r2.py:187-295
          Jacobi-PCG, Fletcher-Reeves beta, true residual + residual replacement every
          `check_every` iterations and at maxiter, breakdown exit on pAp == 0.
  pcg_ir  simp_r2.py:298-385
          FP64 residual, correction solved by `pcg` in FP32 on the residual scaled to unit
          norm, FP64 update, divergence guard (undo + tenfold tighter inner tol, give up
          below 1e-5).

Parity status: reference-held known answers of results/session-a-20260923 (smoke-q4: 100
PCG iterations at check_every 50, native true residual ~7e-14, compliance 22009.41647...,
3 IR outer iterations with 150 inner), and the stored smoke-q4 solution vectors.
"""
from __future__ import annotations

import numpy as np


def pcg(matvec, b, M_inv, x0=None, tol=1e-5, maxiter=20000, mask=None, check_every=50,
        residual_history=None):
    """Returns (x, iters, converged, rel) with rel = ||b - A x|| / ||b|| recomputed."""
    dt = b.dtype
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=dt)
    if mask is not None:
        x = x * mask
    r = b - matvec(x)
    if mask is not None:
        r = r * mask
    z = M_inv * r
    p = z.copy()
    rz = dt.type(np.dot(r, z))
    b_norm = float(np.linalg.norm(b))
    if b_norm <= 0:
        return x, 0, True, 0.0
    tol_abs = tol * b_norm
    r_norm = float(np.linalg.norm(r))
    if residual_history is not None:
        residual_history.append(r_norm / b_norm)
    if r_norm <= tol_abs:
        return x, 0, True, r_norm / b_norm
    k = max(1, int(check_every))
    converged = False
    iters = 0
    for _ in range(maxiter):
        Ap = matvec(p)
        if mask is not None:
            Ap = Ap * mask
        pAp = dt.type(np.dot(p, Ap))
        if float(pAp) == 0.0:
            break
        alpha = rz / pAp
        x = x + alpha * p
        iters += 1
        if iters % k == 0 or iters >= maxiter:
            r = b - matvec(x)
            if mask is not None:
                r = r * mask
            r_norm = float(np.linalg.norm(r))
            if residual_history is not None:
                residual_history.append(r_norm / b_norm)
            if r_norm <= tol_abs:
                converged = True
                break
        else:
            r = r - alpha * Ap
        z = M_inv * r
        rz_new = dt.type(np.dot(r, z))
        beta = rz_new / rz
        p = z + beta * p
        rz = rz_new
    if not converged:
        r = b - matvec(x)
        if mask is not None:
            r = r * mask
        r_norm = float(np.linalg.norm(r))
    return x, iters, converged, r_norm / b_norm


def pcg_ir(matvec64, b64, matvec32, M_inv32, tol=1e-5, mask32=None, inner_tol=1e-1,
           max_outer=40, inner_maxiter=4000, check_every=50, trace=None, x0=None):
    """Returns (x64, outer, inner_total, converged, rel)."""
    nb = float(np.linalg.norm(b64))
    if nb <= 0:
        return np.zeros_like(b64), 0, 0, True, 0.0
    x = np.zeros_like(b64) if x0 is None else np.array(x0, dtype=np.float64)
    inner_total = 0
    outer = 0
    tol_in = inner_tol
    rel_prev = None
    d_last = None
    scale_last = 0.0
    for outer in range(1, max_outer + 1):
        r = b64 - matvec64(x)
        rel = float(np.linalg.norm(r)) / nb
        if trace is not None:
            trace.append(rel)
        if rel <= tol:
            return x, outer - 1, inner_total, True, rel
        if rel_prev is not None and rel > rel_prev and d_last is not None:
            x = x - d_last.astype(np.float64) * scale_last
            d_last = None
            tol_in *= 0.1
            if tol_in < 1e-5:
                return x, outer, inner_total, False, rel_prev
            continue
        rel_prev = rel
        scale = float(np.linalg.norm(r))
        d32, it32, _, _ = pcg(matvec32, (r / scale).astype(np.float32), M_inv32, tol=tol_in,
                              maxiter=inner_maxiter, mask=mask32, check_every=check_every)
        inner_total += it32
        d_last, scale_last = d32, scale
        x = x + d32.astype(np.float64) * scale
    r = b64 - matvec64(x)
    rel = float(np.linalg.norm(r)) / nb
    if trace is not None:
        trace.append(rel)
    return x, outer, inner_total, bool(rel <= tol), rel
