"""SIMP pieces kept device-resident: cone density filter, top88 OC update, compliance with its
hand-written adjoint behind torch.autograd, and the 100-step optimisation loop.

  ConeFilter(shape, rmin)            forward(rho) = H rho / s, transpose(g) = H (g / s)
                                     (experiments/session_a/donor/src/gpu_fem/filter_r2.py:93-183)
  oc_top88(x, dc, dv, volfrac, move, filt=None, l2=1e9, tol=1e-3)
                                     bisection of experiments/session_h/run_w1b.py:46-62,82-94;
                                     the filtered volume mean(H x_new / s) is evaluated as
                                     sum_j x_new_j w_j with w = H(1/s)/n (H symmetric), one fused
                                     kernel per probe
  Compliance.apply(rho, problem)     c = f^T u(rho); backward dc/drho = -p (E0 - Emin) rho^(p-1)
                                     u_e^T ke u_e (donor_yang_gmg solver_v4.py:246-251) from the
                                     element-energy kernel
  simp_run(problem, steps, ...)      density-filtered SIMP with the top88 OC
                                     (experiments/session_h/run_w1b.py:163-222, update "top88_gpu")
  SimpProblem(shape, ke, mask, force, ...)
                                     the state problem of simp_run on the GMG-FCG engine: every
                                     step rebuilds the hierarchy for the new moduli and solves
                                     K(E) u = f to the true relative residual `tol`
"""
from __future__ import annotations

import time

import torch

from . import _lib as L
from . import ops


class ConeFilter:
    def __init__(self, shape, rmin, dtype=torch.float64, device=None):
        self.shape, self.rmin = tuple(int(s) for s in shape), float(rmin)
        self.n = ops.n_elements(self.shape)
        dev = device or torch.device("cuda", torch.cuda.current_device())
        ones = torch.ones(self.n, dtype=dtype, device=dev)
        self.s = torch.empty_like(ones)
        ops.cone_filter(ones, self.s, self.shape, self.rmin)
        # weights of the filtered-volume identity: mean(H x / s) = sum_j x_j (H (1/s))_j / n
        self.w_vol = torch.empty_like(ones)
        ops.cone_filter(ones, self.w_vol, self.shape, self.rmin, pre_div=self.s)
        ops.axpby(1.0 / self.n, self.w_vol, 0.0, self.w_vol)

    def forward(self, rho, out=None):
        y = out if out is not None else torch.empty_like(rho)
        return ops.cone_filter(rho, y, self.shape, self.rmin, post_div=self.s)

    density = forward

    def transpose(self, g, out=None):
        y = out if out is not None else torch.empty_like(g)
        return ops.cone_filter(g, y, self.shape, self.rmin, pre_div=self.s)

    gradient = transpose

    def check_adjoint(self, seed=0):
        gen = torch.Generator(device=self.s.device).manual_seed(seed)
        a = torch.rand(self.n, generator=gen, device=self.s.device, dtype=self.s.dtype)
        b = torch.rand(self.n, generator=gen, device=self.s.device, dtype=self.s.dtype)
        lhs = float(self.forward(a) @ b)
        rhs = float(a @ self.transpose(b))
        return abs(lhs - rhs) / max(abs(lhs), abs(rhs), 1e-300)


def oc_top88(x, dc, dv, volfrac, move, filt=None, l2=1e9, tol=1e-3):
    """Returns x_new of the last bisection probe, as top88 does."""
    n = x.numel()
    nparts = min(1024, max(1, (n + 255) // 256))
    partial = torch.zeros(nparts, dtype=torch.float64, device=x.device)
    w = filt.w_vol if filt is not None else None
    xnew = torch.empty_like(x)
    l1 = 0.0
    probes = 0
    while (l2 - l1) / (l1 + l2) > tol:
        lmid = 0.5 * (l2 + l1)
        ops.oc_probe(x, dc, dv, w, lmid, move, xnew, partial, nparts)
        vol = float(partial.sum()) if w is not None else float(partial.sum()) / n
        probes += 1
        if vol > volfrac:
            l1 = lmid
        else:
            l2 = lmid
    return xnew


def modulus(rho, penal=3.0, E0=1.0, Emin=1e-9):
    return Emin + (E0 - Emin) * rho ** penal


class Compliance(torch.autograd.Function):
    """c(rho) = f . u with K(E(rho)) u = f.  `problem` provides .solve(E) -> u (device tensor),
    .force, .blocks, .shape, .mask (ops.Mask), .penal, .E0, .Emin."""

    @staticmethod
    def forward(ctx, rho, problem):
        E = modulus(rho.detach(), problem.penal, problem.E0, problem.Emin)
        u = problem.solve(E)
        ctx.problem = problem
        ctx.save_for_backward(rho, u)
        return (problem.force @ u).reshape(())

    @staticmethod
    def backward(ctx, grad):
        rho, u = ctx.saved_tensors
        pb = ctx.problem
        ce = torch.empty_like(rho)
        dc = torch.empty_like(rho)
        ops.element_energy(pb.blocks, u, ce, pb.shape, mask=pb.mask, rho=rho.detach().contiguous(), dc=dc,
                           penal=pb.penal, dE=pb.E0 - pb.Emin)
        return grad * dc, None


class SimpProblem:
    """State problem of the SIMP loop: K(E) u = f on the device-resident GMG-FCG engine
    (voxelhip.engine.GMGSolver; precision "mixed" = FP32 V-cycle in the FP64 outer iteration).
    solve(E) rebuilds the hierarchy for E and returns a copy of u; the iteration count and the
    residual of the last solve are kept in last_iterations / last_rel."""

    def __init__(self, shape, ke, mask, force, *, precision="mixed", tol=1e-8, maxiter=200, penal=3.0, E0=1.0,
                 Emin=1e-9, graph=False, **solver_kw):
        from .engine import GMGSolver
        import voxel_kernels as vk
        self.shape = tuple(int(s) for s in shape)
        self.penal, self.E0, self.Emin = float(penal), float(E0), float(Emin)
        self.tol, self.maxiter = float(tol), int(maxiter)
        self.blocks = vk.Modal(ke).blocks
        self.force = force if isinstance(force, torch.Tensor) else torch.from_numpy(force)
        self.force = self.force.to("cuda", torch.float64).contiguous()
        self.mask = mask if isinstance(mask, ops.Mask) else ops.Mask.from_values(mask, self.shape)
        solver_kw.setdefault("smoother", "chebyshev")
        self.solver = GMGSolver(self.shape, ke, self.mask, None, precision=precision, graph=graph, **solver_kw)
        self.last_iterations, self.last_rel = None, None

    def solve(self, E):
        self.solver.update(E)
        x, it, ok, rel, _ = self.solver.solve(self.force, tol=self.tol, maxiter=self.maxiter)
        if not ok:
            raise RuntimeError(f"state solve did not reach {self.tol:g} in {it} iterations (rel {rel:.3e})")
        self.last_iterations, self.last_rel = it, rel
        return x.clone()


def simp_run(problem, steps, volfrac=0.3, rmin=1.5, move=0.2, stop_change=None):
    """Density-filtered SIMP with the top88 OC, everything on the device.  Returns (history, rho)."""
    filt = ConeFilter(problem.shape, rmin, device=problem.force.device)
    n = filt.n
    x = torch.full((n,), float(volfrac), dtype=torch.float64, device=problem.force.device)
    dv_x = filt.transpose(torch.ones_like(x))
    rho = filt.forward(x)
    history = []
    for k in range(steps):
        t0 = time.perf_counter()
        r = rho.detach().requires_grad_(True)
        c = Compliance.apply(r, problem)
        c.backward()
        dc_x = filt.transpose(r.grad.contiguous())
        x_new = oc_top88(x, dc_x, dv_x, volfrac, move, filt=filt)
        rho = filt.forward(x_new)
        change = float((x_new - x).abs().max())
        x = x_new
        history.append({"step": k, "compliance": float(c), "change": change, "volume": float(rho.mean()),
                        "iterations": getattr(problem, "last_iterations", None),
                        "seconds": time.perf_counter() - t0})
        if stop_change is not None and change <= stop_change:
            break
    return history, rho
