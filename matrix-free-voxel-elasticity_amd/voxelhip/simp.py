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
  run_simp(problem, ...)             the donor's full driver (simp_r2.py:388-433, 552-918): the
                                     four-phase continuation of penalty, Heaviside sharpness,
                                     move limit and filter radius, the projected OC update with a
                                     warm-started multiplier bracket, warm-started state solves
                                     with the fail-closed retry from zero, the physical-change and
                                     compliance-window stopping test, and the closing solve on the
                                     final design
  project / dproject, schedule_at    Heaviside projection (vh_heaviside) and the continuation table
  FilterFn, ProjectFn, design_compliance
                                     the design gradient end to end through filter and projection
                                     behind torch.autograd
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

    def set_rmin(self, rmin):
        """Change the radius (element units) and recompute the row sums (filter_r2.py:115-129)."""
        rmin = float(rmin)
        if abs(rmin - self.rmin) < 1e-12:
            return
        self.rmin = rmin
        ones = torch.ones_like(self.s)
        ops.cone_filter(ones, self.s, self.shape, self.rmin)
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
    """Returns x_new of the last bisection probe, as top88 does.

    The bisection's probes are the ones of the reference loop, in its order and with its
    midpoints 0.5 (l2 + l1); they are evaluated several at a time: while l1 = 0 the next midpoints
    are l2/2, l2/4, ... whatever the volumes, and afterwards the midpoints of the next levels of
    the bisection tree (every branch) are known in advance.  One pass of vh_oc_probe_multi gives
    the volumes of up to 16 of them; the host walks the reference loop on those volumes and stops
    where it stops.  The last probe's x_new is then computed by one vh_oc_probe pass."""
    n = x.numel()
    nparts = min(1024, max(1, (n + 255) // 256))
    K = ops.OC_MAX_PROBES
    partial = torch.zeros(K * nparts, dtype=torch.float64, device=x.device)
    w = filt.w_vol if filt is not None else None

    def volumes(lams):
        ops.oc_probe_multi(x, dc, dv, w, lams, move, partial, nparts)
        v = partial[:len(lams) * nparts].view(len(lams), nparts).sum(dim=1).tolist()
        return v if w is not None else [vk / n for vk in v]

    l1, l2 = 0.0, float(l2)
    lmid = None
    # while l1 = 0 the probes halve l2 until the first volume above the target
    while l1 == 0.0 and (l2 - l1) / (l1 + l2) > tol:
        lams, lb = [], l2
        for _ in range(K):
            lb = 0.5 * (lb + l1)
            lams.append(lb)
        for lm_, vol in zip(lams, volumes(lams)):
            lmid = lm_
            if vol > volfrac:
                l1 = lmid
                break
            l2 = lmid
    # then the tree of the next levels of midpoints (depth 4: 15 probes per pass)
    depth = 4
    while (l2 - l1) / (l1 + l2) > tol:
        nodes = [(l1, l2)]                      # breadth-first: node j has children 2j+1, 2j+2
        lams = []
        for j in range(2 ** depth - 1):
            a, b = nodes[j]
            m = 0.5 * (b + a)
            lams.append(m)
            nodes.append((a, m))                # volume <= target: l2 = m
            nodes.append((m, b))                # volume > target:  l1 = m
        vols = volumes(lams)
        j = 0
        while j < len(lams) and (l2 - l1) / (l1 + l2) > tol:
            lmid = lams[j]
            if vols[j] > volfrac:
                l1, j = lmid, 2 * j + 2
            else:
                l2, j = lmid, 2 * j + 1
    xnew = torch.empty_like(x)
    if lmid is not None:
        ops.oc_probe(x, dc, dv, w, lmid, move, xnew, partial, nparts)
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


class _JacobiPCG:
    """K(E) u = f by the device Jacobi-PCG (voxelhip.solvers.pcg, true residual every
    `check_every` iterations) on the fused FP64 product: the state solver for grids whose extents
    do not halve far enough for a multigrid hierarchy with a small coarsest level."""

    def __init__(self, shape, ke, mask, check_every=50):
        import voxel_mg as mg
        self._mg, self.shape, self.mask, self.check_every = mg, shape, mask, check_every
        self.product = mg.HipProduct(shape, ke)
        self.level = None

    def update(self, E):
        from . import xp
        self.level = self._mg.Level(xp, self.shape, self.product, E, self.mask, torch.float64)

    def solve(self, b, tol=1e-8, maxiter=20000, x0=None):
        from .solvers import pcg
        lvl = self.level
        x, it, ok, rel = pcg(lvl.apply, b, lvl.inv_diag, x0=x0, tol=tol, maxiter=maxiter, mask=lvl.mask,
                             check_every=self.check_every)
        return x, it, ok, rel, None


class SimpProblem:
    """State problem of the SIMP loop: K(E) u = f on the device-resident GMG-FCG engine
    (voxelhip.engine.GMGSolver; precision "mixed" = FP32 V-cycle in the FP64 outer iteration),
    or, with method="pcg", the device Jacobi-PCG in FP64 (for grids that do not coarsen).
    solve(E) rebuilds the hierarchy for E and returns a copy of u; the iteration count and the
    residual of the last solve are kept in last_iterations / last_rel, whether the last solve
    fell back from a warm start to a solve from zero in last_retried."""

    def __init__(self, shape, ke, mask, force, *, precision="mixed", tol=1e-8, maxiter=200, penal=3.0, E0=1.0,
                 Emin=1e-9, graph=False, warm_start=False, method="gmg", **solver_kw):
        import voxel_kernels as vk
        self.shape = tuple(int(s) for s in shape)
        self.penal, self.E0, self.Emin = float(penal), float(E0), float(Emin)
        self.tol, self.maxiter = float(tol), int(maxiter)
        self.blocks = vk.Modal(ke).blocks
        self.force = force if isinstance(force, torch.Tensor) else torch.from_numpy(force)
        self.force = self.force.to("cuda", torch.float64).contiguous()
        self.mask = mask if isinstance(mask, ops.Mask) else ops.Mask.from_values(mask, self.shape)
        if method == "pcg":
            if precision != "fp64":
                raise ValueError("method='pcg' solves in FP64")
            self.solver = _JacobiPCG(self.shape, ke, self.mask, **solver_kw)
        elif method == "gmg":
            from .engine import GMGSolver
            solver_kw.setdefault("smoother", "chebyshev")
            if warm_start:       # the spectral bounds also continue from the previous state's
                solver_kw.setdefault("bound_warm_iterations", 5)
            self.solver = GMGSolver(self.shape, ke, self.mask, None, precision=precision, graph=graph, **solver_kw)
        else:
            raise ValueError(f"method: {method!r}")
        self.method = method
        # warm start: each solve starts from the previous state (the tolerance stays relative to
        # |f|), with the fail-closed retry from zero of simp_r2.py:680-708
        self.warm_start = bool(warm_start)
        self._u = None
        self.last_iterations, self.last_rel, self.last_retried = None, None, False

    def _cold_retry(self, E):
        """The retry from zero after a warm start missed the tolerance.  A GMG rebuild whose
        Chebyshev bounds continued the previous power iteration re-estimates them from the seeded
        vector first: a short continued iteration can lag a spectrum the new design has moved."""
        s = self.solver
        if getattr(s, "bound_warm_iterations", None) is not None and getattr(s, "_bound_vectors", None) is not None:
            s._bound_vectors = None
            s.update(E)

    def solve(self, E):
        self.solver.update(E)
        x0 = self._u if self.warm_start else None
        x, it, ok, rel, _ = self.solver.solve(self.force, tol=self.tol, maxiter=self.maxiter, x0=x0)
        self.last_retried = False
        if not ok and x0 is not None:
            self.last_retried = True
            self._cold_retry(E)
            x, it2, ok, rel, _ = self.solver.solve(self.force, tol=self.tol, maxiter=self.maxiter)
            it += it2
        if not ok:
            raise RuntimeError(f"state solve did not reach {self.tol:g} in {it} iterations (rel {rel:.3e})")
        self.last_iterations, self.last_rel = it, rel
        u = x.clone()
        if self.warm_start:
            self._u = u.clone()
        return u


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
        history.append({"step": k, "compliance": float(c.detach()), "change": change, "volume": float(rho.mean()),
                        "iterations": getattr(problem, "last_iterations", None),
                        "retried": bool(getattr(problem, "last_retried", False)),
                        "seconds": time.perf_counter() - t0})
        if stop_change is not None and change <= stop_change:
            break
    return history, rho


# --------------------------------------------------------------------------- projected SIMP driver
# (until_iter, penal, beta, move, rmin): the donor's four-phase continuation (simp_r2.py:411-417)
DEFAULT_SCHEDULE = [
    (15, 1.5, 1.0, 0.20, 1.50),
    (40, 3.5, 4.0, 0.15, 1.35),
    (65, 4.5, 16.0, 0.08, 1.25),
    (120, 4.5, 32.0, 0.05, 1.20),
]


def schedule_at(it, schedule=None):
    """(penal, beta, move, rmin, in_last_phase) at outer iteration `it` (1-based); past the last
    phase the parameters stay fixed (simp_r2.py:420-433)."""
    sch = schedule or DEFAULT_SCHEDULE
    for until, p, beta, move, rmin in sch:
        if it <= until:
            return p, beta, move, rmin, until == sch[-1][0]
    p, beta, move, rmin = sch[-1][1:]
    return p, beta, move, rmin, True


def project(rho_bar, beta, eta=0.5, out=None):
    y = out if out is not None else torch.empty_like(rho_bar)
    ops.heaviside(rho_bar, beta, eta, rho=y)
    return y


def dproject(rho_bar, beta, eta=0.5, out=None):
    y = out if out is not None else torch.empty_like(rho_bar)
    ops.heaviside(rho_bar, beta, eta, drho=y)
    return y


def _parts(n):
    return min(1024, max(1, (n + 255) // 256))


def oc_update(rho, dc, dv, volfrac, move, filt, beta, eta, rho_lb, bisect_tol=1e-8, max_bisect=80,
              lam_hint=None):
    """One OC step bisecting on the physical (projected) volume, with the bracket seeded from
    the previous multiplier (simp_r2.py:467-549).  Returns (rho_new, achieved_volume, lam, probes)."""
    n = rho.numel()
    nparts = _parts(n)
    partial = torch.zeros(nparts, dtype=torch.float64, device=rho.device)
    cand, fb = torch.empty_like(rho), torch.empty_like(rho)

    def design(lmid):
        ops.oc_design(rho, dc, dv, lmid, move, rho_lb, cand)
        filt.forward(cand, out=fb)
        ops.heaviside(fb, beta, eta, partial=partial)
        return float(partial.sum()) / n

    probes = 0
    if lam_hint is not None and lam_hint > 0:
        l1, l2 = lam_hint / 4.0, lam_hint * 4.0
        for _ in range(30):              # widen both ends 8x until the volumes straddle volfrac
            v1 = design(l1)
            v2 = design(l2)
            probes += 2
            if v1 > volfrac and v2 <= volfrac:
                break
            if v1 <= volfrac:
                l1 /= 8.0
            if v2 > volfrac:
                l2 *= 8.0
            if l1 < 1e-30 or l2 > 1e30:
                l1, l2 = 0.0, 1e9
                break
    else:
        l1, l2 = 0.0, 1e9
    lam = 0.5 * (l1 + l2)
    achieved = design(lam)
    probes += 1
    for _ in range(max_bisect):
        if (l2 - l1) <= bisect_tol * (l1 + l2 + 1e-30):
            break
        lam = 0.5 * (l1 + l2)
        achieved = design(lam)
        probes += 1
        if achieved > volfrac:
            l1 = lam
        else:
            l2 = lam
    return cand.clone(), achieved, lam, probes


class LinearSolveNotConverged(RuntimeError):
    """A state solve missed its tolerance from zero as well (the fail-closed rule)."""


def _state_solve(problem, E, x0, tol):
    """Warm-started solve with the fail-closed retry from zero (simp_r2.py:680-708).  Returns
    (u, iterations including a retry, rel, retried)."""
    s = problem.solver
    s.update(E)
    x, it, ok, rel, _ = s.solve(problem.force, tol=tol, maxiter=problem.maxiter, x0=x0)
    retried = False
    if not ok and x0 is not None:
        retried = True
        problem._cold_retry(E)
        x, it2, ok, rel, _ = s.solve(problem.force, tol=tol, maxiter=problem.maxiter)
        it += it2
    if not ok:
        raise LinearSolveNotConverged(f"state solve missed tol={tol:g} (relative residual {rel:.3e} after {it} "
                                      f"iterations)")
    return x.clone(), it, rel, retried


def run_simp(problem, volfrac=0.30, rmin=1.5, penal=None, E0=1.0, Emin=1e-9, rho_lb=1e-3, eta=0.5, cg_tol=1e-5,
             warm_start=True, max_outer=400, schedule=None, conv_drho=0.01, conv_dc_rel=1e-4, conv_window=10,
             fixed_budget=None, verbose=False):
    """The donor's SIMP driver (simp_r2.py:552-918) on the device: filter -> Heaviside projection ->
    modulus, warm-started GMG-FCG state solve (fail-closed), adjoint sensitivities through the
    projection and the filter, projected OC update; stops after the last continuation phase when
    the physical change and the compliance spread over `conv_window` steps fall below their
    bounds (or after `fixed_budget` steps), then solves once more on the final design.  `problem`
    is a SimpProblem (its solver, force, mask and element blocks are used).  Returns a dict with
    the history, the solve log and the final design's compliance, volume and grayness."""
    t_start = time.perf_counter()
    dev = problem.force.device
    filt = ConeFilter(problem.shape, rmin, device=dev)
    n = filt.n
    rho = torch.full((n,), float(volfrac), dtype=torch.float64, device=dev)
    rho_bar, rho_phys, dprj, E = (torch.empty_like(rho) for _ in range(4))
    ce, dc_drp = torch.empty_like(rho), torch.empty_like(rho)
    u = None
    history, solve_log = [], []
    total_cg, retries = 0, 0
    converged, stop_reason = False, "guard"
    lam_prev, rho_phys_prev = None, None
    budget = fixed_budget or max_outer
    it = 0
    while it < budget:
        it += 1
        p_it, beta_it, move_it, rmin_it, last_phase = schedule_at(it, schedule)
        if penal is not None:
            p_it = penal
        filt.set_rmin(rmin_it)
        filt.forward(rho, out=rho_bar)
        ops.heaviside(rho_bar, beta_it, eta, p_it, E0, Emin, rho=rho_phys, drho=dprj, E=E)
        t0 = time.perf_counter()
        u, iters, rel, retried = _state_solve(problem, E, u if (warm_start and it > 1) else None, cg_tol)
        t_solve = time.perf_counter() - t0
        total_cg += iters
        retries += int(retried)
        solve_log.append({"iter": it, "cg_iters": iters, "rel_resid": rel, "retried_from_zero": retried,
                          "solve_s": t_solve, "penal": p_it, "beta": beta_it, "rmin": filt.rmin})
        c = float(problem.force @ u)
        ops.element_energy(problem.blocks, u, ce, problem.shape, mask=problem.mask, rho=rho_phys, dc=dc_drp,
                           penal=p_it, dE=E0 - Emin)
        dc = filt.transpose(dc_drp * dprj)
        dv = filt.transpose(dprj / n)
        rho_old = rho
        rho, achieved, lam, probes = oc_update(rho, dc, dv, volfrac, move_it, filt, beta_it, eta, rho_lb,
                                               lam_hint=lam_prev)
        lam_prev = lam
        change = float((rho - rho_old).abs().max())
        change_phys = float((rho_phys - rho_phys_prev).abs().max()) if rho_phys_prev is not None else float("inf")
        rho_phys_prev = rho_phys.clone()
        gray = float((4.0 * rho_phys * (1.0 - rho_phys)).mean())
        history.append({"iter": it, "compliance": c, "change": change,
                        "change_phys": None if change_phys == float("inf") else change_phys,
                        "vol_phys": achieved, "vol_raw": float(rho.mean()), "grayness": gray, "cg_iters": iters,
                        "rel_resid": rel, "penal": p_it, "beta": beta_it, "move": move_it, "rmin": filt.rmin,
                        "oc_probes": probes, "retried": retried})
        if verbose:
            print(f"  it {it:4d}  c={c:.6f}  dr={change:.4f}  drp={change_phys:.4f}  V={achieved:.4f}  "
                  f"g={gray:.4f}  cg={iters:4d}  p={p_it:.1f} b={beta_it:.0f}", flush=True)
        if fixed_budget is None and last_phase and len(history) > conv_window:
            recent = [h["compliance"] for h in history[-(conv_window + 1):]]
            spread = (max(recent) - min(recent)) / max(abs(recent[-1]), 1e-30)
            if change_phys <= conv_drho and spread <= conv_dc_rel:
                converged, stop_reason = True, "outer_criterion"
                break
    if fixed_budget is not None:
        stop_reason = "fixed_budget"
    # one more solve, on the design the run ends on (simp_r2.py:834-880)
    p_end, beta_end = schedule_at(it, schedule)[:2]
    if penal is not None:
        p_end = penal
    filt.forward(rho, out=rho_bar)
    ops.heaviside(rho_bar, beta_end, eta, p_end, E0, Emin, rho=rho_phys, E=E)
    u_f, it_f, rel_f, _ = _state_solve(problem, E, None, cg_tol)
    total_cg += it_f
    solve_log.append({"iter": it + 1, "cg_iters": it_f, "rel_resid": rel_f, "retried_from_zero": False,
                      "final_iterate": True, "penal": p_end, "beta": beta_end, "rmin": filt.rmin})
    return {
        "outer_iters": it, "outer_converged": converged, "stop_reason": stop_reason,
        "final_compliance": float(problem.force @ u_f),
        "final_compliance_pre_update": history[-1]["compliance"] if history else float("nan"),
        "final_vol_phys": float(rho_phys.mean()),
        "final_grayness": float((4.0 * rho_phys * (1.0 - rho_phys)).mean()),
        "rho_phys_min": float(rho_phys.min()), "E_min_achieved": float(E.min()),
        "total_cg_iters": total_cg, "warm_start_retries": retries,
        "max_cg_iters_single_solve": max(s_["cg_iters"] for s_ in solve_log),
        "max_rel_resid": max(s_["rel_resid"] for s_ in solve_log),
        "wall_s": time.perf_counter() - t_start,
        "history": history, "solve_log": solve_log,
        "rho": rho, "rho_phys": rho_phys,
    }


class FilterFn(torch.autograd.Function):
    """rho_bar = H x / s with backward H (g / s)."""

    @staticmethod
    def forward(ctx, x, filt):
        ctx.filt = filt
        return filt.forward(x.detach().contiguous())

    @staticmethod
    def backward(ctx, g):
        return ctx.filt.transpose(g.contiguous()), None


class ProjectFn(torch.autograd.Function):
    """Heaviside projection with its analytic derivative (vh_heaviside)."""

    @staticmethod
    def forward(ctx, rho_bar, beta, eta):
        rb = rho_bar.detach().contiguous()
        r, d = torch.empty_like(rb), torch.empty_like(rb)
        ops.heaviside(rb, beta, eta, rho=r, drho=d)
        ctx.save_for_backward(d)
        return r

    @staticmethod
    def backward(ctx, g):
        (d,) = ctx.saved_tensors
        return g * d, None, None


def design_compliance(x, problem, filt, beta=0.0, eta=0.5):
    """c(x) = f . u for the design variable x through filter, projection and the SIMP modulus:
    differentiable end to end (the adjoint kernel, the projection derivative and the filter
    transpose in backward)."""
    return Compliance.apply(ProjectFn.apply(FilterFn.apply(x, filt), float(beta), float(eta)), problem)
