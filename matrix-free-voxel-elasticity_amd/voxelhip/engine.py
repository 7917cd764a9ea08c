"""Device-resident GMG-preconditioned flexible CG solver with optional HIP-graph capture.

`GMGSolver` is the solve path behind the benchmark configurations: it builds the hierarchy of
voxel_mg (rediscretised levels with ke * 2^l, 2x2x2 modulus averaging, injected masks, a dense
coarsest level), picks the smoother (damped Jacobi as src/voxel_mg.py:161-175, or Chebyshev-
Jacobi), and runs the flexible CG of src/voxel_mg.py:181-217 with every vector in fixed
buffers, so that one iteration (A p with p.q, the update, the true-residual test, the V-cycle,
the Polak-Ribiere beta and the direction) is a fixed sequence of kernel launches on one stream.
With `graph=True` that sequence is captured once in two HIP graphs (the z / z_old roles swap
every iteration) and replayed; the host only reads the residual norm at the check points.

precision:
  "fp64"          FP64 everywhere (the reference's plain FP64 GMG-FCG path)
  "mixed"         FP64 outer flexible CG and true residual, FP32 V-cycle (matvec, smoothing,
                  transfers, coarsest solve in FP32), FP64 dot accumulation.  With
                  fp32_direction (the default) the search direction p = z + beta p_old and the
                  product p.Ap that gives alpha are formed in FP32 as well (z is already an FP32
                  V-cycle output); the iterate x, the residual b - A x formed every iteration and
                  the convergence test stay FP64, so an FP32 rounding of p or alpha only changes
                  the step taken, never the residual it is judged by.
  "ir"            FP64 iterative refinement around an FP32 flexible CG with the FP32 V-cycle
                  (the scheme of pcg_ir, experiments/session_a/donor/src/gpu_fem/simp_r2.py:298-385,
                  with GMG in place of Jacobi): FP64 only for the outer residual b - A x, the
                  correction update x += |r| d, the dot accumulation and the coarsest solve;
                  the inner solve of A d = r / |r| stops at max(inner_tol, 0.5 tol / rel).

coarse_degree / coarse_level: smoothing steps on the levels from coarse_level down (the donor's
  coarse_smooth_iters, multigrid_v4.py:466,934-967); None keeps degree / pre / post there.

vcycle_storage="bf16" (mixed / ir): the finest level's smoothing and residual passes read the
  V-cycle iterate through a bfloat16 copy (FP32 arithmetic): the BF16-storage ablation.

omega_seed: the seed of the power iteration behind the Jacobi omegas (src/voxel_mg.py:228-244).

bound_warm_iterations: for a sequence of rebuilds with slowly changing moduli (a SIMP run), the
  Chebyshev bounds of each rebuild after the first continue the previous rebuild's power iteration
  for this many steps instead of 30 steps from the seeded vector (None: always from the seed).

coarse:
  "rediscretize"  every coarse level rediscretised (ke * 2^l with averaged moduli; src/voxel_mg.py)
  "galerkin"      rediscretised levels, the coarsest one the Galerkin P^T A P of the level above
  "galerkin-all"  every coarse operator Galerkin, A_l = M (P^T A_(l-1) P) M + (I - M)
                  (donor_yang_gmg multigrid_v4.py:132-270,747-772): the non-coarsest levels as
                  assembled 27-point stencils (StencilProduct), the coarsest dense
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib as L
from . import ops

import voxel_mg as mg


class GMGSolver:
    def __init__(self, shape, ke, mask, E=None, *, max_coarse_x=16, n_levels=None, smoother="jacobi",
                 degree=3, pre=4, post=4, precision="fp64", coarse="rediscretize", graph=False,
                 lower_frac=1.0 / 30.0, omega_iterations=30, inner_tol=1e-3, coarse_fp64=True, tune=False,
                 coarse_degree=None, coarse_level=2, fp32_direction=True, bound_warm_iterations=None,
                 omega_seed=20260923, vcycle_storage=None):
        self.shape = tuple(int(s) for s in shape)
        self.ke = np.asarray(ke, dtype=np.float64)
        self.dev = torch.device("cuda", torch.cuda.current_device())
        shapes = mg.hierarchy_shapes(self.shape, max_coarse_x)
        if n_levels is not None:
            shapes = shapes[:n_levels]
        self.shapes = shapes
        self.smoother, self.degree, self.pre, self.post = smoother, degree, pre, post
        # smoothing steps on the levels from coarse_level down (None: degree / pre / post there too)
        self.coarse_degree, self.coarse_level = coarse_degree, coarse_level
        if precision not in ("fp64", "mixed", "ir"):
            raise ValueError(precision)
        self.precision = precision
        self.fp32_direction = bool(fp32_direction)
        # Chebyshev bounds of a rebuild continue the previous rebuild's power iteration for this many
        # steps (None: 30 steps from the seeded vector every time)
        self.bound_warm_iterations = bound_warm_iterations
        self._bound_vectors = None
        self.pdtype = torch.float64 if precision == "fp64" else torch.float32
        self.inner_tol = inner_tol
        # the FP64 coarsest solve inside the FP32 V-cycle of the refinement scheme
        self.coarse_dtype = torch.float64 if (precision == "ir" and coarse_fp64) else None
        self.last_outer = 0
        # launch tuning of every product pass (voxelhip.tuner): False, True (a per-process
        # LaunchTuner) or a LaunchTuner instance
        if tune is True:
            from .tuner import LaunchTuner
            tune = LaunchTuner()
        self.tuner = tune or None
        if coarse not in ("rediscretize", "galerkin", "galerkin-all"):
            raise ValueError(f"coarse: {coarse!r}")
        self.coarse = coarse
        self.graph = graph
        self.lower_frac = lower_frac
        self.omega_iterations = omega_iterations
        self.omega_seed = int(omega_seed)
        if vcycle_storage not in (None, "bf16") or (vcycle_storage and precision == "fp64"):
            raise ValueError("vcycle_storage='bf16' is for the FP32 V-cycle of the mixed and ir precisions")
        self.vcycle_storage = vcycle_storage
        self.mask = mask if isinstance(mask, ops.Mask) else ops.Mask.from_values(mask, self.shape, self.dev)
        self.fine = None
        self.M = None
        self._graphs = None
        self._start_graph = None
        if E is not None:
            self.update(E)

    # ------------------------------------------------------------------ setup per modulus field
    def update(self, E):
        """(Re)build every level for a new modulus field (device or host array)."""
        E64 = E.to(self.dev, torch.float64).contiguous() if isinstance(E, torch.Tensor) else \
            torch.from_numpy(np.ascontiguousarray(E, dtype=np.float64)).to(self.dev)
        galerkin = self._galerkin_coarsest(E64) if (self.coarse == "galerkin" and len(self.shapes) > 1) else None
        stencils = None
        if self.coarse == "galerkin-all" and len(self.shapes) > 1:
            stencils, galerkin = galerkin_hierarchy(self.shapes, self.ke, self.mask, E64)
        self.fine, self.M = build_hierarchy(self.shape, self.ke, self.mask, E64, self.shapes, self.pdtype,
                                            smoother=self.smoother, degree=self.degree, pre=self.pre,
                                            post=self.post, lower_frac=self.lower_frac,
                                            omega_iterations=self.omega_iterations, omega_seed=self.omega_seed,
                                            storage=self.vcycle_storage,
                                            coarse_dtype=self.coarse_dtype, coarse_matrix=galerkin,
                                            stencils=stencils, coarse_degree=self.coarse_degree,
                                            coarse_level=self.coarse_level, **self._bound_args())
        if getattr(self, "_pending_bounds", None) is not None:     # the rebuild completed
            self._bound_vectors, self._pending_bounds = self._pending_bounds, None
        # the last post-smoothing step of the V-cycle forms the Polak-Ribiere dots itself (and
        # writes z in the V-cycle precision, which the DOT pass reads directly)
        self.fused_pr = (self.smoother == "chebyshev" and len(self.shapes) > 1 and self.degree >= 1
                         and self.precision != "ir")
        # FP32 search direction (mixed, with z in FP32 from the fused V-cycle)
        self.p32 = self.fp32_direction and self.precision == "mixed" and self.fused_pr
        if self.tuner is not None:
            self.M.tune(self.tuner)
            self.fine.tune(self.tuner, ([] if self.p32 else [(L.MODE_DOT, True)]) +
                           [(L.MODE_RESID, True), (L.MODE_RESID, False)])
            if self.fused_pr:
                self.M.levels[0].tune(self.tuner, [(L.MODE_CHEB3D, False)] + ([(L.MODE_DOT, True)] if self.p32 else []))
            if self.precision == "ir":
                self.M.levels[0].tune(self.tuner, [(L.MODE_DOT, True), (L.MODE_RESID, True)])
        self._alloc()
        self._graphs = None
        self._start_graph = None

    def _bound_args(self):
        if self.bound_warm_iterations is None or self.smoother != "chebyshev":
            return {}
        prev, self._bound_vectors = self._bound_vectors, None    # (a rebuild that fails keeps none)
        warm = prev is not None and len(prev) == len(self.shapes) - 1
        keep = []
        self._pending_bounds = keep
        return dict(bound_start=prev if warm else None, bound_keep=keep,
                    bound_iterations=int(self.bound_warm_iterations) if warm else 30)

    def _galerkin_coarsest(self, E64):
        """A_c = P^T A_{L-2} P with the coarse mask re-imposed, assembled on the device."""
        fshape = self.shapes[-2]
        lvl = len(self.shapes) - 2
        Ef, m = E64, self.mask
        for i in range(1, lvl + 1):
            Ef = mg.coarsen_modulus(None, Ef, self.shapes[i - 1])
            m = m.coarsen(self.shapes[i - 1])
        return galerkin_coarsest(fshape, self.ke * 2 ** lvl, Ef, m)

    def _alloc(self):
        n = self.fine.n
        z = lambda dt=torch.float64: torch.zeros(n, dtype=dt, device=self.dev)  # noqa: E731
        self.b, self.r = z(), z()
        self.parts = ops.reduce_parts(n)
        # per-tile partial counts of the fused passes (they depend on each pass's launch plan)
        f, f32 = self.fine, self.M.levels[0]
        self.t_dot = (f32 if self.p32 else f).tiles(L.MODE_DOT, True)
        self.t_res = f.tiles(L.MODE_RESID, True)
        self.t_res0 = f.tiles(L.MODE_RESID, False)
        if self.precision == "ir":
            self.t_dot32, self.t_res32 = f32.tiles(L.MODE_DOT, True), f32.tiles(L.MODE_RESID, True)
        else:
            self.t_dot32 = self.t_res32 = 0
        self.t_pr = f32.tiles(L.MODE_CHEB3D, False) if self.fused_pr else 0
        self.partial = torch.zeros(2 * max(self.parts, self.t_dot, self.t_res, self.t_res0, self.t_dot32,
                                           self.t_res32, self.t_pr), dtype=torch.float64, device=self.dev)
        self.scal = torch.zeros(16, dtype=torch.float64, device=self.dev)
        self.tol2 = torch.zeros(1, dtype=torch.float64, device=self.dev)     # tol^2 of the running solve
        # right-hand side of the finest V-cycle level: the FP32 copy of r written by the
        # residual pass (mixed), the inner residual (ir), or r itself (FP64)
        self.rhs0 = self.M._work[0]["rhs"] if self.pdtype != torch.float64 else self.r
        if self.precision == "ir":
            f32 = torch.float32
            self.x = z()
            self.b32 = z(f32)
            self.ds, self.ps, self.zs = [z(f32), z(f32)], [z(f32), z(f32)], [z(f32), z(f32)]
        else:
            zdt = self.pdtype if self.fused_pr else torch.float64
            pdt = torch.float32 if self.p32 else torch.float64
            self.xs, self.ps, self.zs = [z(), z()], [z(pdt), z(pdt)], [z(zdt), z(zdt)]

    # ------------------------------------------------------------------ one iteration
    def _iteration(self, i):
        """Launch flexible-CG iteration with buffer parity i (no host synchronisation).

        DOT pass:   p_i = z_i + beta p_o (read on the fly), p_i.A p_i (A p_i not stored) -> alpha
                    (FP32 with fp32_direction, FP64 otherwise)
        RESID pass: x_i = x_o + alpha p_i (on the fly), r = b - A x_i, |r|^2, r -> fp32 (FP64)
        V-cycle on r into z_o; its last smoothing step forms r.z_o and r.(z_o - z_i) -> beta, r.z
        (z in the V-cycle precision; the Jacobi smoother converts and dots in a separate pass)
        """
        o = 1 - i
        f, L_ = self.fine, L
        (self.M.levels[0] if self.p32 else f).matvec(self.zs[i], None, mode=L_.MODE_DOT, partial=self.partial,
                                                     w=self.ps[o], sdev=self.scal, sidx=L_.S_BETA,
                                                     u_out=self.ps[i])
        ops.finish(self.partial, self.t_dot, 1, self.scal, L_.S_PQ, L_.OP_ALPHA)
        # (with the fused dots and an FP32 V-cycle nothing reads the FP64 residual: only its FP32
        #  copy, the V-cycle right-hand side, is written)
        mixed = self.pdtype != torch.float64
        f.matvec(self.xs[o], None if (mixed and self.fused_pr) else self.r, mode=L_.MODE_RESID, b=self.b,
                 partial=self.partial, w=self.ps[i], sdev=self.scal, sidx=L_.S_ALPHA, u_out=self.xs[i],
                 y32=self.rhs0 if mixed else None)
        ops.finish(self.partial, self.t_res, 1, self.scal, 8, L_.OP_RESNORM)
        if self.fused_pr:
            self.M.vcycle(self.rhs0, out=self.zs[o], dots=(self.partial, self.zs[i]))
            ops.finish(self.partial, self.t_pr, 2, self.scal, 5, L_.OP_BETA_PR)
        else:
            zl = self.M.vcycle(self.rhs0)
            ops.cvt_dotdiff(zl, self.r, self.zs[i], self.zs[o], self.partial)
            ops.finish(self.partial, self.parts, 2, self.scal, 5, L_.OP_BETA_PR)

    def _inner_iteration(self, i):
        """FP32 flexible-CG iteration of the refinement's inner solve A d = b32 (parity i);
        the same pass structure as _iteration with FP32 vectors and FP64 scalars."""
        o = 1 - i
        f = self.M.levels[0]
        f.matvec(self.zs[i], None, mode=L.MODE_DOT, partial=self.partial,
                 w=self.ps[o], sdev=self.scal, sidx=L.S_BETA, u_out=self.ps[i])
        ops.finish(self.partial, self.t_dot32, 1, self.scal, L.S_PQ, L.OP_ALPHA)
        f.matvec(self.ds[o], self.rhs0, mode=L.MODE_RESID, b=self.b32, partial=self.partial,
                 w=self.ps[i], sdev=self.scal, sidx=L.S_ALPHA, u_out=self.ds[i])
        ops.finish(self.partial, self.t_res32, 1, self.scal, 8)
        zl = self.M.vcycle(self.rhs0)
        ops.cvt_dotdiff(zl, self.rhs0, self.zs[i], self.zs[o], self.partial)
        ops.finish(self.partial, self.parts, 2, self.scal, 5, L.OP_BETA_PR)

    def _capture(self, step):
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):          # warm-up on a side stream (state is reset afterwards)
            step(0)
            step(1)
        torch.cuda.current_stream().wait_stream(s)
        graphs = []
        for i in (0, 1):
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                step(i)
            graphs.append(g)
        return graphs

    def _start(self, r0):
        """The start of a solve on the device: z = M r, p = z (beta = 0 against a zero previous
        direction), rz = r.z, and the scalar block (|b|^2 in slot 9, tol^2 |b|^2, iteration count
        and freeze flag cleared).  r0 None: x = 0 and r = b (|b|^2 is formed here); otherwise the
        caller has formed x0, r0 and |b|^2."""
        if r0 is None:
            ops.reduce(L.RED_NORM2, self.b, partial=self.partial)
            ops.finish(self.partial, self.parts, 1, self.scal, 9)
            self.xs[1].zero_()
            r0 = self.b
        self.ps[1].zero_()
        if self.pdtype != torch.float64:
            ops.convert(r0, self.rhs0)
        elif r0 is not self.r:
            self.r.copy_(r0)
        if self.fused_pr:
            self.M.vcycle(self.rhs0, out=self.zs[0])
            ops.reduce(L.RED_DOT, self.rhs0, self.zs[0], partial=self.partial)
        else:
            zl = self.M.vcycle(self.rhs0)
            ops.convert(zl, self.zs[0])
            ops.reduce(L.RED_DOT, r0, self.zs[0], partial=self.partial)
        ops.finish(self.partial, self.parts, 1, self.scal, L.S_RZ)
        # (slices filled on the device: assigning a Python number would copy from the host)
        for slot in (L.S_BETA, L.S_DONE, L.S_ITS):
            self.scal[slot:slot + 1].zero_()
        torch.mul(self.scal[9:10], self.tol2, out=self.scal[L.S_TOL2:L.S_TOL2 + 1])

    def _capture_start(self):
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):          # warm-up on a side stream (the replay redoes its work)
            self._start(None)
        torch.cuda.current_stream().wait_stream(s)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._start(None)
        return g

    # ------------------------------------------------------------------ solve
    def solve(self, b, tol=1e-8, maxiter=200, x0=None):
        """Flexible CG (the true residual is formed and tested every iteration, as the reference
        harness runs fcg with check_every=1) to ||b - A x|| / ||b|| <= tol from x = 0.
        Returns (x, iterations, converged, rel, matvecs); x is a view of the solver's buffer.
        With precision "ir", iterations counts the inner FP32 iterations (self.last_outer the
        outer FP64 corrections).  x0 (optional): the initial iterate (a warm start); a start that
        already meets tol returns it with 0 iterations.

        The convergence test runs on the device (OP_RESNORM freezes the recurrence once
        |r|^2 <= tol^2 |b|^2 and counts the iterations), so while the residual history shows the
        next iteration cannot converge, the host queues the one after it before reading: the GPU
        never waits for the host between iterations, and an iteration queued past convergence
        leaves x unchanged."""
        if self.precision == "ir":
            if x0 is not None:
                raise ValueError("the refinement scheme starts from zero")
            return self._solve_ir(b, tol, maxiter)
        if self.graph and self._graphs is None:
            self._graphs = self._capture(self._iteration)
        self.b.copy_(b)
        self.tol2.fill_(float(tol) ** 2)
        start_mv = 0
        if x0 is None:
            # x = 0 (r = b): the whole start runs without a host synchronisation, captured in a
            # graph of its own (launched from Python, its ~25 small coarse-level passes would each
            # wait for the host)
            if self.graph:
                if self._start_graph is None:
                    self._start_graph = self._capture_start()
                self._start_graph.replay()
            else:
                self._start(None)
        else:
            ops.reduce(L.RED_NORM2, self.b, partial=self.partial)
            ops.finish(self.partial, self.parts, 1, self.scal, 9)
            nb = float(self.scal[9].sqrt())
            self.xs[1].copy_(x0)
            self.fine.matvec(self.xs[1], self.r, mode=L.MODE_RESID, b=self.b, partial=self.partial)
            ops.finish(self.partial, self.t_res0, 1, self.scal, 8)
            start_mv = 1
            rel0 = float(self.scal[8].sqrt()) / nb if nb > 0 else 0.0
            if rel0 <= tol:
                return self.xs[1], 0, True, rel0, start_mv
            self._start(self.r)
        if not hasattr(self, "_hbuf"):
            self._hbuf = [torch.empty(16, dtype=torch.float64, pin_memory=True) for _ in range(2)]
            self._hev = [torch.cuda.Event() for _ in range(2)]
        pending, hist = [], []

        def launch(k):
            i = (k - 1) % 2
            if self._graphs is not None:
                self._graphs[i].replay()
            else:
                self._iteration(i)
            self._hbuf[k % 2].copy_(self.scal, non_blocking=True)
            self._hev[k % 2].record()
            pending.append(k)

        def next_cannot_converge():
            if not hist:
                return False
            rate = max(hist[-1] / (hist[-2] if len(hist) > 1 else 1.0),
                       hist[-2] / hist[-3] if len(hist) > 2 else 0.0)
            return hist[-1] * rate > 2.0 * tol

        launch(1)
        while True:
            if len(pending) == 1 and pending[0] < maxiter and next_cannot_converge():
                launch(pending[0] + 1)
            k = pending.pop(0)
            self._hev[k % 2].synchronize()
            v = self._hbuf[k % 2].tolist()
            nb = float(np.sqrt(v[9]))                 # |b|, from the scalar block of the start
            rel = float(np.sqrt(v[8])) / nb if nb > 0 else 0.0
            hist.append(rel)
            if v[L.S_DONE] != 0.0 or k >= maxiter:
                for q in pending:                     # iterations queued past convergence
                    self._hev[q % 2].synchronize()
                done = v[L.S_DONE] != 0.0
                its = int(v[L.S_ITS]) if done else k
                return self.xs[(its - 1) % 2], its, done, rel, 2 * its + start_mv
            if not pending:
                launch(k + 1)

    def _outer_residual(self, nb):
        self.fine.matvec(self.x, self.r, mode=L.MODE_RESID, b=self.b, partial=self.partial)
        ops.finish(self.partial, self.t_res0, 1, self.scal, 8)
        return float(self.scal[8].sqrt()) / nb

    def _solve_ir(self, b, tol, maxiter):
        if self.graph and self._graphs is None:
            self._graphs = self._capture(self._inner_iteration)
        self.b.copy_(b)
        ops.reduce(L.RED_NORM2, self.b, partial=self.partial)
        ops.finish(self.partial, self.parts, 1, self.scal, 9)
        nb = float(self.scal[9].sqrt())
        self.x.zero_()
        self.r.copy_(self.b)
        rel = 1.0 if nb > 0 else 0.0
        inner_total, matvecs, outer = 0, 0, 0
        tol_floor = self.inner_tol
        rel_prev, last_scale, last_d = None, 0.0, None
        while rel > tol and inner_total < maxiter:
            if rel_prev is not None and rel > rel_prev and last_d is not None:
                ops.convert(last_d, self.x, scale=-last_scale, accumulate=True)    # undo the step
                rel, last_d = rel_prev, None
                tol_floor *= 0.1
                if tol_floor < 1e-7:
                    break
                self._outer_residual(nb)
                matvecs += 1
                continue
            rel_prev = rel
            rnorm = rel * nb
            tau = max(tol_floor, 0.5 * tol / rel)
            # inner: A d = r / |r| in FP32 from d = 0
            ops.convert(self.r, self.b32, scale=1.0 / rnorm)
            self.ds[1].zero_()
            self.ps[1].zero_()
            self.rhs0.copy_(self.b32)
            zl = self.M.vcycle(self.rhs0)
            self.zs[0].copy_(zl)
            ops.reduce(L.RED_DOT, self.rhs0, self.zs[0], partial=self.partial)
            ops.finish(self.partial, self.parts, 1, self.scal, L.S_RZ)
            self.scal[L.S_BETA] = 0.0
            it = 0
            for it in range(1, maxiter - inner_total + 1):
                i = (it - 1) % 2
                if self._graphs is not None:
                    self._graphs[i].replay()
                else:
                    self._inner_iteration(i)
                if float(self.scal[8].sqrt()) <= tau:
                    break
            inner_total += it
            matvecs += 2 * it
            last_d, last_scale = self.ds[(it - 1) % 2], rnorm
            ops.convert(last_d, self.x, scale=rnorm, accumulate=True)            # x += |r| d
            outer += 1
            rel = self._outer_residual(nb)
            matvecs += 1
        self.last_outer = outer
        return self.x, inner_total, bool(rel <= tol), rel, matvecs


def build_hierarchy(shape, ke, mask, E64, shapes, pdtype, *, smoother="chebyshev", degree=3, pre=4, post=4,
                    lower_frac=1.0 / 30.0, omega_iterations=30, coarse_dtype=None, coarse_matrix=None,
                    stencils=None, coarse_degree=None, coarse_level=2, bound_start=None, bound_keep=None,
                    bound_iterations=30, omega_seed=20260923, storage=None):
    """The FP64 fine Level and the Multigrid (levels in pdtype, rediscretised ke * 2^l, 2x2x2
    modulus averaging, injected masks, explicit inverse of the coarsest operator, spectral
    omegas / lambdas) of one problem.  coarse_matrix: a precomputed coarsest operator (Galerkin);
    otherwise the rediscretised one is assembled.  stencils: assembled Galerkin operators
    (StencilProduct) of levels 1 .. L-2, which then replace the rediscretised ones.
    bound_start / bound_keep / bound_iterations: the power iteration of the Chebyshev bounds
    (mg.chebyshev_bounds start, keep and iterations)."""
    dev = E64.device
    prods = [mg.HipProduct(s, ke * 2 ** i, pdtype) for i, s in enumerate(shapes)]
    if stencils is not None:
        if len(stencils) != len(shapes) - 2:
            raise ValueError("one stencil per level between the finest and the coarsest")
        prods[1:-1] = [sp.astype(pdtype) for sp in stencils]
    fine = mg.Level(None, shape, mg.HipProduct(shape, ke, torch.float64), E64, mask, torch.float64)
    if coarse_matrix is None:
        Ec, m = E64, mask
        for i in range(1, len(shapes)):
            Ec = mg.coarsen_modulus(None, Ec, shapes[i - 1])
            m = m.coarsen(shapes[i - 1])
        cshape = shapes[-1]
        kc = torch.from_numpy(np.ascontiguousarray(ke * 2 ** (len(shapes) - 1), dtype=np.float64)).to(dev)
        coarse_matrix = ops.assemble_dense(kc, Ec, cshape, mask=m)
    if smoother == "chebyshev":
        kw = dict(smoother="chebyshev", pre=degree, post=degree, lower_frac=lower_frac)
    else:
        kw = dict(pre=pre, post=post)
    kw.update(coarse_sweeps=coarse_degree, coarse_from=coarse_level, storage=storage)
    # (the coarsest factorisation stays on the current stream: run on a second stream beside the
    #  spectrum estimates, torch's rocSOLVER Cholesky made some solves differ from run to run)
    Lc = torch.linalg.cholesky(coarse_matrix)
    M = mg.Multigrid(None, shapes, prods, E64, mask, Lc, pdtype, coarse_dtype=coarse_dtype, **kw)
    # Jacobi: the reference's spectral omegas; Chebyshev: safe upper bounds of the spectrum
    if smoother == "chebyshev":
        M.lambdas = mg.chebyshev_bounds(None, M.levels, iterations=bound_iterations, start=bound_start,
                                        keep=bound_keep)
    else:
        M.omegas, M.lambdas = mg.spectral_omegas(None, M.levels, iterations=omega_iterations, seed=omega_seed)
    return fine, M


def galerkin_probes(fshape, apply, dev):
    """G (81 x 3nn_c, FP64): R A_f P v_p for the 81 probes v_p of the grid below `fshape`, with
    apply(f, y) the level's constrained operator y = A_f f.  P^T A_f P couples coarse nodes at
    most one apart, so 81 probes (the 27 node residues mod 3 times 3 components) give every
    column: probe p = sum of e_(J, c) over the nodes J of residue class p, pushed through
    prolongation, the level operator and restriction (P and R unmasked)."""
    cshape = tuple(int(s) // 2 for s in fshape)
    cx, cy, cz = (s + 1 for s in cshape)
    nn, nf = cx * cy * cz, 3 * ops.n_nodes(fshape)
    node = torch.arange(nn, device=dev)
    iz, iy, ix = node % cz, (node // cz) % cy, node // (cz * cy)
    cls = ((ix % 3) * 3 + iy % 3) * 3 + iz % 3
    G = torch.empty((81, 3 * nn), dtype=torch.float64, device=dev)
    v = torch.zeros((nn, 3), dtype=torch.float64, device=dev)
    f, y = (torch.empty(nf, dtype=torch.float64, device=dev) for _ in range(2))
    for p in range(81):
        v.zero_()
        v[cls == p // 3, p % 3] = 1.0
        ops.prolong(v.view(-1), f, fshape, assign=True)
        apply(f, y)
        ops.restrict(y, G[p], fshape)
    return G


def galerkin_coarsest(fshape, ke, Ef, mask):
    """Dense FP64 Galerkin operator M_c (P^T A_f P) M_c + (I - M_c) of the grid below `fshape`,
    on the device (multigrid_v4.py:248-270,747-772 build it with host sparse triple products).
    A_f = M K(Ef) M + (I - M) is the level's constrained operator (ke: that level's element
    matrix); its 81 probe products (galerkin_probes) are scattered into A_c by
    vh_galerkin_scatter."""
    dev = Ef.device
    cshape = tuple(int(s) // 2 for s in fshape)
    blocks = mg.HipProduct(fshape, ke, torch.float64).blocks
    mf = mask if isinstance(mask, ops.Mask) else ops.Mask.from_values(mask, fshape, dev)
    G = galerkin_probes(fshape, lambda f, y: ops.matvec(blocks, Ef, f, y, fshape, mask=mf), dev)
    n = 3 * ops.n_nodes(cshape)
    A = torch.zeros((n, n), dtype=torch.float64, device=dev)
    return ops.galerkin_scatter(G, A, cshape, mask=mf.coarsen(fshape))


def galerkin_hierarchy(shapes, ke, mask, E64):
    """Full Galerkin hierarchy A_l = M_l (P^T A_(l-1) P) M_l + (I - M_l) from the fine
    constrained operator A_0 = M K(E) M + (I - M), every operator assembled on the device
    (donor_yang_gmg multigrid_v4.py:132-270 for level 1 and :747-772 below, there as host-built
    CSR triple products): levels 1 .. L-2 as FP64 stencils (StencilProduct, from the 81 probe
    products of the level above through vh_stencil_scatter), the coarsest as the dense matrix of
    vh_galerkin_scatter.  Returns (stencil products of levels 1 .. L-2, dense coarsest)."""
    dev = E64.device
    masks = [mask if isinstance(mask, ops.Mask) else ops.Mask.from_values(mask, shapes[0], dev)]
    for i in range(1, len(shapes)):
        masks.append(masks[-1].coarsen(shapes[i - 1]))
    blocks = mg.HipProduct(shapes[0], ke, torch.float64).blocks
    apply = lambda f, y: ops.matvec(blocks, E64, f, y, shapes[0], mask=masks[0])  # noqa: E731
    stencils = []
    for lvl in range(1, len(shapes) - 1):
        G = galerkin_probes(shapes[lvl - 1], apply, dev)
        U = torch.empty(126 * ops.n_nodes(shapes[lvl]), dtype=torch.float64, device=dev)
        ops.stencil_scatter(G, U, shapes[lvl])
        del G
        sp = mg.StencilProduct(shapes[lvl], U)
        stencils.append(sp)
        apply = (lambda f, y, sp=sp, m=masks[lvl]: sp.raw(f, out=y, mask=m))  # noqa: E731
    G = galerkin_probes(shapes[-2], apply, dev)
    n = 3 * ops.n_nodes(shapes[-1])
    A = torch.zeros((n, n), dtype=torch.float64, device=dev)
    return stencils, ops.galerkin_scatter(G, A, shapes[-1], mask=masks[-1])
