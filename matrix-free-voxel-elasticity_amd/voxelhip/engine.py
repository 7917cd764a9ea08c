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
                  transfers, coarsest solve in FP32), FP64 dot accumulation
  "mixed-bf16"    as "mixed" with the fine-level FP32 products reading u stored in bfloat16
                  (storage ablation; the smoother iterate is rounded to bf16 before each product)
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
                 lower_frac=1.0 / 30.0, omega_iterations=30):
        self.shape = tuple(int(s) for s in shape)
        self.ke = np.asarray(ke, dtype=np.float64)
        self.dev = torch.device("cuda", torch.cuda.current_device())
        shapes = mg.hierarchy_shapes(self.shape, max_coarse_x)
        if n_levels is not None:
            shapes = shapes[:n_levels]
        self.shapes = shapes
        self.smoother, self.degree, self.pre, self.post = smoother, degree, pre, post
        if precision not in ("fp64", "mixed", "mixed-bf16"):
            raise ValueError(precision)
        self.precision = precision
        self.pdtype = torch.float64 if precision == "fp64" else torch.float32
        self.coarse = coarse
        self.graph = graph
        self.lower_frac = lower_frac
        self.omega_iterations = omega_iterations
        self.mask = mask if isinstance(mask, ops.Mask) else ops.Mask.from_values(mask, self.shape, self.dev)
        self.fine = None
        self.M = None
        self._graphs = None
        if E is not None:
            self.update(E)

    # ------------------------------------------------------------------ setup per modulus field
    def update(self, E):
        """(Re)build every level for a new modulus field (device or host array)."""
        E64 = E.to(self.dev, torch.float64).contiguous() if isinstance(E, torch.Tensor) else \
            torch.from_numpy(np.ascontiguousarray(E, dtype=np.float64)).to(self.dev)
        prods = [mg.HipProduct(s, self.ke * 2 ** i, self.pdtype,
                               storage="bf16" if (i == 0 and self.precision == "mixed-bf16") else None)
                 for i, s in enumerate(self.shapes)]
        self.fine = mg.Level(None, self.shape, mg.HipProduct(self.shape, self.ke, torch.float64), E64,
                             self.mask, torch.float64)
        # coarsest operator: rediscretised (ke * 2^L on the averaged modulus), or Galerkin
        Ec = E64
        masks = [self.mask]
        for i in range(1, len(self.shapes)):
            Ec = mg.coarsen_modulus(None, Ec, self.shapes[i - 1])
            masks.append(masks[-1].coarsen(self.shapes[i - 1]))
        cshape = self.shapes[-1]
        cmask_vals = masks[-1].values(cshape, torch.float64, self.dev)
        if self.coarse == "galerkin" and len(self.shapes) > 1:
            A = self._galerkin_coarsest(E64)
        else:
            A = mg.assemble_dense(cshape, self.ke * 2 ** (len(self.shapes) - 1), Ec, cmask_vals)
        Lc = torch.linalg.cholesky(torch.from_numpy(A).to(self.dev))
        kw = {}
        if self.smoother == "chebyshev":
            kw = dict(smoother="chebyshev", pre=self.degree, post=self.degree, lower_frac=self.lower_frac)
        else:
            kw = dict(pre=self.pre, post=self.post)
        self.M = mg.Multigrid(None, self.shapes, prods, E64, self.mask, Lc, self.pdtype, **kw)
        om, lam = mg.spectral_omegas(None, self.M.levels, iterations=self.omega_iterations)
        self.M.omegas = om
        self.M.lambdas = lam
        self._alloc()
        self._graphs = None

    def _galerkin_coarsest(self, E64):
        """A_c = P^T A_{L-2} P with the coarse mask re-imposed (host sparse products, setup)."""
        import scipy.sparse as sp
        fshape, cshape = self.shapes[-2], self.shapes[-1]
        lvl = len(self.shapes) - 2
        Ef = E64
        m = self.mask
        for i in range(1, lvl + 1):
            Ef = mg.coarsen_modulus(None, Ef, self.shapes[i - 1])
            m = m.coarsen(self.shapes[i - 1])
        Ef = Ef.cpu().numpy()
        mf = m.values(fshape, torch.float64, self.dev).cpu().numpy()
        edof = mg.edof_author(fshape)
        nf = 3 * ops.n_nodes(fshape)
        ke = self.ke * 2 ** lvl
        K = sp.coo_matrix(((Ef[:, None, None] * ke[None]).ravel(),
                           (np.repeat(edof, 24, axis=1).ravel(), np.tile(edof, (1, 24)).ravel())),
                          shape=(nf, nf)).tocsr()
        Mf = sp.diags(mf)
        Af = Mf @ K @ Mf + sp.diags(1.0 - mf)
        P = _prolongation_matrix(cshape)
        Ac = (P.T @ Af @ P).toarray()
        mc = m.coarsen(fshape).values(cshape, torch.float64, self.dev).cpu().numpy()
        return mc[:, None] * Ac * mc[None, :] + np.diag(1.0 - mc)

    def _alloc(self):
        n = self.fine.n
        z = lambda: torch.zeros(n, dtype=torch.float64, device=self.dev)  # noqa: E731
        self.b, self.x, self.r, self.p, self.q = z(), z(), z(), z(), z()
        self.za, self.zb = z(), z()
        self.parts = ops.reduce_parts(n)
        self.tiles = self.fine.tiles()
        self.partial = torch.zeros(2 * max(self.parts, self.tiles), dtype=torch.float64, device=self.dev)
        self.scal = torch.zeros(16, dtype=torch.float64, device=self.dev)

    # ------------------------------------------------------------------ one iteration
    def _iteration(self, z_new, z_old, check):
        """Launch one flexible-CG iteration (no host synchronisation)."""
        f = self.fine
        f.matvec(self.p, self.q, mode=L.MODE_DOT, partial=self.partial)          # q = A p, p.q
        ops.finish(self.partial, self.tiles, 1, self.scal, L.S_PQ, L.OP_ALPHA)
        if check:
            ops.cg_update(self.x, self.p, self.scal)                              # x += a p
            f.matvec(self.x, self.r, mode=L.MODE_RESID, b=self.b)                # r = b - A x
            ops.reduce(L.RED_NORM2, self.r, partial=self.partial)
            ops.finish(self.partial, self.parts, 1, self.scal, 8)
        else:
            ops.cg_update(self.x, self.p, self.scal, r=self.r, q=self.q)          # + r -= a q
        self.M(self.r, out=z_new)                                                 # z = M r
        ops.reduce(L.RED_DOTDIFF, self.r, z_new, z_old, partial=self.partial)
        ops.finish(self.partial, self.parts, 2, self.scal, 5, L.OP_BETA_PR)
        ops.xpby(z_new, self.p, self.scal)                                        # p = z + beta p

    def _capture(self, check):
        graphs = []
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):          # warm-up on the capture stream
            self._iteration(self.za, self.zb, check)
        torch.cuda.current_stream().wait_stream(s)
        for zn, zo in ((self.za, self.zb), (self.zb, self.za)):
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._iteration(zn, zo, check)
            graphs.append(g)
        return graphs

    # ------------------------------------------------------------------ solve
    def solve(self, b, tol=1e-8, maxiter=200, check_every=1, x0=None):
        """Flexible CG to ||b - A x|| / ||b|| <= tol.  Returns (x, iterations, converged, rel, matvecs).
        The returned x is a view of the solver's buffer (copy it to keep it across solves)."""
        f = self.fine
        self.b.copy_(b)
        ops.reduce(L.RED_NORM2, self.b, partial=self.partial)
        ops.finish(self.partial, self.parts, 1, self.scal, 9)
        nb = float(self.scal[9].sqrt())
        matvecs = 0
        if x0 is None:
            self.x.zero_()
            self.r.copy_(self.b)
        else:
            self.x.copy_(x0)
            f.matvec(self.x, self.r, mode=L.MODE_RESID, b=self.b)
            matvecs += 1
        self.M(self.r, out=self.zb)
        self.p.copy_(self.zb)
        ops.reduce(L.RED_DOT, self.r, self.zb, partial=self.partial)
        ops.finish(self.partial, self.parts, 1, self.scal, L.S_RZ)
        if self.graph and self._graphs is None and check_every == 1:
            # capture clobbers x / r / p through its warm-up: snapshot and restore them
            saved = [t.clone() for t in (self.x, self.r, self.p, self.zb, self.scal)]
            self._graphs = self._capture(True)
            for t, s in zip((self.x, self.r, self.p, self.zb, self.scal), saved):
                t.copy_(s)
        it, rel = 0, float("nan")
        zpair = [(self.za, self.zb), (self.zb, self.za)]
        for it in range(1, maxiter + 1):
            check = it % check_every == 0
            k = (it - 1) % 2
            if self._graphs is not None and check:
                self._graphs[k].replay()
            else:
                self._iteration(*zpair[k], check)
            matvecs += 2 if check else 1
            if check:
                rel = float(self.scal[8].sqrt()) / nb
                if rel <= tol:
                    return self.x, it, True, rel, matvecs
        f.matvec(self.x, self.r, mode=L.MODE_RESID, b=self.b)
        matvecs += 1
        ops.reduce(L.RED_NORM2, self.r, partial=self.partial)
        ops.finish(self.partial, self.parts, 1, self.scal, 8)
        rel = float(self.scal[8].sqrt()) / nb
        return self.x, it, bool(rel <= tol), rel, matvecs


def _prolongation_matrix(coarse):
    """Sparse trilinear P (fine x coarse) in the author AoS layout (src/voxel_mg.py:74-97)."""
    import scipy.sparse as sp
    nx, ny, nz = coarse
    def axis(nc):
        rows, cols, vals = [], [], []
        for i in range(2 * nc + 1):
            if i % 2 == 0:
                rows.append(i); cols.append(i // 2); vals.append(1.0)
            else:
                rows += [i, i]; cols += [i // 2, i // 2 + 1]; vals += [0.5, 0.5]
        return sp.csr_matrix((vals, (rows, cols)), shape=(2 * nc + 1, nc + 1))
    P = sp.kron(sp.kron(sp.kron(axis(nx), axis(ny)), axis(nz)), sp.identity(3))
    return P.tocsr()
