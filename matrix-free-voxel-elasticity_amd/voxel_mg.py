"""Drop-in for the reference's src/voxel_mg.py on an MI355X: geometric multigrid-preconditioned
flexible CG for voxel elasticity, with every vector operation in HIP kernels.

Same names, signatures and return values as src/voxel_mg.py:27-244.  `xp` is `voxelhip.xp`
(the reference passes numpy or cupy); vectors are torch tensors on the HIP device.

  XYZ, edof_author(shape)                      host index helpers (unchanged meaning)
  HipProduct(shape, ke, dtype)                 the product plugin: raw(u, E), diagonal(E) on the
                                               GPU (the role of NumpyProduct / CustomProduct,
                                               src/voxel_mg.py:41-58, session_d/run_gpu.py:42-65)
  assemble_dense(shape, ke, E, mask)           dense constrained coarsest matrix (host setup,
                                               as the reference does before np.linalg.cholesky)
  prolong / restrict / coarsen_modulus / coarsen_mask   HIP transfer kernels
  Level, Multigrid, fcg, hierarchy_shapes, spectral_omegas

Extensions over the reference (all default to the reference's behaviour):
  Multigrid(..., smoother="chebyshev", lambdas=..., lower_frac=1/30, cheb_variant="saad")
      Chebyshev-Jacobi smoothing of degree pre/post (Saad Alg. 12.1; "donor" reproduces the
      recurrence of donor_yang_gmg multigrid_v4.py:860-889).
  The coarsest solve applies the explicit inverse of the coarsest operator (formed once on the
  device from the Cholesky factor) with one dense matrix-vector kernel, where the reference runs
  two triangular solves (src/voxel_mg.py:152-159).
  fcg keeps alpha, beta and r.z on the device and synchronises with the host only when it tests
  the true residual (every `check_every` iterations).
"""
from __future__ import annotations

import numpy as np
import torch

import voxelhip
from voxelhip import _lib as L
from voxelhip import ops
from voxel_kernels import Modal

XYZ = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])

_TD = {np.float64: torch.float64, np.float32: torch.float32,
       torch.float64: torch.float64, torch.float32: torch.float32}


def _tdt(dtype):
    if isinstance(dtype, torch.dtype):
        return _TD[dtype]
    return _TD[np.dtype(dtype).type]


def edof_author(shape):
    """(n_el, 24) author-layout DOFs in the native node order of ke (src/voxel_mg.py:32-38)."""
    nx, ny, nz = (int(s) for s in shape)
    ex, ey, ez = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    base = (ex.ravel() * (ny + 1) + ey.ravel()) * (nz + 1) + ez.ravel()
    off = (XYZ[:, 0] * (ny + 1) + XYZ[:, 1]) * (nz + 1) + XYZ[:, 2]
    nodes = base[:, None] + off[None, :]
    return (3 * nodes[:, :, None] + np.arange(3)).reshape(-1, 24)


# ------------------------------------------------------------------------------- product
class HipProduct:
    """y = K(E) u with the parity-block HIP kernel.  `raw(u, E)` expects constrained DOFs already
    zeroed (src/voxel_mg.py:42).  `storage="bf16"` stores u in bfloat16 (FP32 compute)."""

    def __init__(self, shape, ke, dtype=np.float64, storage=None):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = _tdt(dtype)
        self.ke = np.asarray(ke, dtype=np.float64).reshape(24, 24)
        # blocks from the FP64 ke, rounded to the product dtype inside the library
        self.blocks = Modal(self.ke).blocks
        d = np.diag(self.ke).reshape(8, 3)
        # per-component diagonal, when equal at every corner (axis-aligned bricks): D^-1 can then
        # be formed from the moduli inside the smoothing pass
        self.kdiag = d[0].copy() if np.allclose(d, d[0], rtol=1e-13, atol=0) else None
        self.ndof = 3 * ops.n_nodes(self.shape)
        self.storage = L.BF16 if storage == "bf16" else None
        hb = np.ascontiguousarray(np.asarray(self.blocks, dtype=np.float64).reshape(72))
        # voxel-cube coefficients: the fused load-time combinations are available
        self.cube = bool(L.load().vh_blocks_are_cube(hb.ctypes.data))

    def raw(self, u, E, out=None, mask=None, fixed_zero=False):
        y = out if out is not None else torch.empty(self.ndof, dtype=self.dtype, device=u.device)
        ops.matvec(self.blocks, E, u, y, self.shape, mask=mask, fixed_zero=fixed_zero, storage=self.storage)
        return y

    def diagonal(self, E, out=None, mask=None, invert=False):
        d = out if out is not None else torch.empty(self.ndof, dtype=self.dtype, device=E.device)
        return ops.diagonal(self.ke, E, d, self.shape, mask=mask, invert=invert)


class StencilProduct:
    """Assembled Galerkin operator K = P^T A_f P of one coarse level as the symmetric 27-point
    node-block stencil U of vh_stencil_scatter (14 x 9 x nn coefficients), applied by
    vh_stencil_apply; it takes the place of HipProduct on the non-coarsest levels of a Galerkin
    hierarchy (donor_yang_gmg multigrid_v4.py:132-270,747-772).  `raw` ignores E."""

    cube = False
    kdiag = None
    storage = None

    def __init__(self, shape, U):
        self.shape = tuple(int(s) for s in shape)
        self.ndof = 3 * ops.n_nodes(self.shape)
        if U.numel() != 42 * self.ndof:
            raise ValueError("stencil size does not match the grid")
        self.U = U.contiguous()
        self.dtype = U.dtype

    def astype(self, dtype):
        return self if _tdt(dtype) == self.dtype else StencilProduct(self.shape, self.U.to(_tdt(dtype)))

    def raw(self, u, E=None, out=None, mask=None, fixed_zero=False):
        y = out if out is not None else torch.empty(self.ndof, dtype=self.dtype, device=u.device)
        return ops.stencil_apply(self.U, u, y, self.shape, mask=mask, fixed_zero=fixed_zero)

    def diagonal(self, E=None, out=None, mask=None, invert=False):
        """Diagonal of K (1 at fixed DOFs under a mask), or its inverse."""
        nn = self.ndof // 3
        d = self.U.view(14, 9, nn)[0, [0, 4, 8], :].t().reshape(-1)
        if mask is not None and mask.kind != L.MASK_NONE:
            d = torch.where(mask.values(self.shape, self.dtype, d.device) != 0, d, torch.ones_like(d))
        d = 1.0 / d if invert else d.clone()
        if out is not None:
            out.copy_(d)
            return out
        return d


def assemble_dense(shape, ke, E, mask):
    """Dense M K M + (I - M) in FP64 on the host (src/voxel_mg.py:61-70); setup only."""
    import scipy.sparse as sp
    E = E.detach().cpu().numpy() if isinstance(E, torch.Tensor) else np.asarray(E)
    m = mask.detach().cpu().numpy() if isinstance(mask, torch.Tensor) else np.asarray(mask)
    m = m.astype(np.float64)
    edof = edof_author(shape)
    n = 3 * ops.n_nodes(shape)
    vals = (E[:, None, None] * np.asarray(ke, dtype=np.float64)[None]).ravel()
    rows = np.repeat(edof, 24, axis=1).ravel()
    cols = np.tile(edof, (1, 24)).ravel()
    K = sp.coo_matrix((vals, (rows, cols)), shape=(n, n)).toarray()
    return m[:, None] * K * m[None, :] + np.diag(1.0 - m)


# ------------------------------------------------------------------------------- transfers
def prolong(xp, uc, coarse_shape):
    nx, ny, nz = (int(s) for s in coarse_shape)
    fine = (2 * nx, 2 * ny, 2 * nz)
    out = torch.empty(3 * ops.n_nodes(fine), dtype=uc.dtype, device=uc.device)
    return ops.prolong(uc.contiguous(), out, fine, assign=True)


def restrict(xp, rf, fine_shape):
    nx, ny, nz = (int(s) for s in fine_shape)
    out = torch.empty(3 * ops.n_nodes((nx // 2, ny // 2, nz // 2)), dtype=rf.dtype, device=rf.device)
    return ops.restrict(rf.contiguous(), out, fine_shape)


def coarsen_modulus(xp, E, shape):
    nx, ny, nz = (int(s) for s in shape)
    out = torch.empty((nx // 2) * (ny // 2) * (nz // 2), dtype=E.dtype, device=E.device)
    return ops.coarsen_modulus(E.contiguous(), out, shape)


def coarsen_mask(xp, mask, shape):
    nx, ny, nz = (int(s) for s in shape)
    return mask.reshape(nx + 1, ny + 1, nz + 1, 3)[::2, ::2, ::2, :].contiguous().reshape(-1)


def hierarchy_shapes(shape, max_coarse_x=16):
    """Halve every direction until x <= max_coarse_x (all extents stay integers) (src/voxel_mg.py:220-225)."""
    shapes = [tuple(int(s) for s in shape)]
    while shapes[-1][0] > max_coarse_x and all(s % 2 == 0 for s in shapes[-1]):
        shapes.append(tuple(s // 2 for s in shapes[-1]))
    return shapes


# ------------------------------------------------------------------------------- hierarchy
class Level:
    """Constrained operator A u = M K M u + (I - M) u (src/voxel_mg.py:120-130)."""

    def __init__(self, xp, shape, product, E, mask, dtype):
        self.xp, self.shape, self.product = xp, tuple(int(s) for s in shape), product
        self.dtype = _tdt(dtype)
        dev = voxelhip.device()
        self.E = voxelhip.to_device(E, self.dtype)
        self.maskobj = mask if isinstance(mask, ops.Mask) else ops.Mask.from_values(mask, self.shape, dev)
        self.mask = self.maskobj.values(self.shape, self.dtype, dev)
        self.fixed = 1 - self.mask
        self.n = 3 * ops.n_nodes(self.shape)
        self.stencil = isinstance(product, StencilProduct)
        self.fused = hasattr(product, "blocks") or self.stencil
        if self.fused:
            self.inv_diag = product.diagonal(self.E, mask=self.maskobj, invert=True)
        else:
            d = product.diagonal(self.E)
            self.inv_diag = (1 / torch.where(self.mask != 0, d, torch.ones_like(d))).to(self.dtype)

    def apply(self, u, out=None):
        y = out if out is not None else torch.empty(self.n, dtype=self.dtype, device=u.device)
        if self.stencil:
            return self.product.raw(u, out=y, mask=self.maskobj)
        if self.fused:
            return ops.matvec(self.product.blocks, self.E, u, y, self.shape, mask=self.maskobj,
                              storage=self.product.storage)
        t = ops.axpby(1.0, u, 0.0, torch.empty_like(u), w=self.mask)
        y.copy_(self.product.raw(t, self.E))
        ops.axpby(1.0, y, 0.0, y, w=self.mask)
        return ops.axpby(1.0, u, 1.0, y, w=self.fixed)

    # fused passes used by the V-cycle and the Krylov loops
    def _plan(self, mode, combine):
        return getattr(self, "launch", {}).get((mode, bool(combine)))

    def tiles(self, mode=L.MODE_DOT, combine=False):
        """Number of per-tile partial sums a DOT (or RESID) pass of this level writes, under the
        launch plan that pass uses (tuned or automatic)."""
        if self.stencil:
            raise TypeError("a stencil level has no per-tile partial sums")
        t, lx = self._plan(mode, combine) or (0, 0)
        return ops.matvec_tiles(self.shape, lx=lx, tile=t, dtype=self.dtype, blocks=self.product.blocks,
                                storage=self.product.storage)[0]

    def matvec(self, u, y, **kw):
        if not self.fused:
            raise TypeError("fused passes need a HipProduct or a StencilProduct")
        if self.stencil:
            kw.pop("kdiag", None)
            return ops.stencil_apply(self.product.U, u, y, self.shape, mask=self.maskobj, **kw)
        if "tile" not in kw and "lx" not in kw:
            plan = self._plan(kw.get("mode", L.MODE_RAW), kw.get("w") is not None)
            if plan is not None:
                kw["tile"], kw["lx"] = plan
        storage = kw.pop("storage", self.product.storage)
        return ops.matvec(self.product.blocks, self.E, u, y, self.shape, mask=self.maskobj, storage=storage, **kw)

    def tune(self, tuner, passes):
        """Empirical launch plans for the given (mode, combine) pass kinds (voxelhip.tuner)."""
        if not self.fused or self.stencil:
            return
        self.launch = getattr(self, "launch", {})
        for mode, combine in passes:
            dinv_array = (mode in (L.MODE_JACOBI, L.MODE_CHEB, L.MODE_CHEB3, L.MODE_CHEB3D)
                          and self.product.kdiag is None)
            self.launch[(mode, bool(combine))] = tuner.tune(self, mode, combine, dinv_array)


def _cheb_params(lam, lower_frac):
    lmin = lower_frac * lam
    return 0.5 * (lam + lmin), 0.5 * (lam - lmin)


class Multigrid:
    """V-cycle preconditioner (src/voxel_mg.py:133-178).  `products` is a list of per-level
    product objects (finest first) exposing raw(u, E) and diagonal(E); `coarse_factor` is the
    lower Cholesky factor of the coarsest constrained operator (host or device), or None when the
    caller sets the coarsest inverse later (set_coarse_inverse, before the first V-cycle)."""

    def __init__(self, xp, shapes, products, E_fine, mask_fine, coarse_factor, dtype,
                 pre=4, post=4, omegas=None, smoother="jacobi", lambdas=None,
                 lower_frac=1.0 / 30.0, cheb_variant="saad", coarse_dtype=None, coarse_sweeps=None, coarse_from=2,
                 storage=None):
        self.xp, self.pre, self.post = xp, pre, post
        # storage="bf16" (FP32 V-cycle): the finest level's smoothing and residual passes read their
        # iterate through a bfloat16 copy, the BF16 storage of displacements (the donor's bf16
        # fine smoother, donor_yang_gmg multigrid_v4.py:851-856); arithmetic stays FP32
        if storage not in (None, "bf16"):
            raise ValueError(f"storage: {storage!r}")
        if storage == "bf16" and _tdt(dtype) != torch.float32:
            raise ValueError("bf16 storage of the iterate needs an FP32 V-cycle")
        self.storage = storage
        # levels >= coarse_from smooth with coarse_sweeps steps instead of pre / post (the donor's
        # separate coarse smoothing count, coarse_smooth_iters of multigrid_v4.py:466,934-967)
        self.coarse_sweeps, self.coarse_from = coarse_sweeps, int(coarse_from)
        self.dtype = _tdt(dtype)
        dev = voxelhip.device()
        self.levels = []
        E = voxelhip.to_device(E_fine, self.dtype)
        mask = mask_fine if isinstance(mask_fine, ops.Mask) else ops.Mask.from_values(mask_fine, shapes[0], dev)
        for i, (shape, prod) in enumerate(zip(shapes, products)):
            if i:
                E = coarsen_modulus(xp, E, shapes[i - 1])
                mask = mask.coarsen(shapes[i - 1])
            self.levels.append(Level(xp, shape, prod, E, mask, self.dtype))
        n_smooth = len(self.levels) - 1
        self.omegas = list(omegas) if omegas is not None else [0.6] * n_smooth
        assert len(self.omegas) == n_smooth
        self.smoother = smoother
        self.lambdas = list(lambdas) if lambdas is not None else None
        self.lower_frac = lower_frac
        if cheb_variant not in ("saad", "donor"):
            raise ValueError(cheb_variant)
        self.cheb_variant = cheb_variant
        self.dinv_on_the_fly = True
        # coarsest: explicit inverse A^-1 = L^-T L^-1 of the constrained operator, formed in FP64
        # and stored in the level precision or, with coarse_dtype=float64, in FP64 also inside an
        # FP32 V-cycle (rows padded to 16-byte multiples for the vector loads of dense_mv)
        self.coarse_dtype = self.dtype if coarse_dtype is None else _tdt(coarse_dtype)
        if self.coarse_dtype not in (self.dtype, torch.float64):
            raise ValueError("coarse_dtype must be the level precision or float64")
        self.coarse_inverse = None
        if coarse_factor is not None:
            Lf = coarse_factor if isinstance(coarse_factor, torch.Tensor) else torch.from_numpy(np.asarray(coarse_factor))
            self.set_coarse_inverse(torch.cholesky_inverse(Lf.to(dev, torch.float64)))
        cl = self.levels[-1]
        self._coarse_mask = ops.Mask(L.MASK_ARRAY, (cl.mask != 0).to(torch.uint8).contiguous())
        # work vectors per level: x, x2 (ping-pong), res, d (Chebyshev), rhs (levels > 0)
        self._work = []
        for lv in self.levels:
            z = lambda: torch.zeros(lv.n, dtype=self.dtype, device=dev)  # noqa: E731
            self._work.append({"x": z(), "x2": z(), "res": z(), "d": z(), "rhs": z()})
        if storage == "bf16":
            if not (self.levels[0].fused and not self.levels[0].stencil):
                raise ValueError("bf16 storage needs the parity product on the finest level")
            self._xb = torch.zeros(self.levels[0].n, dtype=torch.bfloat16, device=dev)

    def set_coarse_inverse(self, inverse64):
        """The coarsest level's explicit inverse (FP64 device matrix), stored for dense_mv."""
        self.coarse_inverse = ops.padded_matrix(inverse64, self.coarse_dtype)

    def tune(self, tuner):
        """Launch plans of every smoothing level's passes (smoother and residual)."""
        sm = L.MODE_CHEB3 if self.smoother == "chebyshev" else L.MODE_JACOBI
        for lv in self.levels[:-1]:
            passes = [(sm, False), (L.MODE_RESID, False)]
            if sm == L.MODE_CHEB3:
                passes.append((L.MODE_JACOBI, False))    # the first post-smoothing step (c = 0)
            if sm == L.MODE_CHEB3 and self.cheb_variant == "saad":
                passes.append((sm, True))          # the first step, fused with x0 = D^-1 r / theta
            lv.tune(tuner, passes)

    # ---------------------------------------------------------------- smoothing
    def _pass(self, level, u, y, **kw):
        """One fused pass of `level` on the iterate u (through its bfloat16 copy with storage="bf16"
        on the finest level)."""
        Lv = self.levels[level]
        if self.storage == "bf16" and level == 0 and kw.get("w") is None:
            ops.convert(u, self._xb)
            return Lv.matvec(self._xb, y, storage=L.BF16, **kw)
        return Lv.matvec(u, y, **kw)

    def _dinv_kw(self, Lv):
        """D^-1 formed from the moduli inside the pass when ke allows it, else the stored array."""
        kd = getattr(Lv.product, "kdiag", None)
        if self.dinv_on_the_fly and kd is not None:
            return {"kdiag": kd}
        return {"dinv": Lv.inv_diag}

    def _smooth(self, level, x, r, sweeps, first, out=None, dots=None):
        """Returns the buffer holding the smoothed x (x or its ping-pong partner).  With `out`
        (post-smoothing, Chebyshev) the last step writes there instead and returns it; with
        dots = (partial, zold) that step also forms the two Polak-Ribiere dots r . out and
        r . (out - zold) per tile (MODE_CHEB3D)."""
        Lv, w = self.levels[level], self._work[level]
        other = w["x2"] if x is w["x"] else w["x"]
        dk = self._dinv_kw(Lv)
        if out is not None and (first or self.smoother != "chebyshev"):
            raise ValueError("out= is for the Chebyshev post-smoothing")

        def last_pass(xin, partner, **kw):
            if dots is not None:
                self._pass(level, xin, out, mode=L.MODE_CHEB3D, b=r, prev=partner, partial=dots[0], zold=dots[1], **kw)
            else:
                self._pass(level, xin, out, mode=L.MODE_CHEB3, b=r, prev=partner, **kw)
            return out

        if self.smoother == "chebyshev":
            # three-term form x_{k+1} = x_k + a_k (z_k + c_k (x_k - x_{k-1})), d_k = x_{k+1} - x_k;
            # the pass overwrites x_{k-1} (held by the ping-pong partner) with x_{k+1}
            theta, delta = _cheb_params(self.lambdas[level], self.lower_frac)
            prev_zero = False
            alpha = 2.0 / theta if self.cheb_variant == "saad" else 1.0 / theta
            start = 1
            if first and Lv.product.cube and sweeps > 1 and self.cheb_variant == "saad":
                # x0 = 0: steps 1 and 2 in one pass, x1 = D^-1 r / theta formed at load time (and
                # written to x, the next step's previous iterate), x2 into the partner buffer
                c = 0.25 * delta * delta * alpha
                alpha = 1.0 / (theta - c)
                Lv.matvec(r, other, mode=L.MODE_CHEB3, b=r, cheb_a=alpha, cheb_c=c, prev_zero=True,
                          w=Lv.inv_diag, load_scale=1.0 / theta, u_out=x, **dk)
                x, other = other, x
                start = 2
            elif first:      # x0 = 0: x1 = D^-1 r / theta, and x_{k-1} = 0 at the next step
                ops.axpby(1.0 / theta, r, 0.0, x, w=Lv.inv_diag)
                prev_zero = True
            elif out is not None and sweeps == 1:
                return last_pass(x, other, cheb_a=1.0 / theta, cheb_c=0.0, **dk)
            else:
                # (c = 0: the step is the damped Jacobi sweep with omega = 1 / theta, the same
                #  arithmetic, and that pass does not read the previous iterate)
                self._pass(level, x, other, mode=L.MODE_JACOBI, b=r, omega=1.0 / theta, **dk)
                x, other = other, x
            for k in range(start, sweeps):
                c = 0.25 * delta * delta * alpha
                alpha = 1.0 / (theta - c)
                cd = c * theta if (self.cheb_variant == "donor" and k == 1) else c
                if out is not None and k == sweeps - 1:
                    return last_pass(x, other, cheb_a=alpha, cheb_c=cd, prev_zero=prev_zero, **dk)
                self._pass(level, x, other, mode=L.MODE_CHEB3, b=r, cheb_a=alpha, cheb_c=cd, prev_zero=prev_zero,
                           **dk)
                prev_zero = False
                x, other = other, x
            return x
        om = self.omegas[level]
        if first:
            ops.axpby(om, r, 0.0, x, w=Lv.inv_diag)
            sweeps -= 1
        for _ in range(sweeps):
            self._pass(level, x, other, mode=L.MODE_JACOBI, b=r, omega=om, **dk)
            x, other = other, x
        return x

    def _sweeps(self, level):
        """(pre, post) smoothing steps of a level: `pre`/`post`, or `coarse_sweeps` (when set) on
        the levels from `coarse_from` down."""
        if self.coarse_sweeps is not None and level >= self.coarse_from:
            return self.coarse_sweeps, self.coarse_sweeps
        return self.pre, self.post

    def coarse_solve(self, r, out=None):
        y = out if out is not None else torch.empty_like(r)
        return ops.dense_mv(self.coarse_inverse, r, y, mask=self._coarse_mask, n=r.numel())

    def vcycle(self, r, level=0, out=None, dots=None):
        """One V-cycle on r.  out / dots (finest level, Chebyshev smoother): the result is
        written to `out` by the last smoothing step, which also forms the dots of _smooth."""
        Lv, w = self.levels[level], self._work[level]
        if level == len(self.levels) - 1:
            return self.coarse_solve(r, w["x"] if out is None else out)
        pre, post = self._sweeps(level)
        x = self._smooth(level, w["x"], r, pre, True)
        self._pass(level, x, w["res"], mode=L.MODE_RESID, b=r, fixed_zero=True)      # (r - A x) M
        nxt = self.levels[level + 1]
        rc = self._work[level + 1]["rhs"]
        ops.restrict(w["res"], rc, Lv.shape, cmask=nxt.maskobj)
        ec = self.vcycle(rc, level + 1)
        ops.prolong(ec, x, Lv.shape, mask=Lv.maskobj)                          # x += P ec M
        return self._smooth(level, x, r, post, False, out=out, dots=dots)

    def __call__(self, r, out=None):
        w0 = self._work[0]
        if r.dtype != self.dtype:
            ops.convert(r, w0["rhs"])
            rin = w0["rhs"]
        else:
            rin = r
        x = self.vcycle(rin)
        y = out if out is not None else torch.empty(r.numel(), dtype=r.dtype, device=r.device)
        if y.dtype != self.dtype:
            ops.convert(x, y)
        else:
            y.copy_(x)
        return y


# ------------------------------------------------------------------------------- Krylov
class _Reducer:
    """Device scalars (FP64) and partial-sum buffers for one vector length."""

    def __init__(self, n, dev, extra_parts=0):
        self.n = n
        self.parts = ops.reduce_parts(n)
        self.partial = torch.zeros(2 * max(self.parts, extra_parts), dtype=torch.float64, device=dev)
        self.scal = torch.zeros(16, dtype=torch.float64, device=dev)

    def dot(self, x, y, dst, op=L.OP_NONE):
        ops.reduce(L.RED_DOT, x, y, partial=self.partial)
        ops.finish(self.partial, self.parts, 1, self.scal, dst, op)

    def norm2(self, x, dst):
        ops.reduce(L.RED_NORM2, x, partial=self.partial)
        ops.finish(self.partial, self.parts, 1, self.scal, dst)


def fcg(xp, apply_A, b, precondition, tol, maxiter, check_every=5, fixed_iterations=None):
    """Flexible PCG (Polak-Ribiere beta), FP64 outer; stops on the TRUE relative residual
    (src/voxel_mg.py:181-217).  Returns (x, iterations, converged, true_relative_residual,
    outer_matvecs).  `apply_A(p)` / `precondition(r)` return device tensors; when apply_A is a
    Level's bound `apply`, q = A p and p.q come from one fused pass."""
    dev = b.device
    n = b.numel()
    level = getattr(apply_A, "__self__", None)
    fused = isinstance(level, Level) and level.fused and apply_A.__func__ is Level.apply
    red = _Reducer(n, dev, level.tiles() if fused else 0)
    S_NORM, S_NB = 8, 9
    red.norm2(b, S_NB)
    nb = float(red.scal[S_NB].sqrt())
    x = torch.zeros_like(b)
    r = b.clone()
    z = precondition(r)
    p = z.clone()
    q = torch.empty_like(b)
    true = torch.empty_like(b)
    red.dot(r, z, L.S_RZ)
    matvecs = 0
    limit = fixed_iterations if fixed_iterations is not None else maxiter
    it = 0
    for it in range(1, limit + 1):
        if fused:
            nt = level.tiles()
            level.matvec(p, q, mode=L.MODE_DOT, partial=red.partial)
            ops.finish(red.partial, nt, 1, red.scal, L.S_PQ, L.OP_ALPHA)
        else:
            q.copy_(apply_A(p))
            red.dot(p, q, L.S_PQ, L.OP_ALPHA)
        matvecs += 1
        ops.cg_update(x, p, red.scal, r=r, q=q)            # x += a p ; r -= a q
        if fixed_iterations is None and it % check_every == 0:
            if fused:
                level.matvec(x, true, mode=L.MODE_RESID, b=b)
            else:
                true.copy_(b - apply_A(x))
            matvecs += 1
            red.norm2(true, S_NORM)
            rel = float(red.scal[S_NORM].sqrt()) / nb
            if rel <= tol:
                return x, it, True, rel, matvecs
            r, true = true, r                              # residual replacement
        z_new = precondition(r)
        ops.reduce(L.RED_DOTDIFF, r, z_new, z, partial=red.partial)   # r.z_new, r.(z_new - z)
        ops.finish(red.partial, red.parts, 2, red.scal, 5, L.OP_BETA_PR)
        ops.xpby(z_new, p, red.scal)                       # p = z_new + beta p
        z = z_new
    if fused:
        level.matvec(x, true, mode=L.MODE_RESID, b=b)
    else:
        true.copy_(b - apply_A(x))
    matvecs += 1
    red.norm2(true, S_NORM)
    rel = float(red.scal[S_NORM].sqrt()) / nb
    return x, it, bool(rel <= tol), rel, matvecs


_STARTS = {}


def _start_vector(n, seed, dtype):
    """default_rng(seed).standard_normal(n) on the device (src/voxel_mg.py:234-235), kept per
    (n, seed, dtype): drawing tens of millions of normals on the host costs more than the power
    iteration itself, and the set-up of every SIMP step asks for the same vectors again."""
    key = (int(n), int(seed), dtype)
    v = _STARTS.get(key)
    if v is None:
        rng = np.random.default_rng(seed)
        v = voxelhip.to_device(rng.standard_normal(int(n)), torch.float64).to(dtype)
        _STARTS[key] = v
    return v


_BOUND_STARTS = {}


def _bound_start(n, seed, dtype):
    """The start vector of the Chebyshev bound estimate: uniform on [-1, 1) from SplitMix64
    (ops.fill_splitmix), drawn on the device and kept per (n, seed, dtype)."""
    key = (int(n), int(seed), dtype)
    v = _BOUND_STARTS.get(key)
    if v is None:
        v = ops.fill_splitmix(torch.empty(int(n), dtype=torch.float64, device=voxelhip.device()), seed).to(dtype)
        _BOUND_STARTS[key] = v
    return v


def spectral_omegas(xp, levels, iterations=30, seed=20260923):
    """omega_l = 4 / (3 lambda_l) for every non-coarsest level, lambda_l from a seeded power
    iteration on D^-1 A restricted to free DOFs (src/voxel_mg.py:228-244).  Returns (omegas, lambdas)."""
    omegas, lambdas = [], []
    for Lv in levels[:-1]:
        v = _start_vector(Lv.n, seed, Lv.dtype).clone()
        v = ops.axpby(1.0, v, 0.0, v, w=Lv.mask)
        red = _Reducer(Lv.n, v.device)
        red.norm2(v, 8)
        ops.axpby(1.0 / float(red.scal[8].sqrt()), v, 0.0, v)
        w = torch.empty_like(v)
        lam = 0.0
        for _ in range(iterations):
            Lv.matvec(v, w, fixed_zero=True)              # A v M = M K M v
            ops.axpby(1.0, w, 0.0, w, w=Lv.inv_diag)
            red.norm2(w, 8)
            lam = float(red.scal[8].sqrt())
            ops.axpby(1.0 / lam, w, 0.0, v)
        lambdas.append(lam)
        omegas.append(4.0 / (3.0 * lam))
    return omegas, lambdas


def chebyshev_bounds(xp, levels, iterations=30, safety=1.1, seed=20260923, start=None, keep=None):
    """Upper bounds of the spectrum of D^-1 A per smoothing level for the Chebyshev smoother:
    safety x the D-weighted Rayleigh quotient v.Av / v.Dv after `iterations` power steps
    v <- D^-1 A v / |D^-1 A v|_D (donor_yang_gmg multigrid_v4.py:273-295), started from a seeded
    uniform vector over the free DOFs (SplitMix64 on the device, _bound_start; the donor's all-ones
    start is a smooth vector, far from the top of the spectrum).  (spectral_omegas' norm ratio
    approaches lambda_max from below slowly on smooth moduli - 30 steps give 2.81 for 3.09 on a
    uniform 256 x 128 x 128 grid - and a Chebyshev interval that ends below the spectrum
    amplifies the top modes; the Rayleigh quotient converges twice as fast and the 10 % margin
    covers the rest.)  The normalisation stays on the device: one host read per level.
    start: per-level vectors to start from instead of the seeded one (a previous call's `keep`,
    for an operator that changed little); keep: a list that receives each level's last iterate."""
    lams = []
    for li, Lv in enumerate(levels[:-1]):
        d = Lv.product.diagonal(Lv.E, mask=Lv.maskobj)               # D, 1 at fixed DOFs
        w = (start[li] if start is not None else _bound_start(Lv.n, seed, Lv.dtype)).clone()
        w = ops.axpby(1.0, w, 0.0, w, w=Lv.mask)
        red = _Reducer(Lv.n, w.device)
        v, t = torch.empty_like(w), torch.empty_like(w)
        for k in range(iterations + 1):
            if k:                                                     # w = D^-1 M K M v
                Lv.matvec(v, w, fixed_zero=True)
                ops.axpby(1.0, w, 0.0, w, w=Lv.inv_diag)
            ops.axpby(1.0, w, 0.0, t, w=d)
            red.dot(w, t, 8, L.OP_INVSQRT)                            # 1 / |w|_D
            ops.convert(w, v, dscale=red.scal[8:9])
        Lv.matvec(v, w, fixed_zero=True)                              # Rayleigh quotient
        red.dot(v, w, 8)
        ops.axpby(1.0, v, 0.0, t, w=d)
        red.dot(v, t, 9)
        s89 = red.scal[8:10].cpu()
        lams.append(safety * float(s89[0]) / float(s89[1]))
        if keep is not None:
            keep.append(v)
    return lams
