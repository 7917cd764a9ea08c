"""CPU oracle for the geometric multigrid and the flexible CG of the voxel solver.

TEST INFRASTRUCTURE ONLY (tests/, __graft_entry__.smoke(), bench.py cpu_baseline).

Restates src/voxel_mg.py:74-244 (transfers, coarsening, Level, Multigrid V-cycle with damped
Jacobi, flexible PCG, hierarchy shapes, spectral damping) and the Chebyshev-Jacobi smoother of
experiments/combined/donor_yang_gmg/src/gpu_fem/multigrid_v4.py:860-889 (Saad Alg. 12.1 with
M^-1 = D^-1 on [lam_max * lower_frac, lam_max]).

Parity status:
  prolong/restrict   P trilinear (interpolation of a linear field is exact), R == P^T,
                     P^T K_f P == K_c for uniform E (tests/test_mg_cpu.py:53-67)
  galerkin_dense     equals the rediscretised operator for uniform E, with and without the
                     clamped face; symmetric; fixed rows are identity rows
  galerkin_hierarchy level 1 equals galerkin_dense; for uniform E and the clamp every level
                     equals the rediscretised constrained operator; symmetric positive definite
                     for random moduli and masks
  Multigrid/fcg      reference-held known answers (smoke-q4 iterations, compliance, solution
                     npz of results/session-d-20260923), golden fixtures of one V-cycle and of
                     spectral_omegas written from the reference in this container
  chebyshev          reduces to damped Jacobi at degree 1 with omega = 2/(lmax+lmin); the
                     error polynomial is checked against the closed-form Chebyshev polynomial
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import solve_triangular

from .hexref import XYZ, NumpyProduct, assemble_dense  # noqa: F401  (re-exported names)


# ------------------------------------------------------------------------- transfers
def _prolong_axis(v, axis):
    v = np.moveaxis(v, axis, 0)
    nc = v.shape[0]
    out = np.empty((2 * nc - 1,) + v.shape[1:], dtype=v.dtype)
    out[0::2] = v
    out[1::2] = 0.5 * (v[:-1] + v[1:])
    return np.moveaxis(out, 0, axis)


def _restrict_axis(f, axis):
    f = np.moveaxis(f, axis, 0)
    out = f[0::2].copy()
    half = 0.5 * f[1::2]
    out[:-1] += half
    out[1:] += half
    return np.moveaxis(out, 0, axis)


def prolong(uc, coarse_shape):
    """Trilinear prolongation of a nodal AoS vector (src/voxel_mg.py:92-97)."""
    nx, ny, nz = coarse_shape
    v = np.asarray(uc).reshape(nx + 1, ny + 1, nz + 1, 3)
    for axis in range(3):
        v = _prolong_axis(v, axis)
    return np.ascontiguousarray(v).ravel()


def restrict(rf, fine_shape):
    """R = P^T applied axis by axis (src/voxel_mg.py:100-105)."""
    nx, ny, nz = fine_shape
    v = np.asarray(rf).reshape(nx + 1, ny + 1, nz + 1, 3)
    for axis in range(3):
        v = _restrict_axis(v, axis)
    return np.ascontiguousarray(v).ravel()


def coarsen_modulus(E, shape):
    """2x2x2 mean of the element moduli (src/voxel_mg.py:108-111)."""
    nx, ny, nz = shape
    return np.ascontiguousarray(
        np.asarray(E).reshape(nx // 2, 2, ny // 2, 2, nz // 2, 2).mean(axis=(1, 3, 5))).ravel()


def prolongation_matrix(coarse_shape):
    """Dense P (fine x coarse): column j is prolong(e_j) (src/voxel_mg.py:92-97)."""
    n = 3 * int(np.prod([s + 1 for s in coarse_shape]))
    return np.stack([prolong(np.eye(n)[j], coarse_shape) for j in range(n)], axis=1)


def galerkin_dense(fine_shape, ke, E, mask):
    """Galerkin coarse operator M_c (P^T A_f P) M_c + (I - M_c), A_f = assemble_dense(fine)
    (multigrid_v4.py:248-270,747-772: K_c = R K_f P with R = P^T, the boundary re-imposed by
    injection, solver_v2.py:716-763)."""
    coarse = tuple(int(s) // 2 for s in fine_shape)
    Af = assemble_dense(fine_shape, ke, E, mask)
    P = prolongation_matrix(coarse)
    Ac = P.T @ Af @ P
    mc = coarsen_mask(mask, fine_shape).astype(np.float64)
    return mc[:, None] * Ac * mc[None, :] + np.diag(1.0 - mc)


def galerkin_hierarchy(shapes, ke, E, mask):
    """Full Galerkin hierarchy (multigrid_v4.py:132-270 level 1, :747-772 deeper levels, there
    R K P in CSR): A_0 = M K(E) M + (I - M) (assemble_dense), K_l = P^T A_(l-1) P and
    A_l = M_l K_l M_l + (I - M_l) with the injected masks.  Dense; small grids only.
    Returns (K_1 .. K_(L-1) unmasked, A_0 .. A_(L-1) constrained)."""
    A = [assemble_dense(shapes[0], ke, E, mask)]
    K, m = [], np.asarray(mask, dtype=np.float64)
    for lvl in range(1, len(shapes)):
        P = prolongation_matrix(shapes[lvl])
        K.append(P.T @ A[-1] @ P)
        m = coarsen_mask(m, shapes[lvl - 1]).astype(np.float64)
        A.append(m[:, None] * K[-1] * m[None, :] + np.diag(1.0 - m))
    return K, A


class DenseProduct:
    """An assembled (unmasked) level operator K in the product interface of Level:
    raw(u, E) = K u and diagonal(E) = diag K (E is not used)."""

    def __init__(self, K, dtype=np.float64):
        self.K, self.dtype = np.asarray(K, dtype=np.float64), dtype

    def raw(self, u, E):
        return (self.K @ np.asarray(u, dtype=np.float64)).astype(self.dtype)

    def diagonal(self, E):
        return np.diag(self.K).astype(self.dtype)


def coarsen_mask(mask, shape):
    """Injection of the nodal mask at even nodes (src/voxel_mg.py:114-116)."""
    nx, ny, nz = shape
    return np.ascontiguousarray(
        np.asarray(mask).reshape(nx + 1, ny + 1, nz + 1, 3)[::2, ::2, ::2, :]).ravel()


def hierarchy_shapes(shape, max_coarse_x=16):
    """Halve every extent while x > max_coarse_x and all extents are even (src/voxel_mg.py:220-225)."""
    shapes = [tuple(int(s) for s in shape)]
    while shapes[-1][0] > max_coarse_x and all(s % 2 == 0 for s in shapes[-1]):
        shapes.append(tuple(s // 2 for s in shapes[-1]))
    return shapes


# ------------------------------------------------------------------------- hierarchy
class Level:
    """Constrained operator A u = M K M u + (I - M) u and its inverse diagonal (src/voxel_mg.py:120-130)."""

    def __init__(self, shape, product, E, mask, dtype):
        self.shape, self.product, self.dtype = tuple(shape), product, dtype
        self.E = np.asarray(E).astype(dtype)
        self.mask = np.asarray(mask).astype(dtype)
        self.fixed = (1 - self.mask).astype(dtype)
        d = product.diagonal(self.E)
        self.inv_diag = (1 / np.where(self.mask != 0, d, dtype(1))).astype(dtype)

    def apply(self, u):
        return self.product.raw(u * self.mask, self.E) * self.mask + u * self.fixed


def chebyshev(apply_A, rhs, x0, inv_diag, lam_max, lower_frac, degree, variant="saad"):
    """Chebyshev-Jacobi smoothing with M^-1 = D^-1 on [lower_frac*lam_max, lam_max].

    variant="saad": Saad, Iterative Methods 2nd ed., Alg. 12.1, written with a_k = 2 rho_k/delta:
        d_0 = z_0/theta, a_0 = 2/theta;  a_k = 1/(theta - delta^2 a_{k-1}/4),
        d_k = a_k (z_k + delta^2 a_{k-1}/4 d_{k-1});  x_{k+1} = x_k + d_k.
      The error after `degree` steps is T_d((theta - lam)/delta)/T_d(theta/delta) (tested).
    variant="donor": the recurrence as written in
      experiments/combined/donor_yang_gmg/src/gpu_fem/multigrid_v4.py:860-889 (first direction
      z_0 unscaled and a_0 = 1/theta), which is a different degree-d polynomial.
    """
    lam_min = lower_frac * lam_max
    theta = 0.5 * (lam_max + lam_min)
    delta = 0.5 * (lam_max - lam_min)
    x = x0
    z = inv_diag * (rhs - apply_A(x))
    if variant == "saad":
        d, alpha = z / theta, 2.0 / theta
        x = x + d
    elif variant == "donor":
        d, alpha = z, 1.0 / theta
        x = x + alpha * d
    else:
        raise ValueError(variant)
    for _ in range(1, degree):
        z = inv_diag * (rhs - apply_A(x))
        c = 0.25 * delta * delta * alpha
        alpha = 1.0 / (theta - c)
        d = alpha * (z + c * d)
        x = x + d
    return x


class Multigrid:
    """V-cycle preconditioner (src/voxel_mg.py:133-178).

    smoother="jacobi": the reference's damped Jacobi, `pre`/`post` sweeps with omegas[level].
    smoother="chebyshev": one Chebyshev-Jacobi polynomial of degree `pre`/`post` on
    [lam/30, lam] per level, lam = lambdas[level] (multigrid_v4.py:860-889 with the fine
    lambda used on every level).
    """

    def __init__(self, shapes, products, E_fine, mask_fine, coarse_factor, dtype,
                 pre=4, post=4, omegas=None, smoother="jacobi", lambdas=None,
                 lower_frac=1.0 / 30.0, cheb_variant="saad"):
        self.dtype, self.pre, self.post = dtype, pre, post
        self.levels = []
        E, mask = E_fine, mask_fine
        for i, (shape, prod) in enumerate(zip(shapes, products)):
            if i:
                E = coarsen_modulus(E, shapes[i - 1])
                mask = coarsen_mask(mask, shapes[i - 1])
            self.levels.append(Level(shape, prod, E, mask, dtype))
        self.coarse_factor = np.asarray(coarse_factor).astype(dtype)
        n_smooth = len(self.levels) - 1
        self.omegas = list(omegas) if omegas is not None else [0.6] * n_smooth
        assert len(self.omegas) == n_smooth
        self.smoother = smoother
        self.lambdas = list(lambdas) if lambdas is not None else None
        self.lower_frac = lower_frac
        self.cheb_variant = cheb_variant

    def coarse_solve(self, r):
        y = solve_triangular(self.coarse_factor, r, lower=True)
        return solve_triangular(self.coarse_factor.T, y, lower=False)

    def _smooth(self, level, x, r, sweeps, first):
        L = self.levels[level]
        if self.smoother == "chebyshev":
            lam = self.lambdas[level]
            x0 = np.zeros_like(r) if first else x
            return chebyshev(L.apply, r, x0, L.inv_diag, lam, self.lower_frac, sweeps,
                             self.cheb_variant)
        w = self.omegas[level]
        if first:
            x = w * L.inv_diag * r
            sweeps -= 1
        for _ in range(sweeps):
            x = x + w * L.inv_diag * (r - L.apply(x))
        return x

    def vcycle(self, r, level=0):
        L = self.levels[level]
        if level == len(self.levels) - 1:
            return self.coarse_solve(r) * L.mask
        pre, post = self.pre, self.post
        cs = getattr(self, "coarse_sweeps", None)    # the donor's coarse smoothing count
        if cs is not None and level >= getattr(self, "coarse_from", 2):
            pre = post = cs
        x = self._smooth(level, None, r, pre, True)
        res = (r - L.apply(x)) * L.mask
        rc = restrict(res, L.shape) * self.levels[level + 1].mask
        ec = self.vcycle(rc, level + 1)
        x = x + prolong(ec, self.levels[level + 1].shape) * L.mask
        x = self._smooth(level, x, r, post, False)
        return x

    def __call__(self, r):
        return self.vcycle(np.asarray(r).astype(self.dtype)).astype(np.asarray(r).dtype)


def fcg(apply_A, b, precondition, tol, maxiter, check_every=5, fixed_iterations=None):
    """Flexible PCG with Polak-Ribiere beta and true-residual stopping/replacement
    (src/voxel_mg.py:181-217). Returns (x, iterations, converged, rel, outer_matvecs)."""
    nb = float(np.linalg.norm(b))
    x = np.zeros_like(b)
    r = b.copy()
    z = precondition(r)
    p = z.copy()
    rz = float(r @ z)
    matvecs = 0
    limit = fixed_iterations if fixed_iterations is not None else maxiter
    it = 0
    for it in range(1, limit + 1):
        q = apply_A(p)
        matvecs += 1
        alpha = rz / float(p @ q)
        x = x + alpha * p
        r_new = r - alpha * q
        if fixed_iterations is None and it % check_every == 0:
            true = b - apply_A(x)
            matvecs += 1
            rel = float(np.linalg.norm(true)) / nb
            if rel <= tol:
                return x, it, True, rel, matvecs
            r_new = true
        z_new = precondition(r_new)
        rz_new = float(r_new @ z_new)
        beta = float(r_new @ (z_new - z)) / rz
        p = z_new + beta * p
        r, z, rz = r_new, z_new, rz_new
    true = b - apply_A(x)
    matvecs += 1
    rel = float(np.linalg.norm(true)) / nb
    return x, it, bool(rel <= tol), rel, matvecs


def spectral_omegas(levels, iterations=30, seed=20260923):
    """omega_l = 4 / (3 lambda_l), lambda_l from a seeded power iteration on D^-1 A over the free
    DOFs (src/voxel_mg.py:228-244). Returns (omegas, lambdas)."""
    omegas, lambdas = [], []
    for L in levels[:-1]:
        rng = np.random.default_rng(seed)
        v = rng.standard_normal(L.mask.size).astype(L.dtype) * L.mask
        v = v / np.linalg.norm(v)
        lam = 0.0
        for _ in range(iterations):
            w = L.inv_diag * L.apply(v) * L.mask
            lam = float(np.linalg.norm(w))
            v = w / lam
        lambdas.append(lam)
        omegas.append(4.0 / (3.0 * lam))
    return omegas, lambdas


def splitmix_uniform(n, seed):
    """2 u_i - 1 on [-1, 1), u_i = (s_i >> 11) 2^-53 with s_i the (i+1)-th output of the SplitMix64
    generator seeded with `seed` (Steele, Lea, Flood, "Fast splittable pseudorandom number
    generators", OOPSLA 2014): s_i = mix(seed + (i + 1) 0x9E3779B97F4A7C15), mix the generator's
    finaliser.  Unsigned 64-bit arithmetic wraps, as the generator's does."""
    with np.errstate(over="ignore"):
        z = np.uint64(int(seed) & (2 ** 64 - 1)) + (np.arange(1, int(n) + 1, dtype=np.uint64)
                                                     * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return 2.0 * ((z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53) - 1.0


def chebyshev_bounds(levels, iterations=30, safety=1.1, seed=20260923):
    """Upper bounds of the spectrum of D^-1 A per smoothing level for the Chebyshev smoother:
    safety x the D-weighted Rayleigh quotient v.Av / v.Dv after `iterations` power steps
    v <- D^-1 A v / |D^-1 A v|_D (donor_yang_gmg multigrid_v4.py:273-295), started from a seeded
    uniform vector (splitmix_uniform; the donor's all-ones start is a smooth vector, far from the
    top of the spectrum), over the free DOFs."""
    lams = []
    for L in levels[:-1]:
        d = np.where(L.mask != 0, 1.0 / L.inv_diag, 1.0).astype(L.dtype)
        v = splitmix_uniform(L.mask.size, seed).astype(L.dtype) * L.mask
        v = v / np.sqrt(float(v @ (d * v)))
        for _ in range(iterations):
            w = L.inv_diag * (L.apply(v) * L.mask)
            v = w / np.sqrt(float(w @ (d * w)))
        Kv = L.apply(v) * L.mask
        lams.append(safety * float(v @ Kv) / float(v @ (d * v)))
    return lams


def build_hierarchy(shape, ke, E, mask, dtype=np.float64, max_coarse_x=16, pre=4, post=4,
                    smoother="jacobi", degree=None):
    """Hierarchy as the reference tests/harness build it (tests/test_mg_cpu.py:98-108,
    experiments/session_d/run_gpu.py:223-236): rediscretised levels with ke * 2^l, host
    Cholesky of the dense coarsest operator, spectral omegas."""
    shapes = hierarchy_shapes(shape, max_coarse_x)
    products = [NumpyProduct(s, ke * 2 ** i, dtype) for i, s in enumerate(shapes)]
    Ec, mc = E, mask
    for i in range(1, len(shapes)):
        Ec, mc = coarsen_modulus(Ec, shapes[i - 1]), coarsen_mask(mc, shapes[i - 1])
    A = assemble_dense(shapes[-1], ke * 2 ** (len(shapes) - 1), Ec, mc)
    factor = np.linalg.cholesky(A)
    if smoother == "chebyshev":
        M = Multigrid(shapes, products, E, mask, factor, dtype, pre=degree or pre,
                      post=degree or post, smoother="chebyshev")
        M.omegas, _ = spectral_omegas(M.levels)
        M.lambdas = chebyshev_bounds(M.levels)
    else:
        M = Multigrid(shapes, products, E, mask, factor, dtype, pre=pre, post=post)
        M.omegas, _ = spectral_omegas(M.levels)
    return shapes, M
