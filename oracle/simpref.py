"""CPU oracle for the SIMP pieces around the solve, and the synthetic problem generator.

TEST INFRASTRUCTURE ONLY (tests/, __graft_entry__.smoke(), bench.py cpu_baseline).

Restates:
  cone_stencil      y_e = sum_{|d| < r} (r - |d|) x_{e+d} over the in-grid neighbours
                    (experiments/session_a/donor/src/gpu_fem/filter_r2.py:40-76), author
                    element order e = (ex*ny + ey)*nz + ez
  ConeFilter        forward = H x / s, transpose = H (g / s), s = H 1 (filter_r2.py:93-183)
  oc_top88          bisection on the multiplier until (l2-l1)/(l1+l2) <= tol, volume on the
                    filtered field (experiments/session_h/run_w1b.py:46-62)
  sensitivity       dc/drho = -p (E0 - Emin) rho^(p-1) u_e^T ke u_e
                    (donor_yang_gmg/src/gpu_fem/solver_v4.py:246-251)
  regenerate/upsample  clamp and consistent line load, nearest-neighbour upsampling
                    (experiments/session_i/make_inputs.py:31-50)
  cantilever        the synthetic voxel cantilever of the benchmark configurations (seeded
                    iid densities, x=0 clamp, unit downward line load on the x=nx, z=nz/2 edge)

Parity status: cone filter pinned by brute force against an explicit sparse H built
independently (DensityFilter of run_w1b.py:97-126) and by <Hx, y> == <x, Hy>; OC pinned by
its invariants (bounds, move limit, achieved volume within tolerance) and by a closed-form
case (uniform sensitivities); sensitivity pinned by central finite differences of the
compliance; regenerate pinned by the smoke-q4 / optimized-q64 fixtures' mask and force.
"""
from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp

from .hexref import to_author


def cone_offsets(rmin):
    """(dx, dy, dz, w) with w = rmin - |d| > 0 for |d_i| <= ceil(rmin) - 1."""
    R = max(0, int(math.ceil(rmin)) - 1)
    out = []
    for dx in range(-R, R + 1):
        for dy in range(-R, R + 1):
            for dz in range(-R, R + 1):
                w = rmin - math.sqrt(dx * dx + dy * dy + dz * dz)
                if w > 0:
                    out.append((dx, dy, dz, w))
    return out


def cone_stencil(x, shape, rmin):
    nx, ny, nz = shape
    xg = np.asarray(x, dtype=np.float64).reshape(nx, ny, nz)
    y = np.zeros_like(xg)
    for dx, dy, dz, w in cone_offsets(rmin):
        dst = tuple(slice(max(0, -d), n - max(0, d)) for d, n in zip((dx, dy, dz), shape))
        src = tuple(slice(max(0, d), n - max(0, -d)) for d, n in zip((dx, dy, dz), shape))
        y[dst] += w * xg[src]
    return y.ravel()


def cone_matrix(shape, rmin):
    """Explicit sparse H, built row by row as the CPU density filter does (run_w1b.py:97-126)."""
    nx, ny, nz = shape
    ex, ey, ez = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ex, ey, ez = ex.ravel(), ey.ravel(), ez.ravel()
    rows, cols, vals = [], [], []
    for dx, dy, dz, w in cone_offsets(rmin):
        jx, jy, jz = ex + dx, ey + dy, ez + dz
        ok = (jx >= 0) & (jx < nx) & (jy >= 0) & (jy < ny) & (jz >= 0) & (jz < nz)
        rows.append(((ex * ny + ey) * nz + ez)[ok])
        cols.append(((jx * ny + jy) * nz + jz)[ok])
        vals.append(np.full(int(ok.sum()), w))
    n = nx * ny * nz
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(n, n))


class ConeFilter:
    def __init__(self, shape, rmin):
        self.shape, self.rmin = tuple(shape), float(rmin)
        self.s = cone_stencil(np.ones(int(np.prod(shape))), self.shape, self.rmin)

    def forward(self, rho):
        return cone_stencil(rho, self.shape, self.rmin) / self.s

    def transpose(self, g):
        return cone_stencil(np.asarray(g) / self.s, self.shape, self.rmin)


def oc_top88(x, dc, dv, volfrac, move, density=None, l2=1e9, tol=1e-3):
    """Returns x_new (run_w1b.py:46-62)."""
    density = density or (lambda v: v)
    l1 = 0.0
    x_new = x
    while (l2 - l1) / (l1 + l2) > tol:
        lmid = 0.5 * (l2 + l1)
        x_new = np.clip(x * np.sqrt(np.maximum(-dc / dv / lmid, 0.0)),
                        np.maximum(x - move, 0.0), np.minimum(x + move, 1.0))
        if density(x_new).mean() > volfrac:
            l1 = lmid
        else:
            l2 = lmid
    return x_new


def modulus(rho, penal=3.0, E0=1.0, Emin=1e-9):
    return Emin + (E0 - Emin) * np.asarray(rho) ** penal


def sensitivity(rho, ce, penal=3.0, E0=1.0, Emin=1e-9):
    return -penal * (E0 - Emin) * np.asarray(rho) ** (penal - 1) * ce


# ------------------------------------------------------------------------- problems
def regenerate(shape):
    """Natural-layout x=0 clamp and -z line load at x = nx, z = 0 (make_inputs.py:31-39)."""
    nx, ny, nz = (int(n) for n in shape)
    mask = np.ones((nz + 1, ny + 1, nx + 1, 3))
    mask[:, :, 0, :] = 0.0
    force = np.zeros((nz + 1, ny + 1, nx + 1, 3))
    line = np.full(ny + 1, -1.0 / ny)
    line[[0, -1]] = -0.5 / ny
    force[0, :, nx, 2] = line
    return mask.ravel(), force.ravel()


def upsample(data, factor):
    """Nearest-neighbour upsampling with regenerated boundary data (make_inputs.py:42-50)."""
    nx, ny, nz = (int(n) for n in data["shape"])
    rho = np.asarray(data["rho"]).reshape(nz, ny, nx)
    for axis in range(3):
        rho = np.repeat(rho, factor, axis=axis)
    shape = np.array([nx * factor, ny * factor, nz * factor])
    mask, force = regenerate(shape)
    return {"shape": shape, "rho": np.ascontiguousarray(rho.ravel()), "mask": mask,
            "force": force, "ke": data["ke"]}


def fixture_case(data, emin=None):
    """Author-layout (shape, E, mask, force) of a reference .npz state (tests/test_mg_cpu.py:80-95)."""
    shape = tuple(int(s) for s in data["shape"])
    floor = 1e-6 if emin is None else emin
    E = to_author(floor + (1 - floor) * np.asarray(data["rho"]) ** 3, shape)
    return (shape, E, to_author(data["mask"], shape, True), to_author(data["force"], shape, True))


def cantilever(shape, seed, rmin=None, rho_lo=0.01, rho_hi=1.0, E0=1.0, Emin=1e-9, penal=3.0):
    """Author-layout synthetic cantilever: (rho, E, mask, force).

    rho iid U[rho_lo, rho_hi] from default_rng(seed) in author element order, optionally
    smoothed by the cone filter of radius rmin; every DOF of the x=0 node plane clamped; a unit
    total -z line load along y on nodes (nx, iy, nz//2): -1/ny inside, -1/(2 ny) at both ends.
    """
    nx, ny, nz = (int(s) for s in shape)
    rng = np.random.default_rng(seed)
    rho = rng.uniform(rho_lo, rho_hi, nx * ny * nz)
    if rmin is not None:
        rho = ConeFilter((nx, ny, nz), rmin).forward(rho)
    E = modulus(rho, penal, E0, Emin)
    mask = np.ones((nx + 1, ny + 1, nz + 1, 3))
    mask[0] = 0.0
    force = np.zeros((nx + 1, ny + 1, nz + 1, 3))
    line = np.full(ny + 1, -1.0 / ny)
    line[[0, -1]] = -0.5 / ny
    force[nx, :, nz // 2, 2] = line
    return rho, E, mask.ravel(), force.ravel()
