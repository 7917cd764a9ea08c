"""CPU oracle for the SIMP pieces around the solve, and the synthetic problem generator.

TEST INFRASTRUCTURE ONLY (tests/, __graft_entry__.smoke(), bench.py cpu_baseline).

Restates:
  cone_stencil      y_e = sum_{|d| < r} (r - |d|) x_{e+d} over the in-grid neighbours
                    (experiments/session_a/donor/src/gpu_fem/filter_r2.py:40-76), author
                    element order e = (ex*ny + ey)*nz + ez
  ConeFilter        forward = H x / s, transpose = H (g / s), s = H 1 (filter_r2.py:93-183)
  oc_top88          bisection on the multiplier until (l2-l1)/(l1+l2) <= tol, volume on the
                    filtered field (experiments/session_h/run_w1b.py:46-62)
  top88_run         the density-filtered top88 SIMP loop (run_w1b.py:163-222) with direct
                    state solves (state_solve)
  point_load_case   clamp and point load of the donor's problem presets
                    (donor_yang_gmg/src/gpu_fem/bc_generator.py:381-447)
  sensitivity       dc/drho = -p (E0 - Emin) rho^(p-1) u_e^T ke u_e
                    (donor_yang_gmg/src/gpu_fem/solver_v4.py:246-251)
  regenerate/upsample  clamp and consistent line load, nearest-neighbour upsampling
                    (experiments/session_i/make_inputs.py:31-50)
  cantilever        the synthetic voxel cantilever of the benchmark configurations (seeded
                    iid densities, x=0 clamp, unit downward line load on the x=nx, z=nz/2 edge)
  project/dproject, schedule_at, oc_update, run_simp
                    the donor's projected driver (simp_r2.py:388-433, 467-549, 552-918) with a
                    direct sparse state solve

Parity status: cone filter pinned by brute force against an explicit sparse H built
independently (DensityFilter of run_w1b.py:97-126) and by <Hx, y> == <x, Hy>; OC pinned by
its invariants (bounds, move limit, achieved volume within tolerance) and by a closed-form
case (uniform sensitivities); sensitivity pinned by central finite differences of the
compliance; regenerate pinned by the smoke-q4 / optimized-q64 fixtures' mask and force;
projection pinned by its closed values (0 -> 0, eta -> eta, 1 -> 1), monotonicity, the
beta -> 0 limit and central differences of dproject; the projected OC by its invariants
(box, move limit, achieved physical volume at the bisection's end) and by the closed form of
uniform sensitivities (the multiplier that keeps a uniform design); run_simp by the volume
constraint at every step and by the compliance decrease over its first phase.
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


# ------------------------------------------------------------------ projected SIMP driver
# simp_r2.py:388-433 (projection, continuation), :467-549 (OC on the physical volume),
# :552-918 (run_simp), restated in NumPy with a direct sparse state solve.
DEFAULT_SCHEDULE = [(15, 1.5, 1.0, 0.20, 1.50), (40, 3.5, 4.0, 0.15, 1.35),
                    (65, 4.5, 16.0, 0.08, 1.25), (120, 4.5, 32.0, 0.05, 1.20)]


def schedule_at(it, schedule=None):
    sch = schedule or DEFAULT_SCHEDULE
    for until, p, beta, move, rmin in sch:
        if it <= until:
            return p, beta, move, rmin, until == sch[-1][0]
    return (*sch[-1][1:], True)


def project(rho_bar, beta, eta=0.5):
    if beta <= 1e-12:
        return np.array(rho_bar, dtype=np.float64, copy=True)
    tb = math.tanh(beta * eta)
    den = tb + math.tanh(beta * (1.0 - eta))
    return (tb + np.tanh(beta * (np.asarray(rho_bar) - eta))) / den


def dproject(rho_bar, beta, eta=0.5):
    if beta <= 1e-12:
        return np.ones_like(np.asarray(rho_bar, dtype=np.float64))
    tb = math.tanh(beta * eta)
    den = tb + math.tanh(beta * (1.0 - eta))
    ch = np.cosh(beta * (np.asarray(rho_bar) - eta))
    return beta / (den * ch * ch)


def oc_update(rho, dc, dv, volfrac, move, filt, beta, eta, rho_lb, bisect_tol=1e-8, max_bisect=80,
              lam_hint=None):
    dcn = np.minimum(dc, -1e-30)
    dvp = np.maximum(dv, 1e-30)

    def design(lmid):
        step = np.sqrt(-dcn / (lmid * dvp))
        rn = np.clip(np.clip(rho * step, rho - move, rho + move), rho_lb, 1.0)
        return rn, float(project(filt.forward(rn), beta, eta).mean())

    probes = 0
    if lam_hint is not None and lam_hint > 0:
        l1, l2 = lam_hint / 4.0, lam_hint * 4.0
        for _ in range(30):
            _, v1 = design(l1)
            _, v2 = design(l2)
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
    rho_new, achieved = design(lam)
    probes += 1
    for _ in range(max_bisect):
        if (l2 - l1) <= bisect_tol * (l1 + l2 + 1e-30):
            break
        lam = 0.5 * (l1 + l2)
        rho_new, achieved = design(lam)
        probes += 1
        if achieved > volfrac:
            l1 = lam
        else:
            l2 = lam
    return rho_new, achieved, lam, probes


def state_solve(shape, ke, mask, force, E):
    """Direct sparse solve of K(E) u = f on the free DOFs (u = 0 at fixed DOFs); returns
    (u, compliance f.u, element energies u_e^T ke u_e)."""
    from scipy.sparse.linalg import spsolve
    from .hexref import edof_author
    edof = edof_author(shape)
    free = np.flatnonzero(np.asarray(mask) != 0)
    nd = 3 * (shape[0] + 1) * (shape[1] + 1) * (shape[2] + 1)
    K = sp.coo_matrix(((np.asarray(E)[:, None, None] * ke[None]).ravel(),
                       (np.repeat(edof, 24, axis=1).ravel(), np.tile(edof, (1, 24)).ravel())),
                      shape=(nd, nd)).tocsr()
    u = np.zeros(nd)
    u[free] = spsolve(K[free][:, free].tocsc(), np.asarray(force)[free])
    ue = u[edof]
    return u, float(np.asarray(force) @ u), np.einsum("ei,ij,ej->e", ue, ke, ue)


def top88_run(shape, ke, mask, force, steps, volfrac=0.3, rmin=1.5, move=0.2, penal=3.0, E0=1.0,
              Emin=1e-9, stop_change=None):
    """The density-filtered SIMP loop with the top88 OC of experiments/session_h/run_w1b.py:163-222
    (update "top88"): x starts at volfrac, rho = filtered x; each step solves the state on rho,
    takes dc/drho = -p (E0-Emin) rho^(p-1) c_e, maps dc and dv = 1 back through the filter
    transpose, bisects the OC multiplier on the filtered volume, and records compliance, the
    largest design change and the filtered volume.  State solves are direct.
    Returns (history [(compliance, change, volume)], x, rho)."""
    filt = ConeFilter(shape, rmin)
    n = int(np.prod(shape))
    x = np.full(n, float(volfrac))
    rho = filt.forward(x)
    dv_x = filt.transpose(np.ones(n))
    hist = []
    for _ in range(steps):
        _, c, ce = state_solve(shape, ke, mask, force, modulus(rho, penal, E0, Emin))
        dc_x = filt.transpose(sensitivity(rho, ce, penal, E0, Emin))
        x_new = oc_top88(x, dc_x, dv_x, volfrac, move, density=filt.forward)
        rho = filt.forward(x_new)
        change = float(np.max(np.abs(x_new - x)))
        x = x_new
        hist.append((c, change, float(rho.mean())))
        if stop_change is not None and change <= stop_change:
            break
    return hist, x, rho


def run_simp(shape, ke, mask, force, steps, volfrac=0.30, rmin=1.5, penal=None, E0=1.0, Emin=1e-9,
             rho_lb=1e-3, eta=0.5, schedule=None):
    """`steps` outer iterations of the projected SIMP loop with exact state solves; returns the
    per-step (compliance, physical volume) history and the final design variable."""
    n = int(np.prod(shape))
    rho = np.full(n, float(volfrac))
    hist = []
    lam_prev = None
    for it in range(1, steps + 1):
        p_it, beta, move, r_it, _ = schedule_at(it, schedule)
        if penal is not None:
            p_it = penal
        filt = ConeFilter(shape, r_it)
        rho_bar = filt.forward(rho)
        rp = project(rho_bar, beta, eta)
        E = modulus(rp, p_it, E0, Emin)
        _, c, ce = state_solve(shape, ke, mask, force, E)
        dc_drp = sensitivity(rp, ce, p_it, E0, Emin)
        dprj = dproject(rho_bar, beta, eta)
        dc = filt.transpose(dc_drp * dprj)
        dv = filt.transpose(dprj / n)
        rho, vol, lam_prev, _ = oc_update(rho, dc, dv, volfrac, move, filt, beta, eta, rho_lb, lam_hint=lam_prev)
        hist.append((c, vol))
    return hist, rho


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


def point_load_case(dims, lengths, point, load, clamp="left"):
    """Author-layout (mask, force) of a donor problem preset with one face clamped and one point
    load (donor_yang_gmg/src/gpu_fem/bc_generator.py:381-447): node coordinates i*L/n, the load
    goes to the node nearest `point` (first minimum of the Euclidean distance), every DOF of the
    nodes on the x = 0 ("left") face is fixed.  The author node numbering
    (ix*(ny+1) + iy)*(nz+1) + iz is the donor's."""
    nx, ny, nz = (int(n) for n in dims)
    if clamp != "left":
        raise ValueError(clamp)
    ix, iy, iz = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    coords = np.stack([ix.ravel() * (lengths[0] / nx), iy.ravel() * (lengths[1] / ny),
                       iz.ravel() * (lengths[2] / nz)], axis=1)
    node = int(np.argmin(np.linalg.norm(coords - np.asarray(point, dtype=np.float64)[None], axis=1)))
    mask = np.ones((nx + 1, ny + 1, nz + 1, 3))
    mask[0] = 0.0
    force = np.zeros(3 * (nx + 1) * (ny + 1) * (nz + 1))
    force[3 * node:3 * node + 3] += np.asarray(load, dtype=np.float64)
    force *= mask.ravel()
    return mask.ravel(), force


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
