"""Pins the Jacobi-PCG / mixed IR oracle (oracle/solvers.py) on the reference's own smoke-q4
state and stored GPU solutions, and the SIMP pieces (oracle/simpref.py) by brute force,
adjointness, invariants and finite differences."""
import json

import numpy as np
import pytest

from oracle import hexref as hx
from oracle import simpref as sr
from oracle import solvers as sv


@pytest.fixture(scope="module")
def smoke_ctx(golden):
    d = np.load(golden / "smoke-q4.npz")
    shape, E, mask, force = sr.fixture_case(d)
    prod = hx.NumpyProduct(shape, d["ke"])
    diag = np.where(mask != 0, prod.diagonal(E), 1.0)
    ka = json.loads((golden / "known_answers.json").read_text())
    return dict(d=d, shape=shape, E=E, mask=mask, force=force, prod=prod, inv=1 / diag, ka=ka)


def _action(ctx, dtype=np.float64):
    prod32 = hx.NumpyProduct(ctx["shape"], ctx["d"]["ke"], dtype)
    m = ctx["mask"].astype(dtype)
    E = ctx["E"].astype(dtype)
    return lambda u: prod32.raw(u * m, E) * m + u * (1 - m)


def test_pcg_known_answers_on_smoke_q4(smoke_ctx, golden):
    c = smoke_ctx
    ka = c["ka"]["session_a/pcg64-modal8"]
    x, it, ok, rel = sv.pcg(_action(c), c["force"], c["inv"], tol=1e-8, maxiter=20000,
                            mask=c["mask"], check_every=50)
    assert ok and it == ka["iterations"] and rel < 1e-12
    sol = hx.to_author(np.load(golden / "smoke-q4--pcg64-modal8--tol-00.npz")["solution"],
                       c["shape"], True)
    assert np.linalg.norm(x - sol) / np.linalg.norm(sol) < 1e-10
    assert float(c["force"] @ x) == pytest.approx(ka["force_work"], rel=1e-11)
    assert np.all(x[c["mask"] == 0] == 0)


def test_pcg_fixed_work_matvec_count(smoke_ctx):
    """fixed-work solves run tol=0, maxiter=100: 104 matvecs (1 initial + 100 + 2 checks + 1 exit)."""
    c = smoke_ctx
    calls = [0]
    A = _action(c)

    def counted(u):
        calls[0] += 1
        return A(u)
    _, it, ok, _ = sv.pcg(counted, c["force"], c["inv"], tol=0.0, maxiter=100, mask=c["mask"],
                          check_every=50)
    ka = c["ka"]["session_a/fixed/modal8"]
    assert it == ka["iterations"] and calls[0] == ka["matvec_calls"] and not ok


def test_pcg_ir_known_answers_on_smoke_q4(smoke_ctx, golden):
    c = smoke_ctx
    ka = c["ka"]["session_a/ir-modal8-modal8_32"]
    x, outer, inner, ok, rel = sv.pcg_ir(_action(c), c["force"], _action(c, np.float32),
                                         c["inv"].astype(np.float32), tol=1e-8,
                                         mask32=c["mask"].astype(np.float32), inner_tol=0.1,
                                         max_outer=40, inner_maxiter=4000, check_every=50)
    assert ok and rel <= 1e-8
    assert outer == ka["outer_iterations"] and inner == ka["inner_iterations"]
    sol = hx.to_author(np.load(golden / "smoke-q4--pcg64-modal8--tol-00.npz")["solution"],
                       c["shape"], True)
    assert np.linalg.norm(x - sol) / np.linalg.norm(sol) < 5e-5      # protocol cross-state limit


def test_pcg_reports_nonconvergence(smoke_ctx):
    c = smoke_ctx
    _, it, ok, rel = sv.pcg(_action(c), c["force"], c["inv"], tol=1e-8, maxiter=1, mask=c["mask"])
    assert it == 1 and not ok and rel > 1e-8


def test_cone_filter_matches_the_explicit_matrix():
    shape = (5, 4, 3)
    x = np.random.default_rng(0).random(60)
    for rmin in (1.5, 2.2, 1.0):
        H = sr.cone_matrix(shape, rmin)
        assert np.abs(sr.cone_stencil(x, shape, rmin) - H @ x).max() < 1e-14
    assert len(sr.cone_offsets(1.5)) == 19
    w = sorted({round(o[3], 12) for o in sr.cone_offsets(1.5)})
    assert np.allclose(w, [1.5 - np.sqrt(2), 0.5, 1.5])


def test_cone_filter_adjointness_and_constants():
    shape = (6, 5, 4)
    f = sr.ConeFilter(shape, 1.5)
    rng = np.random.default_rng(2)
    a, b = rng.random(120), rng.random(120)
    lhs, rhs = f.forward(a) @ b, a @ f.transpose(b)
    assert abs(lhs - rhs) / abs(lhs) < 1e-12
    assert np.allclose(f.forward(np.full(120, 0.37)), 0.37, rtol=1e-14)


def test_oc_top88_invariants():
    rng = np.random.default_rng(3)
    n = 200
    x = rng.uniform(0.2, 0.6, n)
    dc = -rng.random(n)
    dv = np.ones(n)
    xn = sr.oc_top88(x, dc, dv, 0.4, 0.2)
    assert np.all(xn >= np.maximum(x - 0.2, 0) - 1e-15) and np.all(xn <= np.minimum(x + 0.2, 1) + 1e-15)
    assert abs(xn.mean() - 0.4) < 5e-3
    # uniform sensitivities: the update is uniform scaling by the volume fraction ratio
    x0 = np.full(n, 0.5)
    xu = sr.oc_top88(x0, -np.ones(n), dv, 0.45, 0.2)
    assert np.allclose(xu, xu[0]) and abs(xu[0] - 0.45) < 1e-3


def _interior_ratio_check(x, xn, B, lo, hi):
    """For every element the bisection left strictly inside its box, x_new = x sqrt(B / lam) with
    one common lam: x_new^2 / (x^2 B) is the same constant 1/lam (the sqrt step of the OC update,
    eta = 1/2).  Returns the largest relative spread of that constant."""
    inside = (xn > lo + 1e-9) & (xn < hi - 1e-9)
    assert inside.sum() >= 10
    k = xn[inside] ** 2 / (x[inside] ** 2 * B[inside])
    return float((k.max() - k.min()) / k.mean())


def _oc_top88_without_sqrt(x, dc, dv, volfrac, move, l2=1e9, tol=1e-3):
    l1 = 0.0
    while (l2 - l1) / (l1 + l2) > tol:
        lmid = 0.5 * (l2 + l1)
        x_new = np.clip(x * np.maximum(-dc / dv / lmid, 0.0), np.maximum(x - move, 0.0), np.minimum(x + move, 1.0))
        if x_new.mean() > volfrac:
            l1 = lmid
        else:
            l2 = lmid
    return x_new


def test_oc_top88_step_is_the_square_root_of_the_sensitivity_ratio():
    rng = np.random.default_rng(11)
    n = 400
    x = rng.uniform(0.2, 0.6, n)
    dc = -rng.uniform(0.5, 2.0, n)
    dv = rng.uniform(0.8, 1.2, n)
    B = -dc / dv
    xn = sr.oc_top88(x, dc, dv, 0.4, 0.3)
    assert _interior_ratio_check(x, xn, B, np.maximum(x - 0.3, 0), np.minimum(x + 0.3, 1)) < 1e-12
    # a step that drops the square root fails the same check
    bad = _oc_top88_without_sqrt(x, dc, dv, 0.4, 0.3)
    inside = (bad > np.maximum(x - 0.3, 0) + 1e-9) & (bad < np.minimum(x + 0.3, 1) - 1e-9)
    k = bad[inside] ** 2 / (x[inside] ** 2 * B[inside])
    assert (k.max() - k.min()) / k.mean() > 0.1


def test_projected_oc_step_is_the_square_root_of_the_sensitivity_ratio():
    shape = (10, 6, 6)
    rng = np.random.default_rng(12)
    n = int(np.prod(shape))
    filt = sr.ConeFilter(shape, 1.5)
    rho = rng.uniform(0.2, 0.6, n)
    dc, dv = -rng.uniform(0.5, 2.0, n), rng.uniform(0.8, 1.2, n) / n
    move = 0.3
    for beta in (1.0, 8.0):
        new, _, lam, _ = sr.oc_update(rho, dc, dv, 0.4, move, filt, beta, 0.5, 1e-3)
        lo, hi = np.maximum(rho - move, 1e-3), np.minimum(rho + move, 1.0)
        assert _interior_ratio_check(rho, new, -dc / dv, lo, hi) < 1e-12
        inside = (new > lo + 1e-9) & (new < hi - 1e-9)
        # and the common constant is the returned multiplier
        assert np.allclose(new[inside] ** 2 / (rho[inside] ** 2 * (-dc / dv)[inside]), 1.0 / lam, rtol=1e-12)


def test_continuation_schedule_matches_the_donor_module(golden):
    """Against schedule_at evaluated by the donor's own simp_r2.py (tests/golden/make_schedule_golden.py)."""
    g = json.loads((golden / "schedule_golden.json").read_text())
    for it, row in enumerate(g["default"], start=1):
        assert list(sr.schedule_at(it)) == row
    table = [tuple(r) for r in g["custom_table"]]
    for it, row in enumerate(g["custom"], start=1):
        assert list(sr.schedule_at(it, table)) == row


def test_point_load_case_of_the_donor_cantilever_preset():
    """cantilever_gpu_large (presets.py:146-152): 120x60x30 on 2 x 1 x 0.5, left face clamped, fy = -1 at
    (2.0, 0.5, 0.25) -> node (120, 30, 15)."""
    mask, force = sr.point_load_case((120, 60, 30), (2.0, 1.0, 0.5), (2.0, 0.5, 0.25), (0.0, -1.0, 0.0))
    node = (120 * 61 + 30) * 31 + 15
    assert force[3 * node + 1] == -1.0 and np.count_nonzero(force) == 1
    m = mask.reshape(121, 61, 31, 3)
    assert np.all(m[0] == 0) and np.all(m[1:] == 1)


def test_top88_oracle_loop_keeps_the_volume_and_lowers_the_compliance():
    shape = (12, 6, 6)
    ke = 2 * hx.quadrature([1, 1, 1], 0.3)
    mask, force = sr.point_load_case(shape, (2.0, 1.0, 1.0), (2.0, 0.5, 0.5), (0.0, -1.0, 0.0))
    hist, x, rho = sr.top88_run(shape, ke, mask, force, steps=6, volfrac=0.3, move=0.15)
    c = [h[0] for h in hist]
    assert all(b < a for a, b in zip(c, c[1:]))
    assert all(abs(h[2] - 0.3) < 2e-3 for h in hist)
    assert all(h[1] <= 0.15 + 1e-15 for h in hist)
    assert np.allclose(rho, sr.ConeFilter(shape, 1.5).forward(x))
    # step 0 is the uniform design: compliance of E = 0.3^3 everywhere
    _, c0, _ = sr.state_solve(shape, ke, mask, force, np.full(x.size, sr.modulus(0.3)))
    assert c[0] == pytest.approx(c0, rel=1e-12)


def test_sensitivity_matches_central_finite_differences():
    shape = (4, 2, 2)
    ke = 2 * hx.quadrature([1, 1, 1], 0.3)
    rho, E, mask, force = sr.cantilever(shape, seed=9, rho_lo=0.3)
    A = lambda r: hx.assemble_dense(shape, ke, sr.modulus(r), mask)  # noqa: E731

    def compliance(r):
        return force @ np.linalg.solve(A(r), force)
    u = np.linalg.solve(A(rho), force)
    dc = sr.sensitivity(rho, hx.element_energy(shape, ke, u))
    h = 1e-6
    for e in (0, 5, 11):
        rp, rm = rho.copy(), rho.copy()
        rp[e] += h; rm[e] -= h
        fd = (compliance(rp) - compliance(rm)) / (2 * h)
        assert fd == pytest.approx(dc[e], rel=1e-6)


def test_regenerate_reproduces_the_fixture_boundary_data(golden):
    for name in ("smoke-q4.npz", "optimized-q64.npz"):
        d = np.load(golden / name)
        m, f = sr.regenerate(d["shape"])
        assert np.array_equal(m, d["mask"]) and np.array_equal(f, d["force"])


def test_cantilever_load_is_unit_and_clamp_is_x0():
    rho, E, mask, force = sr.cantilever((8, 4, 4), seed=1)
    assert force.sum() == pytest.approx(-1.0)
    m = mask.reshape(9, 5, 5, 3)
    assert np.all(m[0] == 0) and np.all(m[1:] == 1)
    assert np.all((rho >= 0.01) & (rho <= 1.0))


# ------------------------------------------------------------------ projected SIMP driver (oracle)
@pytest.mark.parametrize("beta", [1.0, 4.0, 32.0])
def test_projection_closed_values_monotonicity_and_derivative(beta):
    x = np.linspace(0.0, 1.0, 101)
    y = sr.project(x, beta)
    assert abs(y[0]) < 1e-15 and abs(y[-1] - 1.0) < 1e-15 and abs(sr.project(np.array([0.5]), beta)[0] - 0.5) < 1e-15
    assert np.all(np.diff(y) > 0)
    h = 1e-6
    xm = np.linspace(0.05, 0.95, 19)
    fd = (sr.project(xm + h, beta) - sr.project(xm - h, beta)) / (2 * h)
    assert np.allclose(sr.dproject(xm, beta), fd, rtol=1e-6, atol=1e-9)
    # the beta -> 0 limit is the identity
    assert np.abs(sr.project(x, 1e-6) - x).max() < 1e-12 and np.array_equal(sr.project(x, 0.0), x)


def test_projected_oc_invariants_and_uniform_case():
    shape = (10, 6, 6)
    rng = np.random.default_rng(7)
    n = int(np.prod(shape))
    filt = sr.ConeFilter(shape, 1.5)
    rho = rng.uniform(0.05, 0.9, n)
    dc, dv = -rng.uniform(0.1, 2.0, n), rng.uniform(0.5, 1.5, n) / n
    for beta, lam_hint in ((1.0, None), (8.0, 3.0)):
        new, vol, lam, _ = sr.oc_update(rho, dc, dv, 0.4, 0.1, filt, beta, 0.5, 1e-3, lam_hint=lam_hint)
        assert new.min() >= 1e-3 and new.max() <= 1.0
        assert np.all(np.abs(new - rho) <= 0.1 + 1e-15)
        assert abs(vol - 0.4) < 1e-6
        assert abs(vol - sr.project(filt.forward(new), beta).mean()) < 1e-15
    # uniform design and sensitivities: the volume-preserving multiplier keeps the design
    u = np.full(n, 0.4)
    new, vol, lam, _ = sr.oc_update(u, np.full(n, -2.0), np.full(n, 1.0 / n), 0.4, 0.2, filt, 0.0, 0.5, 1e-3)
    assert abs(lam - 2.0 * n) / (2.0 * n) < 1e-6 and np.abs(new - 0.4).max() < 1e-6


def test_projected_simp_oracle_keeps_the_volume_and_lowers_the_compliance():
    shape = (8, 4, 4)
    ke = hx.quadrature([1, 1, 1], 0.3)
    _, _, mask, force = sr.cantilever(shape, seed=1)
    hist, rho = sr.run_simp(shape, ke, mask, force, steps=8)
    c = [h[0] for h in hist]
    assert all(abs(v - 0.3) < 1e-6 for _, v in hist)
    assert all(b < a for a, b in zip(c, c[1:]))
    assert rho.min() >= 1e-3 and rho.max() <= 1.0
