"""Pins the multigrid / flexible CG oracle (oracle/mgref.py): the reference's own CPU test
cases (tests/test_mg_cpu.py), the golden V-cycle / spectral / fcg vectors written from the
reference in this container, the reference's stored smoke-q4 GPU solutions and known answers,
and closed forms for the Chebyshev smoother."""
import json

import numpy as np
import pytest

from oracle import hexref as hx
from oracle import mgref as mg
from oracle import simpref as sr


@pytest.fixture(scope="module")
def KE(golden):
    return np.load(golden / "smoke-q4.npz")["ke"]


def smoke(golden, q=4, seed=None):
    """tests/test_mg_cpu.py:80-95 of the reference."""
    d = np.load(golden / "smoke-q4.npz")
    if q == 4 and seed is None:
        return sr.fixture_case(d)[:4]
    shape = (2 * q, q, q)
    nx, ny, nz = shape
    rng = np.random.default_rng(seed or 0)
    rho = 0.05 + 0.95 * rng.random(nx * ny * nz) ** 2
    m = np.ones((nz + 1, ny + 1, nx + 1, 3)); m[:, :, 0, :] = 0
    f = np.zeros((nz + 1, ny + 1, nx + 1, 3)); f[0, :, nx, 2] = -1.0 / (ny + 1)
    E = hx.to_author(1e-6 + (1 - 1e-6) * rho ** 3, shape)
    return shape, E, hx.to_author(m.ravel(), shape, True), hx.to_author(f.ravel(), shape, True)


def apply_fine(KE, shape, E, mask):
    prod = hx.NumpyProduct(shape, KE)
    return lambda u: prod.raw(u * mask, E) * mask + u * (1 - mask)


def P_matrix(coarse):
    n = 3 * hx.n_nodes(coarse)
    return np.stack([mg.prolong(np.eye(n)[i], coarse) for i in range(n)], axis=1)


def test_restriction_is_the_exact_transpose_of_prolongation():
    rng = np.random.default_rng(1)
    coarse, fine = (2, 1, 1), (4, 2, 2)
    c, f = rng.standard_normal(3 * 3 * 2 * 2), rng.standard_normal(3 * 5 * 3 * 3)
    assert np.isclose(mg.prolong(c, coarse) @ f, c @ mg.restrict(f, fine), rtol=1e-13)


def test_prolongation_interpolates_linear_fields_exactly():
    coarse = (3, 2, 2)
    X = np.stack(np.meshgrid(*[np.arange(n + 1) * 2.0 for n in coarse], indexing="ij"), -1)
    Xf = np.stack(np.meshgrid(*[np.arange(2 * n + 1) * 1.0 for n in coarse], indexing="ij"), -1)
    A = np.array([[1.0, 2, -1], [0.5, 0, 3], [-2, 1, 1]])
    uc = (X.reshape(-1, 3) @ A.T).ravel()
    assert np.abs(mg.prolong(uc, coarse) - (Xf.reshape(-1, 3) @ A.T).ravel()).max() < 1e-13


def test_rediscretisation_equals_galerkin_for_uniform_modulus(KE):
    fine, coarse = (4, 2, 2), (2, 1, 1)
    Kf = mg.assemble_dense(fine, KE, np.ones(16), np.ones(3 * 5 * 3 * 3))
    Kc = mg.assemble_dense(coarse, 2 * KE, np.ones(2), np.ones(3 * 3 * 2 * 2))
    P = P_matrix(coarse)
    assert np.abs(P.T @ Kf @ P - Kc).max() / np.abs(Kc).max() < 1e-13


@pytest.mark.parametrize("clamped", [False, True])
def test_galerkin_operator_equals_rediscretisation_for_uniform_modulus(KE, clamped):
    fine = (4, 2, 4)
    coarse = (2, 1, 2)
    mask = np.ones(3 * hx.n_nodes(fine))
    if clamped:
        mask[: 3 * 3 * 5] = 0                       # the x = 0 face
    Ac = mg.galerkin_dense(fine, KE, np.full(32, 0.7), mask)
    Kc = mg.assemble_dense(coarse, 2 * KE, np.full(4, 0.7), mg.coarsen_mask(mask, fine))
    assert np.abs(Ac - Kc).max() / np.abs(Kc).max() < 1e-13
    assert np.abs(Ac - Ac.T).max() < 1e-14 * np.abs(Ac).max()


def test_galerkin_operator_of_random_moduli_is_spd_with_identity_rows(KE):
    fine = (4, 4, 2)
    rng = np.random.default_rng(5)
    mask = (rng.random(3 * hx.n_nodes(fine)) > 0.2).astype(float)
    Ac = mg.galerkin_dense(fine, KE, rng.uniform(0.1, 1, 32), mask)
    mc = mg.coarsen_mask(mask, fine)
    fixed = np.flatnonzero(mc == 0)
    assert np.array_equal(Ac[fixed], np.eye(Ac.shape[0])[fixed])
    assert np.linalg.eigvalsh(Ac).min() > 0


def test_galerkin_hierarchy_level_one_is_galerkin_dense(KE):
    shapes = [(8, 4, 4), (4, 2, 2), (2, 1, 1)]
    rng = np.random.default_rng(11)
    mask = (rng.random(3 * hx.n_nodes(shapes[0])) > 0.15).astype(float)
    E = rng.uniform(0.01, 1, 128)
    K, A = mg.galerkin_hierarchy(shapes, KE, E, mask)
    ref = mg.galerkin_dense(shapes[0], KE, E, mask)
    assert np.abs(A[1] - ref).max() < 1e-13 * np.abs(ref).max()
    for a in A[1:]:
        assert np.abs(a - a.T).max() < 1e-13 * np.abs(a).max()
        assert np.linalg.eigvalsh(a).min() > 0


@pytest.mark.parametrize("clamped", [False, True])
def test_galerkin_hierarchy_equals_rediscretisation_for_uniform_modulus(KE, clamped):
    """P^T K_h P = K_2h for trilinear P and uniform moduli, level after level; the clamp's fixed
    fine nodes interpolate only from fixed coarse nodes, so the masks nest."""
    shapes = [(8, 4, 4), (4, 2, 2), (2, 1, 1)]
    mask = np.ones(3 * hx.n_nodes(shapes[0]))
    if clamped:
        mask[: 3 * 5 * 5] = 0
    _, A = mg.galerkin_hierarchy(shapes, KE, np.full(128, 0.7), mask)
    m = mask
    for lvl in range(1, 3):
        m = mg.coarsen_mask(m, shapes[lvl - 1])
        ne = int(np.prod(shapes[lvl]))
        R = mg.assemble_dense(shapes[lvl], KE * 2 ** lvl, np.full(ne, 0.7), m)
        assert np.abs(A[lvl] - R).max() < 1e-12 * np.abs(R).max(), lvl
    # a plausible slip - restricting with P instead of P^T scaled to the coarse size - is caught
    P = mg.prolongation_matrix(shapes[1])
    wrong = (P / 8.0).T @ A[0] @ P
    assert np.abs(mg.coarsen_mask(mask, shapes[0])[:, None] * wrong - A[1]).max() > 1e-3 * np.abs(A[1]).max()


def test_vcycle_is_symmetric(golden, KE):
    shape, E, mask, _ = smoke(golden)
    _, M = mg.build_hierarchy(shape, KE, E, mask, max_coarse_x=2)
    rng = np.random.default_rng(2)
    r1, r2 = rng.standard_normal(mask.size) * mask, rng.standard_normal(mask.size) * mask
    a, b = M(r1) @ r2, r1 @ M(r2)
    assert abs(a - b) / abs(a) < 1e-10


def test_mg_fcg_on_the_real_smoke_input(golden, KE):
    shape, E, mask, force = smoke(golden)
    _, M = mg.build_hierarchy(shape, KE, E, mask, max_coarse_x=2)
    x, it, ok, rel, _ = mg.fcg(apply_fine(KE, shape, E, mask), force, M, 1e-8, 200)
    assert ok and rel <= 1e-8 and it <= 30, (it, rel)
    assert np.all(x[mask == 0] == 0)


@pytest.mark.parametrize("q", [4, 8])
def test_iteration_count_is_roughly_mesh_independent(golden, KE, q):
    shape, E, mask, force = smoke(golden, q=q, seed=3)
    _, M = mg.build_hierarchy(shape, KE, E, mask, max_coarse_x=2)
    _, it, ok, rel, _ = mg.fcg(apply_fine(KE, shape, E, mask), force, M, 1e-8, 300)
    assert ok and it <= 60, (q, it, rel)


def test_fp32_preconditioner_still_reaches_1e8_in_fp64_outer(golden, KE):
    shape, E, mask, force = smoke(golden, q=8, seed=4)
    _, M32 = mg.build_hierarchy(shape, KE, E, mask, dtype=np.float32, max_coarse_x=2)
    _, it, ok, rel, _ = mg.fcg(apply_fine(KE, shape, E, mask), force, M32, 1e-8, 300)
    assert ok and rel <= 1e-8, (it, rel)


def test_fixed_iterations_do_exactly_the_requested_work(golden, KE):
    shape, E, mask, force = smoke(golden)
    _, M = mg.build_hierarchy(shape, KE, E, mask, max_coarse_x=2)
    _, it, _, _, mv = mg.fcg(apply_fine(KE, shape, E, mask), force, M, 1e-8, 999, fixed_iterations=7)
    assert it == 7 and mv == 8


def test_single_level_degenerates_to_a_direct_solve(golden, KE):
    shape, E, mask, force = smoke(golden)
    _, M = mg.build_hierarchy(shape, KE, E, mask, max_coarse_x=64)
    _, it, ok, rel, _ = mg.fcg(apply_fine(KE, shape, E, mask), force, M, 1e-8, 10, check_every=1)
    assert ok and it <= 2, (it, rel)


def test_fixed_omega_0_6_is_unstable_on_heterogeneous_density(golden, KE):
    shape, E, mask, _ = smoke(golden, q=8, seed=3)
    _, M = mg.build_hierarchy(shape, KE, E, mask, max_coarse_x=2)
    _, lambdas = mg.spectral_omegas(M.levels)
    assert 0.6 * max(lambdas) > 2.0
    assert all(w * lam == pytest.approx(4 / 3) for w, lam in zip(M.omegas, lambdas))


@pytest.mark.parametrize("tag", ["smoke", "q8s3"])
def test_matches_the_reference_golden_vectors(golden, KE, tag):
    g = np.load(golden / "mg_golden.npz")
    shape = tuple(int(s) for s in g[f"{tag}/shape"])
    E, mask, force = g[f"{tag}/E"], g[f"{tag}/mask"], g[f"{tag}/force"]
    _, M = mg.build_hierarchy(shape, KE, E, mask, max_coarse_x=2)
    _, lams = mg.spectral_omegas(M.levels)
    assert np.allclose(lams, g[f"{tag}/lambdas"], rtol=1e-12, atol=0)
    out = M(g[f"{tag}/vcycle_in"])
    ref = g[f"{tag}/vcycle_out"]
    assert np.linalg.norm(out - ref) / np.linalg.norm(ref) < 1e-12
    x, it, ok, rel, mv = mg.fcg(apply_fine(KE, shape, E, mask), force, M, 1e-8, 300)
    st = g[f"{tag}/fcg_stats"]
    assert (it, int(ok), mv) == (int(st[0]), int(st[1]), int(st[3]))
    assert np.linalg.norm(x - g[f"{tag}/fcg_x"]) / np.linalg.norm(x) < 1e-10
    if tag == "q8s3":
        _, M32 = mg.build_hierarchy(shape, KE, E, mask, dtype=np.float32, max_coarse_x=2)
        assert np.allclose(mg.spectral_omegas(M32.levels)[1], g[f"{tag}/lambdas32"], rtol=1e-5)
        ref32 = g[f"{tag}/vcycle32_out"]
        assert np.linalg.norm(M32(g[f"{tag}/vcycle_in"]) - ref32) / np.linalg.norm(ref32) < 1e-5


def test_mg_solution_matches_the_reference_gpu_run(golden, KE):
    """results/session-d-20260923 smoke-q4--mg64-modal8 (max_coarse_x = 16 -> one level):
    one iteration; same state and compliance."""
    ka = json.loads((golden / "known_answers.json").read_text())["session_d/mg64-modal8"]
    d = np.load(golden / "smoke-q4.npz")
    shape, E, mask, force = sr.fixture_case(d)
    _, M = mg.build_hierarchy(shape, KE, E, mask, max_coarse_x=16)
    x, it, ok, rel, mv = mg.fcg(apply_fine(KE, shape, E, mask), force, M, 1e-8, 200, check_every=1)
    assert ok and it == ka["iterations"] and mv == ka["outer_matvecs"]
    sol = hx.to_author(np.load(golden / "smoke-q4--mg64-modal8--tol-00.npz")["solution"], shape, True)
    assert np.linalg.norm(x - sol) / np.linalg.norm(sol) < 1e-12
    assert float(force @ x) == pytest.approx(ka["force_work"], rel=1e-12)


def test_chebyshev_is_the_chebyshev_polynomial():
    lam = np.linspace(0.05, 2.0, 40)
    lmax, frac = 2.0, 1.0 / 30.0
    for deg in (1, 2, 3, 5):
        x = mg.chebyshev(lambda v: lam * v, np.ones_like(lam), np.zeros_like(lam),
                         np.ones_like(lam), lmax, frac, deg)
        th, de = lmax * (1 + frac) / 2, lmax * (1 - frac) / 2
        T = np.polynomial.chebyshev.Chebyshev.basis(deg)
        assert np.abs((1 - lam * x) - T((th - lam) / de) / T(th / de)).max() < 1e-13
    # degree one is damped Jacobi with omega = 1/theta
    x1 = mg.chebyshev(lambda v: lam * v, np.ones_like(lam), np.zeros_like(lam), np.ones_like(lam),
                      lmax, frac, 1)
    assert np.allclose(x1, 1.0 / (lmax * (1 + frac) / 2))


def test_donor_chebyshev_variant_is_a_different_polynomial():
    """The donor recurrence (multigrid_v4.py:860-889) is kept selectable; it departs from the
    Chebyshev polynomial from degree two on."""
    lam = np.linspace(0.05, 2.0, 40)
    a = mg.chebyshev(lambda v: lam * v, np.ones_like(lam), np.zeros_like(lam), np.ones_like(lam),
                     2.0, 1 / 30, 3, variant="donor")
    b = mg.chebyshev(lambda v: lam * v, np.ones_like(lam), np.zeros_like(lam), np.ones_like(lam),
                     2.0, 1 / 30, 3)
    assert np.abs(a - b).max() > 1e-2


def test_donor_chebyshev_variant_has_the_hand_derived_degree_two_error():
    """Pins mgref.chebyshev(variant="donor") by the error polynomial of the donor's d-form
    (multigrid_v4.py:874-886) worked out by hand for a diagonal operator (D = I, A = lam,
    rhs = lam, x0 = 0, exact solution 1).  Step 1: x1 = z0/theta, e1 = 1 - lam/theta, and the
    direction d0 = z0 = lam is kept unscaled.  Step 2: c = delta^2 a0/4 with a0 = 1/theta,
    q = theta - c, a1 = 1/q, d1 = a1 (lam e1 + c lam), so
        e2 = (1 - lam/theta)(1 - lam/q) - lam c/q.
    Saad's a0 = 2/theta (the first coefficient the donor's comment names) gives a different e2."""
    lmax, frac = 2.0, 1 / 30
    lam = np.linspace(0.0, 2.2, 57)
    theta, delta = 0.5 * lmax * (1 + frac), 0.5 * lmax * (1 - frac)
    c = delta * delta / (4 * theta)
    q = theta - c
    e2 = (1 - lam / theta) * (1 - lam / q) - lam * c / q
    x = mg.chebyshev(lambda v: lam * v, lam.copy(), np.zeros_like(lam), np.ones_like(lam), lmax, frac, 2,
                     variant="donor")
    assert np.abs((1 - x) - e2).max() < 1e-13
    c_saad = delta * delta / (2 * theta)
    e2_saad = (1 - lam / theta) * (1 - lam / (theta - c_saad)) - lam * c_saad / (theta * (theta - c_saad))
    assert np.abs(e2 - e2_saad).max() > 1e-2
    xs = mg.chebyshev(lambda v: lam * v, lam.copy(), np.zeros_like(lam), np.ones_like(lam), lmax, frac, 2)
    assert np.abs((1 - xs) - e2_saad).max() < 1e-13


def test_donor_chebyshev_variant_on_a_diagonal_operator_with_jacobi_scaling():
    """Degree 3 on a diagonal operator with a non-trivial D: the donor's variant equals a
    scalar-by-scalar evaluation of the same d-form in exact rational arithmetic."""
    from fractions import Fraction as F
    rng = np.random.default_rng(5)
    a = rng.uniform(0.5, 3.0, 12)          # diagonal of A
    dscale = rng.uniform(0.8, 1.2, 12)     # D^-1 = dscale / a, so D^-1 A = dscale
    rhs = rng.standard_normal(12)
    inv = dscale / a
    lmax, frac, deg = 1.25, 0.1, 3
    got = mg.chebyshev(lambda v: a * v, rhs, np.zeros(12), inv, lmax, frac, deg, variant="donor")
    lm, lo = F(lmax), F(lmax) * F(frac)
    sig, dl = (lm + lo) / 2, (lm - lo) / 2
    for i in range(12):
        ai, bi, wi = F(a[i]), F(rhs[i]), F(inv[i])
        x = F(0)
        z = wi * (bi - ai * x)
        alpha = 2 / (lm + lo)
        dvec = z
        x = x + alpha * dvec
        for _ in range(1, deg):
            z = wi * (bi - ai * x)
            alpha_new = 1 / (sig - dl * dl * alpha / 4)
            dvec = alpha_new * (z + dl * dl * alpha / 4 * dvec)
            x = x + dvec
            alpha = alpha_new
        assert abs(float(x) - got[i]) <= 1e-13 * max(1.0, abs(float(x)))


def test_chebyshev_multigrid_converges(golden, KE):
    shape, E, mask, force = smoke(golden, q=8, seed=3)
    _, M = mg.build_hierarchy(shape, KE, E, mask, max_coarse_x=2, smoother="chebyshev", degree=3)
    _, it, ok, rel, _ = mg.fcg(apply_fine(KE, shape, E, mask), force, M, 1e-8, 300)
    assert ok and it <= 40, (it, rel)


def test_splitmix_start_vector_reproduces_the_published_generator():
    """splitmix_uniform maps the SplitMix64 stream (Steele, Lea, Flood 2014) to [-1, 1): its first
    values are the generator's published outputs for seed 0 (0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4)
    and seed 1234567 (6457827717110365317, 3203168211198807973, 9817491932198370423,
    4593380528125082431, 16408922859458223821), top 53 bits scaled by 2^-53."""
    to_unit = lambda s: 2.0 * ((s >> 11) * 2.0 ** -53) - 1.0  # noqa: E731
    assert np.array_equal(mg.splitmix_uniform(2, 0), [to_unit(0xE220A8397B1DCDAF), to_unit(0x6E789E6AA1B965F4)])
    pub = [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
           16408922859458223821]
    assert np.array_equal(mg.splitmix_uniform(5, 1234567), [to_unit(v) for v in pub])
    v = mg.splitmix_uniform(200_000, 20260923)
    assert v.min() >= -1.0 and v.max() < 1.0 and abs(v.mean()) < 0.01 and abs(v.var() - 1.0 / 3.0) < 0.01


def test_chebyshev_bounds_enclose_the_spectrum(golden, KE):
    """The Chebyshev upper bound against the dense spectrum of D^-1 A (free DOFs) on the smoke
    state and on a uniform-modulus grid: the Rayleigh estimate is within 5 % of lambda_max, and
    the bound (x 1.1) lies above it."""
    d = np.load(golden / "mg_golden.npz")
    cases = [(tuple(int(v) for v in d["smoke/shape"]), d["smoke/E"], d["smoke/mask"])]
    shape = (8, 4, 4)
    m = np.ones((9, 5, 5, 3)); m[0] = 0
    cases.append((shape, np.full(128, 0.027), m.reshape(-1)))
    for shape, E, mask in cases:
        _, M = mg.build_hierarchy(shape, KE, E, mask, max_coarse_x=2, smoother="chebyshev", degree=3)
        lv = M.levels[0]
        A = hx.assemble_dense(shape, KE, E, mask)
        free = mask != 0
        Af = A[np.ix_(free, free)]
        dg = np.diag(Af)
        lam_true = np.linalg.eigvalsh(Af / np.sqrt(np.outer(dg, dg))).max()
        bound = mg.chebyshev_bounds(M.levels)[0]
        assert lam_true <= bound and bound / 1.1 == pytest.approx(lam_true, rel=0.05)
        assert np.allclose(np.diag(A)[free], 1.0 / lv.inv_diag[free])
