"""GPU parity of the multigrid and flexible CG drop-ins (voxel_mg on the HIP kernels) against
the oracle and the reference's golden vectors / known answers.

Tolerances: a V-cycle differs from the oracle only by rounding and by applying the coarsest
solve through the explicit inverse instead of two triangular solves: relative L2 <= 1e-10 in
FP64, <= 1e-5 in FP32.  Solves: same iteration counts as the reference where it publishes them,
true relative residual <= 1e-8, solution within 1e-9 of the oracle's.
"""
import json

import numpy as np
import pytest
import torch

from oracle import hexref as hx
from oracle import mgref as mg_ref
from oracle import simpref as sr

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else a
    return float(np.linalg.norm(np.asarray(a, dtype=np.float64) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def KE(golden):
    return np.load(golden / "smoke-q4.npz")["ke"]


def build(KE, shape, E, mask, dtype=np.float64, max_coarse_x=2, smoother="jacobi", degree=None, coarse_dtype=None):
    import voxelhip
    import voxel_mg as mg
    xp = voxelhip.xp
    shapes = mg.hierarchy_shapes(shape, max_coarse_x)
    products = [mg.HipProduct(s, KE * 2 ** i, dtype) for i, s in enumerate(shapes)]
    Ec, mc = E, mask
    for i in range(1, len(shapes)):
        Ec, mc = mg_ref.coarsen_modulus(Ec, shapes[i - 1]), mg_ref.coarsen_mask(mc, shapes[i - 1])
    A = mg.assemble_dense(shapes[-1], KE * 2 ** (len(shapes) - 1), Ec, mc)
    kw = {}
    if smoother == "chebyshev":
        kw = dict(smoother="chebyshev", pre=degree, post=degree)
    M = mg.Multigrid(xp, shapes, products, E, mask, np.linalg.cholesky(A), dtype, coarse_dtype=coarse_dtype, **kw)
    om, lam = mg.spectral_omegas(xp, M.levels)
    M.omegas = om
    if smoother == "chebyshev":
        lam = mg.chebyshev_bounds(xp, M.levels)
        M.lambdas = lam
    return shapes, M, lam


def test_spectral_omegas_and_vcycle_match_the_golden_vectors(golden, KE):
    g = np.load(golden / "mg_golden.npz")
    for tag in ("smoke", "q8s3"):
        shape = tuple(int(s) for s in g[f"{tag}/shape"])
        print("case", tag, shape, flush=True)
        E, mask = g[f"{tag}/E"], g[f"{tag}/mask"]
        _, M, lam = build(KE, shape, E, mask)
        assert np.allclose(lam, g[f"{tag}/lambdas"], rtol=1e-12, atol=0)
        r = torch.from_numpy(g[f"{tag}/vcycle_in"]).cuda()
        assert rel(M(r), g[f"{tag}/vcycle_out"]) < 1e-10


def test_fp32_vcycle_matches_the_golden_vector(golden, KE):
    g = np.load(golden / "mg_golden.npz")
    shape = tuple(int(s) for s in g["q8s3/shape"])
    _, M32, lam = build(KE, shape, g["q8s3/E"], g["q8s3/mask"], dtype=np.float32)
    assert np.allclose(lam, g["q8s3/lambdas32"], rtol=1e-5)
    r = torch.from_numpy(g["q8s3/vcycle_in"]).cuda()
    out = M32(r)
    assert out.dtype == torch.float64 and rel(out, g["q8s3/vcycle32_out"]) < 1e-5


@pytest.mark.parametrize("tag", ["smoke", "q8s3"])
def test_fcg_matches_the_golden_solve(golden, KE, tag):
    import voxelhip
    import voxel_mg as mg
    g = np.load(golden / "mg_golden.npz")
    shape = tuple(int(s) for s in g[f"{tag}/shape"])
    E, mask, force = g[f"{tag}/E"], g[f"{tag}/mask"], g[f"{tag}/force"]
    _, M, _ = build(KE, shape, E, mask)
    lvl = M.levels[0]
    x, it, ok, relres, mv = mg.fcg(voxelhip.xp, lvl.apply, torch.from_numpy(force).cuda(), M, 1e-8, 300)
    st = g[f"{tag}/fcg_stats"]
    assert ok and (it, mv) == (int(st[0]), int(st[3])) and relres <= 1e-8
    assert rel(x, g[f"{tag}/fcg_x"]) < 1e-9
    xm = x.cpu().numpy()
    assert np.all(xm[mask == 0] == 0)


def test_fcg_known_answer_on_the_reference_smoke_state(golden, KE):
    """results/session-d-20260923 smoke-q4--mg64-modal8: 1 iteration, 2 outer matvecs, same state."""
    import voxelhip
    import voxel_mg as mg
    ka = json.loads((golden / "known_answers.json").read_text())["session_d/mg64-modal8"]
    shape, E, mask, force = sr.fixture_case(np.load(golden / "smoke-q4.npz"))
    _, M, _ = build(KE, shape, E, mask, max_coarse_x=16)
    x, it, ok, relres, mv = mg.fcg(voxelhip.xp, M.levels[0].apply, torch.from_numpy(force).cuda(), M, 1e-8, 200,
                                   check_every=1)
    assert ok and it == ka["iterations"] and mv == ka["outer_matvecs"]
    sol = hx.to_author(np.load(golden / "smoke-q4--mg64-modal8--tol-00.npz")["solution"], shape, True)
    assert rel(x, sol) < 1e-11
    assert float(torch.from_numpy(force).cuda() @ x) == pytest.approx(ka["force_work"], rel=1e-11)


def test_fixed_iterations_work_accounting(golden, KE):
    import voxelhip
    import voxel_mg as mg
    shape, E, mask, force = sr.fixture_case(np.load(golden / "smoke-q4.npz"))
    _, M, _ = build(KE, shape, E, mask)
    _, it, _, _, mv = mg.fcg(voxelhip.xp, M.levels[0].apply, torch.from_numpy(force).cuda(), M, 1e-8, 999,
                             fixed_iterations=7)
    assert it == 7 and mv == 8


def test_coarse_smoothing_degree_matches_the_oracle(golden, KE):
    """Fewer smoothing steps on the levels from coarse_from down (the donor's coarse_smooth_iters):
    one V-cycle against the oracle with the same per-level counts."""
    g = np.load(golden / "mg_golden.npz")
    shape = tuple(int(s) for s in g["q8s3/shape"])
    E, mask = g["q8s3/E"], g["q8s3/mask"]
    _, M, lam = build(KE, shape, E, mask, smoother="chebyshev", degree=3)
    _, Mo = mg_ref.build_hierarchy(shape, KE, E, mask, max_coarse_x=2, smoother="chebyshev", degree=3)
    assert len(M.levels) >= 3
    r = g["q8s3/vcycle_in"]
    base = M(torch.from_numpy(r).cuda()).cpu().numpy()
    for cf, cs in ((1, 2), (2, 1)):
        M.coarse_sweeps, M.coarse_from = cs, cf
        Mo.coarse_sweeps, Mo.coarse_from = cs, cf
        y = M(torch.from_numpy(r).cuda())
        assert rel(y, Mo(r)) < 1e-10
        assert rel(y, base) > 1e-6                     # the count does change the cycle


def test_chebyshev_vcycle_matches_the_oracle(golden, KE):
    g = np.load(golden / "mg_golden.npz")
    shape = tuple(int(s) for s in g["q8s3/shape"])
    E, mask = g["q8s3/E"], g["q8s3/mask"]
    _, M, lam = build(KE, shape, E, mask, smoother="chebyshev", degree=3)
    _, Mo = mg_ref.build_hierarchy(shape, KE, E, mask, max_coarse_x=2, smoother="chebyshev", degree=3)
    assert np.allclose(lam, Mo.lambdas, rtol=1e-12)
    r = g["q8s3/vcycle_in"]
    assert rel(M(torch.from_numpy(r).cuda()), Mo(r)) < 1e-10
    for variant in ("donor",):
        M.cheb_variant = variant
        Mo.cheb_variant = variant
        assert rel(M(torch.from_numpy(r).cuda()), Mo(r)) < 1e-10


@pytest.mark.parametrize("adt,vdt,n", [(torch.float64, torch.float64, 131), (torch.float64, torch.float32, 4131),
                                       (torch.float32, torch.float32, 517), (torch.float64, torch.float64, 64)])
def test_dense_mv_matches_torch(adt, vdt, n):
    """Coarsest solve kernel: padded rows (the padding is never read: NaN there must not leak),
    FP64 matrix with FP32 vectors, masked rows."""
    from voxelhip import _lib as L
    from voxelhip import ops
    g = torch.Generator(device="cuda").manual_seed(n)
    A = torch.randn(n, n, generator=g, device="cuda", dtype=torch.float64)
    r = torch.randn(n, generator=g, device="cuda", dtype=torch.float64).to(vdt)
    m = torch.rand(n, generator=g, device="cuda") > 0.2
    mask = ops.Mask(L.MASK_ARRAY, m.to(torch.uint8).contiguous())
    Ap = ops.padded_matrix(A, adt)
    Ap[:, n:] = float("nan")
    y = torch.empty(n, dtype=vdt, device="cuda")
    ops.dense_mv(Ap, r, y, mask=mask)
    ref = (A.to(adt).double() @ r.double()) * m.double()
    tol = 1e-13 if (adt, vdt) == (torch.float64, torch.float64) else (1e-6 if adt == torch.float64 else 1e-5)
    assert float((y.double() - ref).norm() / ref.norm()) < tol
    assert torch.all(y[~m] == 0)


@pytest.mark.parametrize("zdt,vdt", [(torch.float32, torch.float32), (torch.float32, torch.float64),
                                     (torch.float64, torch.float64)])
def test_cvt_dotdiff_matches_torch(zdt, vdt):
    from voxelhip import ops
    n = 100003
    g = torch.Generator(device="cuda").manual_seed(7)
    zin, r, zold = (torch.randn(n, generator=g, device="cuda", dtype=torch.float64) for _ in range(3))
    zin, r, zold = zin.to(zdt), r.to(vdt), zold.to(vdt)
    znew = torch.empty(n, dtype=vdt, device="cuda")
    parts = ops.reduce_parts(n)
    partial = torch.zeros(2 * parts, dtype=torch.float64, device="cuda")
    ops.cvt_dotdiff(zin, r, zold, znew, partial)
    zc = zin.to(vdt).double()
    assert torch.equal(znew, zin.to(vdt))
    s0, s1 = float(partial[:parts].sum()), float(partial[parts:].sum())
    assert s0 == pytest.approx(float(r.double() @ zc), rel=1e-12)
    assert s1 == pytest.approx(float(r.double() @ (zc - zold.double())), rel=1e-12)


def test_fp64_coarsest_inside_the_fp32_vcycle(golden, KE):
    """coarse_dtype=float64: the FP32 V-cycle with its coarsest solve in FP64 (FP64 inverse, FP32
    vectors) stays within FP32 rounding of the reference's FP32 and FP64 V-cycles."""
    g = np.load(golden / "mg_golden.npz")
    shape = tuple(int(s) for s in g["q8s3/shape"])
    _, M, _ = build(KE, shape, g["q8s3/E"], g["q8s3/mask"], dtype=np.float32, coarse_dtype=np.float64)
    assert M.coarse_inverse.dtype == torch.float64
    r = torch.from_numpy(g["q8s3/vcycle_in"]).cuda()
    out = M(r)
    assert rel(out, g["q8s3/vcycle32_out"]) < 1e-5 and rel(out, g["q8s3/vcycle_out"]) < 1e-5


@pytest.mark.parametrize("kind", ["none", "x0", "array"])
def test_device_dense_assembly_is_bitwise_the_oracle(KE, kind):
    from voxelhip import ops
    shape = (4, 3, 5)
    rng = np.random.default_rng(21)
    E = rng.uniform(0.05, 1.0, int(np.prod(shape)))
    mask = np.ones(3 * hx.n_nodes(shape))
    if kind == "x0":
        mask[: 3 * 4 * 6] = 0
    elif kind == "array":
        mask = (rng.random(mask.size) > 0.25).astype(float)
    M = ops.Mask.from_values(mask, shape)
    ke = torch.from_numpy(np.ascontiguousarray(KE, dtype=np.float64)).cuda()
    A = ops.assemble_dense(ke, torch.from_numpy(E).cuda(), shape, mask=M)
    assert np.array_equal(A.cpu().numpy(), hx.assemble_dense(shape, KE, E, mask))


@pytest.mark.parametrize("kind", ["x0", "array"])
@pytest.mark.parametrize("fine", [(4, 2, 2), (8, 4, 6), (6, 8, 4)])
def test_device_galerkin_operator_matches_the_oracle(KE, kind, fine):
    from voxelhip import ops
    from voxelhip.engine import galerkin_coarsest
    rng = np.random.default_rng(22)
    E = rng.uniform(0.01, 1.0, int(np.prod(fine)))
    mask = np.ones(3 * hx.n_nodes(fine))
    if kind == "x0":
        mask[: 3 * (fine[1] + 1) * (fine[2] + 1)] = 0
    else:
        mask = (rng.random(mask.size) > 0.2).astype(float)
    print("case", kind, fine, flush=True)
    A = galerkin_coarsest(fine, 2 * KE, torch.from_numpy(E).cuda(), ops.Mask.from_values(mask, fine))
    ref = mg_ref.galerkin_dense(fine, 2 * KE, E, mask)
    assert A.shape == ref.shape
    assert np.abs(A.cpu().numpy() - ref).max() <= 1e-13 * np.abs(ref).max()


def test_galerkin_coarsest_solve(KE):
    from voxelhip.engine import GMGSolver
    shape = (32, 16, 16)
    _, E, mask, force = sr.cantilever(shape, seed=6, rmin=1.5)
    f = torch.from_numpy(force).cuda()
    out = {}
    for coarse in ("rediscretize", "galerkin"):
        print("case", coarse, flush=True)
        s = GMGSolver(shape, KE, mask, torch.from_numpy(E).cuda(), max_coarse_x=8, smoother="chebyshev", degree=3,
                      precision="fp64", coarse=coarse)
        x, it, ok, r, _ = s.solve(f)
        assert ok and r <= 1e-8
        out[coarse] = (x, it)
    # same solution; the Galerkin coarsest is the better coarse operator for varying moduli
    assert float((out["galerkin"][0] - out["rediscretize"][0]).norm() / out["rediscretize"][0].norm()) < 1e-7
    assert out["galerkin"][1] <= out["rediscretize"][1] + 1


# ------------------------------------------------------------------ full Galerkin hierarchy
def _galerkin_case(kind, shapes, seed=12):
    rng = np.random.default_rng(seed)
    n0 = 3 * hx.n_nodes(shapes[0])
    E = rng.uniform(0.01, 1.0, int(np.prod(shapes[0])))
    mask = np.ones(n0)
    if kind == "x0":
        mask[: 3 * (shapes[0][1] + 1) * (shapes[0][2] + 1)] = 0
    elif kind == "array":
        mask = (rng.random(n0) > 0.15).astype(float)
    return E, mask


@pytest.mark.parametrize("kind", ["none", "x0", "array"])
def test_galerkin_stencils_match_the_oracle_hierarchy(KE, kind):
    """Levels 1 and 2 as assembled stencils and the dense coarsest of level 3 against the dense
    recursion A_l = M (P^T A_(l-1) P) M + (I - M) of the oracle."""
    from voxelhip import ops
    from voxelhip.engine import galerkin_hierarchy
    shapes = [(16, 8, 8), (8, 4, 4), (4, 2, 2), (2, 1, 1)]
    E, mask = _galerkin_case(kind, shapes)
    print("case", kind, flush=True)
    K, A = mg_ref.galerkin_hierarchy(shapes, KE, E, mask)
    stencils, Ac = galerkin_hierarchy(shapes, KE, ops.Mask.from_values(mask, shapes[0]), torch.from_numpy(E).cuda())
    assert len(stencils) == 2
    rng = np.random.default_rng(3)
    m = mask
    for lvl, sp in ((1, stencils[0]), (2, stencils[1])):
        m = mg_ref.coarsen_mask(m, shapes[lvl - 1])
        mo = ops.Mask.from_values(m, shapes[lvl])
        u = rng.standard_normal(A[lvl].shape[0])
        y = sp.raw(torch.from_numpy(u).cuda(), mask=mo)
        assert rel(y, A[lvl] @ u) < 1e-13, lvl
        # the unmasked operator, and its diagonal
        assert rel(sp.raw(torch.from_numpy(u).cuda()), K[lvl - 1] @ u) < 1e-13, lvl
        assert rel(sp.diagonal(), np.diag(K[lvl - 1])) < 1e-14
    assert np.abs(Ac.cpu().numpy() - A[3]).max() <= 1e-12 * np.abs(A[3]).max()


@pytest.mark.parametrize("kind", ["x0", "array"])
def test_stencil_epilogues_match_their_formulas(KE, kind):
    import voxel_mg as mg
    from voxelhip import _lib as L
    from voxelhip import ops
    from voxelhip.engine import galerkin_hierarchy
    shapes = [(12, 8, 8), (6, 4, 4), (3, 2, 2)]
    E, mask = _galerkin_case(kind, shapes, seed=4)
    K, A = mg_ref.galerkin_hierarchy(shapes, KE, E, mask)
    (sp,), _ = galerkin_hierarchy(shapes, KE, ops.Mask.from_values(mask, shapes[0]), torch.from_numpy(E).cuda())
    m = mg_ref.coarsen_mask(mask, shapes[0])
    mo = ops.Mask.from_values(m, shapes[1])
    A1, K1 = A[1], K[0]
    n = A1.shape[0]
    rng = np.random.default_rng(8)
    u, b, pv = (rng.standard_normal(n) for _ in range(3))
    dinv = np.where(m != 0, 1.0 / np.diag(K1), 1.0)
    assert rel(sp.diagonal(mask=mo, invert=True), dinv) < 1e-15
    t = lambda v: torch.from_numpy(np.ascontiguousarray(v)).cuda()  # noqa: E731
    Au = A1 @ u
    y = torch.empty(n, dtype=torch.float64, device="cuda")
    y32 = torch.empty(n, dtype=torch.float32, device="cuda")
    print("case", kind, "resid", flush=True)
    ops.stencil_apply(sp.U, t(u), y, shapes[1], mask=mo, mode=L.MODE_RESID, b=t(b), y32=y32)
    assert rel(y, b - Au) < 1e-13 and rel(y32.double(), b - Au) < 1e-7
    ops.stencil_apply(sp.U, t(u), y, shapes[1], mask=mo, mode=L.MODE_RESID, b=t(b), fixed_zero=True)
    assert rel(y, (b - Au) * m) < 1e-13
    ops.stencil_apply(sp.U, t(u), y, shapes[1], mask=mo, fixed_zero=True)
    assert rel(y, m * (K1 @ (m * u))) < 1e-13
    print("case", kind, "jacobi/cheb3", flush=True)
    ops.stencil_apply(sp.U, t(u), y, shapes[1], mask=mo, mode=L.MODE_JACOBI, b=t(b), dinv=t(dinv), omega=0.6)
    assert rel(y, u + 0.6 * dinv * (b - Au)) < 1e-13
    ops.stencil_apply(sp.U, t(u), y, shapes[1], mask=mo, mode=L.MODE_CHEB3, b=t(b), dinv=t(dinv), cheb_a=0.7,
                      cheb_c=0.2, prev=t(pv))
    assert rel(y, u + 0.7 * (dinv * (b - Au) + 0.2 * (u - pv))) < 1e-13
    y.copy_(t(pv))                    # y holds the previous iterate on entry
    ops.stencil_apply(sp.U, t(u), y, shapes[1], mask=mo, mode=L.MODE_CHEB3, b=t(b), dinv=t(dinv), cheb_a=0.7,
                      cheb_c=0.2)
    assert rel(y, u + 0.7 * (dinv * (b - Au) + 0.2 * (u - pv))) < 1e-13
    ops.stencil_apply(sp.U, t(u), y, shapes[1], mask=mo, mode=L.MODE_CHEB3, b=t(b), dinv=t(dinv), cheb_a=0.7,
                      cheb_c=0.2, prev_zero=True)
    assert rel(y, u + 0.7 * (dinv * (b - Au) + 0.2 * u)) < 1e-13
    # the FP32 copy of the stencil
    s32 = sp.astype(torch.float32)
    y32b = s32.raw(t(u).float(), mask=mo)
    assert rel(y32b.double(), Au) < 1e-5
    with pytest.raises(Exception):
        ops.stencil_apply(sp.U, t(u), y, shapes[1], mask=mo, mode=L.MODE_DOT)
    assert isinstance(mg.StencilProduct(shapes[1], sp.U), mg.StencilProduct)


def test_galerkin_hierarchy_vcycle_matches_the_oracle(KE):
    """One Chebyshev V-cycle over Galerkin levels against the oracle's Multigrid with the dense
    Galerkin operators (same lambdas)."""
    import voxel_mg as mg
    from voxelhip import ops
    from voxelhip.engine import build_hierarchy, galerkin_hierarchy
    shapes = [(16, 8, 8), (8, 4, 4), (4, 2, 2), (2, 1, 1)]
    E, mask = _galerkin_case("x0", shapes, seed=9)
    K, A = mg_ref.galerkin_hierarchy(shapes, KE, E, mask)
    E64 = torch.from_numpy(E).cuda()
    mo = ops.Mask.from_values(mask, shapes[0])
    stencils, Ac = galerkin_hierarchy(shapes, KE, mo, E64)
    _, M = build_hierarchy(shapes[0], KE, mo, E64, shapes, torch.float64, smoother="chebyshev", degree=3,
                           coarse_matrix=Ac, stencils=stencils)
    prods = [hx.NumpyProduct(shapes[0], KE)] + [mg_ref.DenseProduct(k) for k in K[:-1]] + \
        [hx.NumpyProduct(shapes[-1], KE * 8)]
    Mo = mg_ref.Multigrid(shapes, prods, E, mask, np.linalg.cholesky(A[-1]), np.float64, pre=3, post=3,
                          smoother="chebyshev", lambdas=M.lambdas)
    lam_o = mg_ref.chebyshev_bounds(Mo.levels)
    assert np.allclose(M.lambdas, lam_o, rtol=1e-10)
    r = np.random.default_rng(5).standard_normal(mask.size) * mask
    assert rel(M(torch.from_numpy(r).cuda()), Mo(r)) < 1e-11


def test_device_start_vector_equals_the_oracle_restatement():
    """vh_fill_splitmix (the Chebyshev bound estimate's start vector) reproduces the oracle's
    SplitMix64 restatement bit for bit in FP64, and its FP32 rounding in FP32 (5 M values, so the
    counter passes 2^22 and the products wrap)."""
    from voxelhip import ops
    n, seed = 5_000_003, 20260923
    ref = mg_ref.splitmix_uniform(n, seed)
    y = ops.fill_splitmix(torch.empty(n, dtype=torch.float64, device="cuda"), seed)
    assert np.array_equal(y.cpu().numpy(), ref)
    y32 = ops.fill_splitmix(torch.empty(n, dtype=torch.float32, device="cuda"), seed)
    assert np.array_equal(y32.cpu().numpy(), ref.astype(np.float32))


def test_coarsest_inverse_and_an_indefinite_coarsest_operator(KE):
    """The hierarchy's coarsest inverse equals the host inverse of the same operator, and an
    operator that is not positive definite raises LinAlgError."""
    from voxelhip import ops
    from voxelhip.engine import build_hierarchy
    shapes = [(16, 8, 8), (8, 4, 4), (4, 2, 2)]
    rho, E, mask, _ = sr.cantilever(shapes[0], seed=3, rmin=1.5)
    E64 = torch.from_numpy(E).cuda()
    mo = ops.Mask.from_values(mask, shapes[0])
    n = 3 * ops.n_nodes(shapes[-1])
    rng = np.random.default_rng(4)
    B = rng.standard_normal((n, n))
    A = B @ B.T + n * np.eye(n)                       # symmetric positive definite
    _, M = build_hierarchy(shapes[0], KE, mo, E64, shapes, torch.float64, smoother="chebyshev", degree=3,
                           coarse_matrix=torch.from_numpy(A).cuda())
    inv = M.coarse_inverse[:, :n].double().cpu().numpy()
    assert np.abs(inv - np.linalg.inv(A)).max() <= 1e-12 * np.abs(np.linalg.inv(A)).max()
    A[5, 5] = -1.0                                    # leading minor 6 negative
    with pytest.raises(torch.linalg.LinAlgError):
        build_hierarchy(shapes[0], KE, mo, E64, shapes, torch.float64, smoother="chebyshev", degree=3,
                        coarse_matrix=torch.from_numpy(A).cuda())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_full_galerkin_solve(KE, precision):
    from voxelhip.engine import GMGSolver
    shape = (32, 16, 16)
    rho, E, mask, force = sr.cantilever(shape, seed=6, rmin=1.5)
    # a high-contrast design: void below rho = 0.5
    E = np.where(rho > 0.5, E, 1e-6)
    f = torch.from_numpy(force).cuda()
    out = {}
    for coarse in ("rediscretize", "galerkin-all"):
        print("case", precision, coarse, flush=True)
        s = GMGSolver(shape, KE, mask, torch.from_numpy(E).cuda(), max_coarse_x=4, smoother="chebyshev", degree=3,
                      precision=precision, coarse=coarse)
        x, it, ok, r, _ = s.solve(f, tol=1e-8, maxiter=400)
        assert ok and r <= 1e-8, (coarse, it, r)
        out[coarse] = (x, it)
    assert float((out["galerkin-all"][0] - out["rediscretize"][0]).norm() / out["rediscretize"][0].norm()) < 1e-6
    assert out["galerkin-all"][1] <= out["rediscretize"][1], (out["galerkin-all"][1], out["rediscretize"][1])
