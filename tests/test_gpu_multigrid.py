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
