"""GPU parity of the batched solver: problems of different grid sizes and hierarchy depths, with
1-3 load cases each, solved together (one launch per pass) must reproduce the single-problem
GMG-FCG engine on every (problem, load case): same iteration count and the same solution up to
the rounding of differently ordered dot products (FP64: 1e-9 relative; mixed FP32 V-cycle:
1e-7, as the single engine against the oracle), and every instance's solution must meet the
tolerance on the oracle's constrained residual ||b - A x|| / ||b|| (NumPy product, 1.01e-8).
The full 48-problem batch workload (131 load cases, 16^3 ... 96x48x48) is checked instance by
instance with the dense 24x24 node kernel, an arithmetic independent of the parity product."""
import numpy as np
import pytest
import torch

from oracle import hexref as hx
from oracle import simpref as sr

pytestmark = pytest.mark.gpu

KE = 2 * hx.quadrature([1, 1, 1], 0.3)


def point_loads(shape, n_loads, rng):
    """Unit point loads at random nodes of random faces other than the clamped x = 0 face."""
    nx, ny, nz = shape
    loads = []
    for _ in range(n_loads):
        f = np.zeros((nx + 1, ny + 1, nz + 1, 3))
        face = rng.integers(0, 5)
        ix, iy, iz = rng.integers(1, nx + 1), rng.integers(0, ny + 1), rng.integers(0, nz + 1)
        if face == 0: ix = nx
        elif face == 1: iy = 0
        elif face == 2: iy = ny
        elif face == 3: iz = 0
        else: iz = nz
        d = rng.normal(size=3)
        f[ix, iy, iz] = d / np.linalg.norm(d)
        loads.append(f.reshape(-1))
    return loads


def _oracle_residual(pb, x, f):
    prod = hx.NumpyProduct(pb.shape, KE)
    m = np.asarray(pb.mask, dtype=np.float64)
    E = pb.E.detach().cpu().numpy() if isinstance(pb.E, torch.Tensor) else np.asarray(pb.E)
    u = x.detach().cpu().numpy()
    Au = prod.raw(u * m, E) * m + u * (1 - m)
    return float(np.linalg.norm(f - Au) / np.linalg.norm(f))


@pytest.fixture(scope="module")
def batch_problems():
    from voxelhip.batch import BatchProblem
    rng = np.random.default_rng(11)
    out = []
    for shape, nl, seed in (((16, 8, 8), 2, 1), ((24, 12, 12), 1, 2), ((8, 8, 8), 3, 3), ((32, 16, 8), 2, 4)):
        _, E, mask, _ = sr.cantilever(shape, seed=seed, rmin=1.5)
        out.append(BatchProblem(shape, E, mask, point_loads(shape, nl, rng)))
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_solve_matches_single_problem_solves(batch_problems, precision):
    from voxelhip.batch import BatchedGMGSolver, coarse_dof_shapes
    from voxelhip.engine import GMGSolver
    print("case batch", precision, flush=True)
    bs = BatchedGMGSolver(batch_problems, KE, max_coarse_dofs=600, precision=precision, graph=True)
    assert len(set(bs.depth)) > 1                      # the batch mixes hierarchy depths
    xs, its, ok, rel = bs.solve(tol=1e-8, maxiter=100)
    assert ok.all() and (rel <= 1e-8).all()
    k = 0
    for pb in batch_problems:
        shapes = coarse_dof_shapes(pb.shape, 600)
        s = GMGSolver(pb.shape, KE, pb.mask, torch.from_numpy(pb.E).cuda(), smoother="chebyshev", degree=3,
                      precision=precision, n_levels=len(shapes), max_coarse_x=0)
        assert s.shapes == shapes
        for f in pb.loads:
            print("case instance", k, pb.shape, flush=True)
            x, it, conv, r, _ = s.solve(torch.from_numpy(f).cuda(), tol=1e-8, maxiter=100)
            assert conv and it == its[k], (k, it, its[k])
            tol = 1e-9 if precision == "fp64" else 1e-7
            assert float((xs[k] - x).norm() / x.norm()) < tol
            assert _oracle_residual(pb, xs[k], f) <= 1.01e-8
            k += 1


def test_batched_graph_replay_equals_eager(batch_problems):
    from voxelhip.batch import BatchedGMGSolver
    out = []
    for graph in (False, True):
        bs = BatchedGMGSolver(batch_problems, KE, max_coarse_dofs=600, precision="mixed", graph=graph)
        for _ in range(2):
            xs, its, ok, _ = bs.solve(tol=1e-8, maxiter=100)
        out.append((torch.cat([x.clone() for x in xs]), its.copy()))
    assert torch.equal(out[0][0], out[1][0]) and (out[0][1] == out[1][1]).all()


def test_full_batch_workload_meets_the_tolerance_on_an_independent_residual():
    """48 problems from seed 4 (16^3 ... 96x48x48, 131 load cases, 24 M DOF), mixed precision;
    two solves in a row (the loads stay resident) give the same iterates."""
    from voxelhip import ops
    from voxelhip.batch import BatchedGMGSolver, random_cantilevers
    problems = random_cantilevers(seed=4, n=48)
    bs = BatchedGMGSolver(problems, KE, precision="mixed", graph=True, lower_frac=1.0 / 15.0)
    print("case full batch", bs.n, "instances", flush=True)
    xs, its, ok, rel = bs.solve(tol=1e-8, maxiter=200)
    assert bs.n == 131 and ok.all() and (rel <= 1e-8).all()
    first = [x.clone() for x in xs]
    xs2, its2, ok2, _ = bs.solve(tol=1e-8, maxiter=200)
    assert (its2 == its).all() and ok2.all()
    k = 0
    worst = 0.0
    for pb in problems:
        mask = torch.from_numpy(np.asarray(pb.mask, dtype=np.float64)).cuda()
        for f in pb.loads:
            assert torch.equal(xs2[k], first[k])
            x = xs[k] * mask
            y = torch.empty_like(x)
            ops.dense_product(KE, pb.E, x, y, pb.shape)
            Ax = y * mask + xs[k] * (1 - mask)
            fd = torch.from_numpy(f).cuda()
            r = float((fd - Ax).norm() / fd.norm())
            worst = max(worst, r)
            assert r <= 1.01e-8, (k, pb.shape, r)
            k += 1
    print("worst independent residual", worst, "iterations", its.min(), its.max(), flush=True)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_batched_product_equals_single_products(dtype):
    """One batched launch over problems of different sizes (x extents 8 ... 96, so that the
    batch-wide x chunking differs from each problem's own plan), for the raw product and the
    three-term Chebyshev step: the three batched tiles (16 x 16, 16 x 32, 8 x 32 nodes) with the
    automatic chunk length or fixed ones (1, 5) give bitwise the same result (the element
    arithmetic and the gather order do not depend on the plan), and it
    matches the single-problem kernel to rounding (the two kernels are compiled separately, and
    the compiler may fuse a different set of multiply-adds: 1e-13 of max |y| in FP64, 1e-5 in
    FP32)."""
    from voxelhip import _lib as L
    from voxelhip import ops
    from voxelhip.batch import _epilogue, product_batch
    print("case batched product", dtype, flush=True)
    g = torch.Generator(device="cuda").manual_seed(5)
    shapes = [(96, 16, 24), (8, 8, 8), (40, 24, 40), (16, 48, 16), (72, 8, 16)]
    blocks = np.ascontiguousarray(ops.modal_blocks(KE)[0], dtype=np.float64)
    keep = blocks.reshape(72)
    tol = 1e-13 if dtype == torch.float64 else 1e-5
    for mode in (L.MODE_RAW, L.MODE_CHEB3):
        data = []
        for s in shapes:
            nd = 3 * (s[0] + 1) * (s[1] + 1) * (s[2] + 1)
            E = (0.01 + torch.rand(s[0] * s[1] * s[2], generator=g, device="cuda", dtype=torch.float64)).to(dtype)
            u = torch.randn(nd, generator=g, device="cuda", dtype=torch.float64).to(dtype)
            b = torch.randn(nd, generator=g, device="cuda", dtype=torch.float64).to(dtype)
            prev = torch.randn(nd, generator=g, device="cuda", dtype=torch.float64).to(dtype)
            ep = dict(b=b, kdiag=np.diag(KE)[:3], cheb_a=0.7, cheb_c=0.3) if mode == L.MODE_CHEB3 else {}
            y_ref = prev.clone()
            ops.matvec(blocks, E, u, y_ref, s, mask=ops.Mask(L.MASK_X0), mode=mode, **ep)
            data.append((s, E, u, prev, ep, y_ref))
        results = []
        for tile, lx in ((L.TILE_16x16, 0), (L.TILE_16x16, 1), (L.TILE_16x32, 5), (L.TILE_16x32, 0), (L.TILE_8x32, 0)):
            items, outs = [], []
            for s, E, u, prev, ep, _ in data:
                y = prev.clone()                 # CHEB3 reads the previous iterate from y
                it = L.ProductItem()
                it.blocks72 = keep.ctypes.data
                it.E, it.u, it.y = E.data_ptr(), u.data_ptr(), y.data_ptr()
                it.nx, it.ny, it.nz = s
                it.mask = None
                it.ep = _epilogue(mode, lx=lx, **{k: (v.data_ptr() if isinstance(v, torch.Tensor) else v)
                                                  for k, v in ep.items()})
                items.append(it)
                outs.append(y)
            h, _ = product_batch(dtype, L.MASK_X0, tile, items)
            h.launch()
            torch.cuda.synchronize()
            results.append(outs)
        for k, (s, *_, y_ref) in enumerate(data):
            for outs in results[1:]:
                assert torch.equal(outs[k], results[0][k]), (mode, s)
            err = float((results[0][k] - y_ref).abs().max() / y_ref.abs().max())
            assert err <= tol, (mode, s, err)


def test_batched_coarsest_solves_group_the_load_cases_of_a_matrix():
    """Consecutive items that share a matrix are applied in one pass over it, up to four at a time
    (six items of one matrix split 4 + 2): bitwise the result of each item in a handle of its own,
    and the single-vector vh_dense_mv to FP32 rounding."""
    from voxelhip import _lib as L
    from voxelhip import ops
    from voxelhip.batch import dense_batch
    print("case batched coarsest solves", flush=True)
    g = torch.Generator(device="cuda").manual_seed(9)
    mats = []
    for n in (203, 96):
        ld = (n + 3) // 4 * 4
        A = torch.zeros(n, ld, device="cuda", dtype=torch.float32)
        A[:, :n] = torch.randn(n, n, generator=g, device="cuda")
        mask = (torch.rand(n, generator=g, device="cuda") > 0.1).to(torch.uint8)
        mats.append((A, n, ld, mask))
    counts = (6, 3)
    rhs = [[torch.randn(m[1], generator=g, device="cuda") for _ in range(c)] for m, c in zip(mats, counts)]

    def run(order):
        items, outs = [], []
        for mi, j in order:
            A, n, ld, mask = mats[mi]
            y = torch.full((n,), float("nan"), device="cuda")
            it = L.DenseItem()
            it.A, it.ld, it.n = A.data_ptr(), ld, n
            it.r, it.y, it.mask = rhs[mi][j].data_ptr(), y.data_ptr(), mask.data_ptr()
            items.append(it)
            outs.append(((mi, j), y))
        dense_batch(torch.float32, torch.float32, items).launch()
        torch.cuda.synchronize()
        return dict(outs)

    grouped = run([(0, j) for j in range(6)] + [(1, j) for j in range(3)])
    for key, y in grouped.items():
        alone = run([key])[key]                    # a handle of one item: one pass over A for it alone
        assert torch.equal(y, alone), key
        A, n, ld, mask = mats[key[0]]
        ref = ops.dense_mv(A, rhs[key[0]][key[1]], torch.empty(n, device="cuda"), ops.Mask(L.MASK_ARRAY, mask), n)
        assert float((y - ref).abs().max() / ref.abs().max()) < 1e-6, key


def test_set_load_replaces_one_instance_before_and_between_solves(batch_problems):
    """set_load before the first solve and between solves: the replaced instance solves the new
    load (same iterate as a batch built with that load), the others are unaffected."""
    from voxelhip.batch import BatchProblem, BatchedGMGSolver
    print("case set_load", flush=True)
    rng = np.random.default_rng(21)
    k, p = 1, 0                                    # instance 1 = problem 0, load case 1
    shape = batch_problems[p].shape
    f_new = point_loads(shape, 1, rng)[0]
    bs = BatchedGMGSolver(batch_problems, KE, max_coarse_dofs=600, precision="mixed", graph=True)
    bs.set_load(k, f_new)
    xs, its, ok, _ = bs.solve(tol=1e-8, maxiter=100)
    swapped = [BatchProblem(pb.shape, pb.E, pb.mask, list(pb.loads)) for pb in batch_problems]
    swapped[p].loads[1] = f_new
    ref = BatchedGMGSolver(swapped, KE, max_coarse_dofs=600, precision="mixed", graph=True)
    xr, itr, okr, _ = ref.solve(tol=1e-8, maxiter=100)
    assert ok.all() and okr.all() and (its == itr).all()
    for a, b in zip(xs, xr):
        assert torch.equal(a, b)
    assert _oracle_residual(batch_problems[p], xs[k], f_new) <= 1.01e-8
    # back to the original load between solves
    bs.set_load(k, batch_problems[p].loads[1])
    xs2, _, ok2, _ = bs.solve(tol=1e-8, maxiter=100)
    assert ok2.all() and _oracle_residual(batch_problems[p], xs2[k], batch_problems[p].loads[1]) <= 1.01e-8
