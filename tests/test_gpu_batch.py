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
