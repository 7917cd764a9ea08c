"""GPU solutions replayed by the reference's own C++ oracle, as the reference verifies its runs
(experiments/session_a/verify.py:15-20,58-83 with experiments/session_a/inputs/reference.cc:16-39,
compiled by oracle/Makefile into oracle/_ref/).

For each solution x the natural-layout C++ product gives y = M K(rho) M x + (I - M) x and the
element energies ce of the masked x; the replay computes, as verify.py does, the independent true
residual ||f - y|| / ||f||, the compliance sum(E ce) and its agreement with the force work f . x.
Checks: the independent residual agrees with the GPU solver's own residual by verify.py's rule
(|difference| <= 1e-12 + 1e-5 of it) and meets the requested 1e-8 (1.01e-8 for rounding of the
two products); compliance and force work agree to 1e-8 (the reference's limit is 1e-6); x is
finite and zero at every fixed DOF.

Solutions: the session-D multigrid solves of optimized-q64 (FP64 and FP32 V-cycle, Emin 1e-6),
the mixed-precision and FP64 GMG-FCG solves of the 256x128x128 cantilever that bench.py times
(Emin 1e-9), and ten instances of the 48-problem batch."""
import numpy as np
import pytest
import torch

from oracle import cref
from oracle import hexref as hx

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not cref.available(), reason="oracle/_ref not built (needs /root/reference)")]


def _replay(shape, rho, x, force, mask, ke, emin, gpu_rel=None):
    """verify.py's physical checks of the author-layout solution x (host arrays)."""
    nat = lambda a: hx.to_natural(np.asarray(a, dtype=np.float64), shape, nodal=True)  # noqa: E731
    f, xn, mn = nat(force), nat(x), nat(mask)
    rho_n = hx.to_natural(np.asarray(rho, dtype=np.float64), shape)
    y, ce, _ = cref.Reference().product(shape, rho_n, xn, ke, mn, emin)
    rel = float(np.linalg.norm(f - y) / np.linalg.norm(f))
    E = emin + (1.0 - emin) * rho_n ** 3
    compliance, work = float(E @ ce), float(f @ xn)
    print(f"{shape}: independent residual {rel:.4e} (GPU {gpu_rel}), compliance {compliance:.12g}, "
          f"energy/work - 1 = {compliance / work - 1:.2e}", flush=True)
    assert np.isfinite(xn).all() and np.all(xn[mn == 0] == 0)
    assert rel <= 1.01e-8
    if gpu_rel is not None:
        assert abs(rel - gpu_rel) <= 1e-12 + 1e-5 * rel
    assert abs(compliance / work - 1) <= 1e-8
    return rel, compliance


@pytest.mark.parametrize("vcycle", [np.float64, np.float32])
def test_session_d_solutions_replayed_by_the_cpp_oracle(golden, vcycle):
    from voxelhip import harness
    print("case replay session D", vcycle.__name__, flush=True)
    data = harness.load_case(golden, "q64x1")
    rec, x = harness.solve_case(data, "q64x1", vcycle_dtype=vcycle)
    shape = tuple(int(s) for s in data["shape"])
    _, c = _replay(shape, data["rho"], x.cpu().numpy(), data["force"], data["mask"], data["ke"],
                   harness.PROTOCOL["material"]["Emin"], rec["native_true_residual"])
    assert c == pytest.approx(rec["compliance"], rel=1e-10)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_benchmark_solutions_replayed_by_the_cpp_oracle(precision):
    """The 256x128x128 solve of bench.py (seed 3), with its solver settings."""
    import voxel_kernels as vk
    from voxelhip.engine import GMGSolver
    from voxelhip.simp import ConeFilter, modulus
    print("case replay benchmark", precision, flush=True)
    shape = (256, 128, 128)
    nx, ny, nz = shape
    rng = np.random.default_rng(3)
    rho = ConeFilter(shape, 1.5, device="cuda").forward(torch.from_numpy(rng.uniform(0.01, 1.0, nx * ny * nz)).cuda())
    E = modulus(rho, 3.0, 1.0, 1e-9)
    f = torch.zeros((nx + 1, ny + 1, nz + 1, 3), dtype=torch.float64, device="cuda")
    line = torch.full((ny + 1,), -1.0 / ny, dtype=torch.float64, device="cuda")
    line[0] = line[-1] = -0.5 / ny
    f[nx, :, nz // 2, 2] = line
    m = torch.ones_like(f)
    m[0] = 0
    f, m = f.reshape(-1), m.reshape(-1)
    ke = 2 * vk.quadrature([1, 1, 1], 0.3)
    s = GMGSolver(shape, ke, m, E, smoother="chebyshev", degree=3, precision=precision, lower_frac=1.0 / 15.0,
                  coarse_degree=2, coarse_level=2)
    x, it, ok, rel, _ = s.solve(f, tol=1e-8)
    assert ok and it <= 11
    _replay(shape, rho.cpu().numpy(), x.cpu().numpy(), f.cpu().numpy(), m.cpu().numpy(), ke, 1e-9, rel)


def test_batched_solutions_replayed_by_the_cpp_oracle():
    """The first ten instances of the 48-problem batch workload (mixed precision)."""
    import voxel_kernels as vk
    from voxelhip.batch import BatchedGMGSolver, random_cantilevers
    print("case replay batch", flush=True)
    ke = 2 * vk.quadrature([1, 1, 1], 0.3)
    problems = random_cantilevers(seed=4, n=48)
    bs = BatchedGMGSolver(problems, ke, precision="mixed", graph=True, lower_frac=1.0 / 15.0)
    xs, its, ok, rel = bs.solve(tol=1e-8, maxiter=200)
    assert ok.all()
    k = 0
    for pb in problems:
        # E = 1e-9 + (1 - 1e-9) rho^3: the densities behind the batch's moduli
        rho = ((pb.E.detach().double().cpu().numpy() - 1e-9) / (1 - 1e-9)) ** (1.0 / 3.0)
        for f in pb.loads:
            if k >= 10:
                return
            _replay(pb.shape, rho, xs[k].cpu().numpy(), f, np.asarray(pb.mask, dtype=np.float64), ke, 1e-9,
                    float(rel[k]))
            k += 1
