"""GPU checks of the launch tuner: every (tile, lx) candidate computes the same product bit for
bit (the element arithmetic and the gather order do not depend on the launch plan), tuned
plans are valid candidates and persist through the cache file, and a tuned solve matches the
untuned one (only the order of the per-tile partial sums of the dots differs)."""
import numpy as np
import pytest
import torch

from oracle import hexref as hx
from oracle import simpref as sr

pytestmark = pytest.mark.gpu

KE = 2 * hx.quadrature([1, 1, 1], 0.3)


@pytest.mark.parametrize("dtype,shape", [(torch.float64, (24, 20, 36)), (torch.float32, (24, 20, 36)),
                                         (torch.float64, (10, 35, 101)), (torch.float32, (6, 67, 40)),
                                         (torch.float64, (8, 20, 70))])
def test_every_launch_plan_gives_the_same_product(dtype, shape):
    import voxel_kernels as vk
    from voxelhip import ops
    from voxelhip.tuner import _candidates
    g = torch.Generator(device="cuda").manual_seed(3)
    E = torch.rand(int(np.prod(shape)), generator=g, device="cuda", dtype=torch.float64).to(dtype) + 0.01
    u = torch.randn(3 * ops.n_nodes(shape), generator=g, device="cuda", dtype=torch.float64).to(dtype)
    blocks = vk.Modal(KE).blocks
    ref = ops.matvec(blocks, E, u, torch.empty_like(u), shape)
    for t, lx in _candidates(shape[0], int(np.prod(shape))):
        print("case", dtype, shape, t, lx, flush=True)
        y = ops.matvec(blocks, E, u, torch.empty_like(u), shape, tile=t, lx=lx)
        assert torch.equal(y, ref), (t, lx)


def test_tuner_plans_and_cache(tmp_path):
    import voxel_mg as mg
    from voxelhip import _lib as L
    from voxelhip.tuner import LaunchTuner, _candidates
    shape = (32, 16, 16)
    _, E, mask, _ = sr.cantilever(shape, seed=2)
    lvl = mg.Level(None, shape, mg.HipProduct(shape, KE), E, mask, np.float64)
    path = tmp_path / "tune.json"
    tuner = LaunchTuner(path=path, reps=2, rounds=1)
    lvl.tune(tuner, [(L.MODE_CHEB3, False), (L.MODE_DOT, True)])
    for key in ((L.MODE_CHEB3, False), (L.MODE_DOT, True)):
        assert tuple(lvl.launch[key]) in _candidates(shape[0], int(np.prod(shape)))
    assert path.exists()
    again = LaunchTuner(path=path)
    assert again.cache == tuner.cache
    # a cached plan is returned without timing
    lvl2 = mg.Level(None, shape, mg.HipProduct(shape, KE), E, mask, np.float64)
    lvl2.tune(again, [(L.MODE_CHEB3, False)])
    assert lvl2.launch[(L.MODE_CHEB3, False)] == lvl.launch[(L.MODE_CHEB3, False)] and not again.timings


@pytest.mark.parametrize("precision", ["fp64", "mixed", "ir"])
def test_tuned_solve_matches_the_untuned_solve(precision):
    from voxelhip.engine import GMGSolver
    shape = (32, 16, 16)
    _, E, mask, force = sr.cantilever(shape, seed=5, rmin=1.5)
    f = torch.from_numpy(force).cuda()
    out = []
    for tune in (False, True):
        print("case", precision, "tune", tune, flush=True)
        s = GMGSolver(shape, KE, mask, torch.from_numpy(E).cuda(), max_coarse_x=8, smoother="chebyshev", degree=3,
                      precision=precision, graph=True, tune=tune)
        x, it, ok, r, _ = s.solve(f)
        assert ok and r <= 1e-8
        out.append((x.clone(), it))
    assert out[0][1] == out[1][1]
    assert float((out[0][0] - out[1][0]).norm() / out[0][0].norm()) < 1e-9
