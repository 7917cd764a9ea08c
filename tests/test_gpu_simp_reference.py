"""The device top88 SIMP loop (voxelhip.simp.simp_run, the loop of the 384x192x192 benchmark)
against the oracle's top88 loop and against the reference's own SIMP records.

Reference records (tests/golden/session_h_known_answers.json, written by
tests/golden/make_session_h.py from results/session-h-host{1,2}-20260926): the donor's
cantilever_gpu_large (120x60x30) and cantilever_gpu_xlarge (160x80x40) presets, density filter
r = 1.5, top88 OC with move 0.15, penal 3, Emin 1e-9, 100 fixed steps.  The reference solved each
state with its GMG-PCG to a relative residual of 1e-5 (warm-started); the state solves here go
to 1e-10 (FP64) or 1e-8 (mixed), so the trajectories agree to the accuracy of the reference's
solves: compliance within 1e-5 and volume within 1e-6 over the first ten steps (step 0, a solve
from zero on the uniform design, within 1e-10), within 3e-5 / 2e-5 over all 100 steps (an OC
bisection may branch differently once in a run when the filtered volume lands within a solve
tolerance of volfrac), the final compliance within the reference's own equivalence limit of
1e-3 of every FP64 run, and the final physical design within 1e-4 (relative L2) of the
reference's stored design (float32).
"""
import json

import numpy as np
import pytest
import torch

from oracle import hexref as hx
from oracle import simpref as sr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def records(golden):
    return json.loads((golden / "session_h_known_answers.json").read_text())


def _case(records, golden, name):
    c = records["cases"][name]
    shape = tuple(c["shape"])
    mask, force = sr.point_load_case(shape, c["lengths"], c["point"], c["load"])
    ke = np.load(golden / "ke_unit_3d.npy")
    stock = [r for r in c["runs"] if r["arm"] == "stock" and r["host"] == "host1" and r["repetition"] == 1][0]
    return shape, ke, mask, force, c, stock


def _check_trajectory(hist, c, stock, steps_full):
    comp = np.array([h["compliance"] for h in hist])
    vol = np.array([h["volume"] for h in hist])
    ch = np.array([h["change"] for h in hist])
    n = len(hist)
    rc = np.array(stock["compliance"][:n])
    rv = np.array(stock["volume"][:n])
    rel_c = np.abs(comp / rc - 1)
    print("compliance rel diff: first ten max %.2e, all max %.2e; volume: first ten %.2e, all %.2e"
          % (rel_c[:10].max(), rel_c.max(), np.abs(vol - rv)[:10].max(), np.abs(vol - rv).max()), flush=True)
    for run in c["runs"]:                      # step 0: every FP64 run of the reference
        assert comp[0] == pytest.approx(run["compliance"][0], rel=1e-10)
    assert rel_c[:10].max() < 1e-5
    assert np.abs(vol - rv)[:10].max() < 1e-6
    assert np.allclose(ch[:10], stock["change"][:10], atol=1e-12)
    assert rel_c.max() < 3e-5 and np.abs(vol - rv).max() < 2e-5
    if steps_full:
        for run in c["runs"]:                  # the reference's equivalence criterion
            assert abs(comp[-1] / run["compliance"][99] - 1) < 1e-3


def _check_design(rho, golden, name):
    d = np.load(golden / f"session_h-{name}-fp64-stock.npz")["rho"].astype(np.float64)
    r = rho.detach().cpu().numpy()
    err = np.linalg.norm(r - d) / np.linalg.norm(d)
    print(f"final design rel L2 vs the reference's {err:.2e}", flush=True)
    assert err < 1e-4


def test_top88_loop_matches_the_oracle_loop():
    from voxelhip.simp import SimpProblem, simp_run
    shape = (16, 8, 8)
    ke = 2 * hx.quadrature([1, 1, 1], 0.3)
    _, _, mask, force = sr.cantilever(shape, seed=1)
    hist_o, x_o, rho_o = sr.top88_run(shape, ke, mask, force, steps=6, volfrac=0.3, rmin=1.5, move=0.2)
    for prec, tol, warm, tc, tr in (("fp64", 1e-12, False, 1e-9, 1e-8), ("mixed", 1e-8, True, 1e-6, 1e-5)):
        print("case", prec, warm, flush=True)
        pb = SimpProblem(shape, ke, mask, force, precision=prec, tol=tol, max_coarse_x=4, degree=3,
                         warm_start=warm, maxiter=300)
        hist, rho = simp_run(pb, steps=6, volfrac=0.3, rmin=1.5, move=0.2)
        for h, (c, change, vol) in zip(hist, hist_o):
            assert h["compliance"] == pytest.approx(c, rel=tc)
            assert h["volume"] == pytest.approx(vol, abs=tr)
            assert h["change"] == pytest.approx(change, abs=tr)
        assert np.abs(rho.cpu().numpy() - rho_o).max() < tr


@pytest.mark.parametrize("prec,tol,warm", [("fp64", 1e-10, False), ("mixed", 1e-8, True)])
def test_cantilever_512k_reproduces_the_reference_simp_records(records, golden, prec, tol, warm):
    """160x80x40 (1.2 M DOF), 100 steps; the GMG hierarchy halves to 20x10x5.  The mixed,
    warm-started solver is the benchmark's SIMP configuration."""
    from voxelhip.simp import SimpProblem, simp_run
    shape, ke, mask, force, c, stock = _case(records, golden, "cantilever-512k")
    pb = SimpProblem(shape, ke, mask, force, precision=prec, tol=tol, warm_start=warm, maxiter=400)
    hist, rho = simp_run(pb, steps=100, volfrac=records["volfrac"], rmin=1.5, move=records["simp"]["move"])
    assert not any(h["retried"] for h in hist)
    _check_trajectory(hist, c, stock, steps_full=True)
    _check_design(rho, golden, "cantilever-512k")


def test_cantilever_216k_first_ten_steps_reproduce_the_reference(records, golden):
    """120x60x30: z halves once only (30 -> 15), so no multigrid hierarchy reaches a small dense
    coarsest level; the states are solved by the device Jacobi-PCG in FP64 to 1e-10."""
    from voxelhip.simp import SimpProblem, simp_run
    shape, ke, mask, force, c, stock = _case(records, golden, "cantilever-216k")
    pb = SimpProblem(shape, ke, mask, force, precision="fp64", tol=1e-10, method="pcg", maxiter=100000)
    hist, _ = simp_run(pb, steps=10, volfrac=records["volfrac"], rmin=1.5, move=records["simp"]["move"])
    _check_trajectory(hist, c, stock, steps_full=False)


def test_continued_bounds_follow_a_changed_spectrum_and_the_retry_starts_cold():
    """Continued Chebyshev bounds (5 power steps from the previous rebuild's vectors) after a
    clearly changed design (a void block of E = 1e-9 appears) stay within 1 % of the 30-step
    bounds from the seeded vector; a retry from zero on that problem re-estimates the bounds cold."""
    from voxelhip.simp import SimpProblem
    shape = (32, 16, 16)
    ke = 2 * hx.quadrature([1, 1, 1], 0.3)
    rho, E1, mask, force = sr.cantilever(shape, seed=4, rmin=1.5)
    E2 = E1.reshape(shape).copy()
    E2[8:20, 4:12, :10] = 1e-9
    E2 = E2.ravel()
    kw = dict(precision="mixed", tol=1e-8, max_coarse_x=8, degree=3, maxiter=300)
    warm = SimpProblem(shape, ke, mask, force, warm_start=True, **kw)
    cold = SimpProblem(shape, ke, mask, force, warm_start=False, **kw)
    assert warm.solver.bound_warm_iterations == 5
    dev = lambda a: torch.from_numpy(a).cuda()  # noqa: E731
    warm.solve(dev(E1))
    u2 = warm.solve(dev(E2))
    cold.solve(dev(E2))
    lw, lc = warm.solver.M.lambdas, cold.solver.M.lambdas
    print("continued", lw, "cold", lc, flush=True)
    assert all(a >= 0.99 * b for a, b in zip(lw, lc))
    assert not warm.last_retried
    # force the retry: an unreachable iteration limit from a far-off start
    n0 = cold.last_iterations
    warm.maxiter = n0 + 1
    warm._u = torch.full_like(u2, 1e6)
    warm.solve(dev(E2))
    assert warm.last_retried and warm.last_rel <= 1e-8
    assert warm.solver.M.lambdas == pytest.approx(lc, rel=1e-12)
