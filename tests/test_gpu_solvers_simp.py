"""GPU parity of the device Jacobi-PCG / mixed IR and of the SIMP pieces (cone filter, OC,
compliance adjoint) against the oracle and the reference's known answers.

Known answers (results/session-a-20260923, smoke-q4, check_every 50): FP64 Jacobi-PCG 100
iterations; IR with FP32 inner PCG 3 outer / 150 inner iterations; compliance 22009.41647...
Tolerances: converged solutions within 1e-10 (FP64) / 5e-5 (the protocol's cross-state limit,
mixed); unconverged iterates within 1e-9 (CG amplifies rounding differences);
filter / OC / sensitivities at FP64 rounding (1e-13 relative), the adjoint against central
finite differences of the GPU solve to 1e-5 relative.
"""
import json

import numpy as np
import pytest
import torch

from oracle import hexref as hx
from oracle import simpref as sr
from oracle import solvers as osv

pytestmark = pytest.mark.gpu

KE = hx.quadrature([1, 1, 1], 0.3)


def rel(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else a
    return float(np.linalg.norm(np.asarray(a, dtype=np.float64) - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def smoke(golden):
    import voxelhip
    import voxel_mg as mg
    d = np.load(golden / "smoke-q4.npz")
    shape, E, mask, force = sr.fixture_case(d)
    ka = json.loads((golden / "known_answers.json").read_text())
    l64 = mg.Level(voxelhip.xp, shape, mg.HipProduct(shape, d["ke"]), E, mask, np.float64)
    l32 = mg.Level(voxelhip.xp, shape, mg.HipProduct(shape, d["ke"], np.float32), E, mask, np.float32)
    return dict(d=d, shape=shape, E=E, mask=mask, force=force, ka=ka, l64=l64, l32=l32)


def test_device_pcg_known_answers(smoke, golden):
    from voxelhip import solvers
    c = smoke
    f = torch.from_numpy(c["force"]).cuda()
    x, it, ok, r = solvers.pcg(c["l64"].apply, f, c["l64"].inv_diag, tol=1e-8, maxiter=20000,
                               mask=c["l64"].mask, check_every=50)
    ka = c["ka"]["session_a/pcg64-modal8"]
    assert ok and it == ka["iterations"] and r < 1e-12
    sol = hx.to_author(np.load(golden / "smoke-q4--pcg64-modal8--tol-00.npz")["solution"], c["shape"], True)
    assert rel(x, sol) < 1e-10
    assert float(f @ x) == pytest.approx(ka["force_work"], rel=1e-11)


def test_device_pcg_matches_the_oracle_iterates(smoke):
    from voxelhip import solvers
    c = smoke
    prod = hx.NumpyProduct(c["shape"], c["d"]["ke"])
    A = lambda u: prod.raw(u * c["mask"], c["E"]) * c["mask"] + u * (1 - c["mask"])  # noqa: E731
    inv = 1 / np.where(c["mask"] != 0, prod.diagonal(c["E"]), 1.0)
    for maxiter in (7, 50, 63):
        xo, ito, _, ro = osv.pcg(A, c["force"], inv, tol=0.0, maxiter=maxiter, mask=c["mask"], check_every=50)
        x, it, _, r = solvers.pcg(c["l64"].apply, torch.from_numpy(c["force"]).cuda(), c["l64"].inv_diag, tol=0.0,
                                  maxiter=maxiter, mask=c["l64"].mask, check_every=50)
        # unconverged CG iterates amplify rounding differences (E spans 1e-6..1): 1e-9; the true
        # residual near 1e-8 is itself at the rounding level of b - A x, compared to 1e-2
        assert it == ito and rel(x, xo) < 1e-9 and r == pytest.approx(ro, rel=1e-2)


def test_device_pcg_ir_known_answers(smoke, golden):
    from voxelhip import solvers
    c = smoke
    ka = c["ka"]["session_a/ir-modal8-modal8_32"]
    x, outer, inner, ok, r = solvers.pcg_ir(c["l64"].apply, torch.from_numpy(c["force"]).cuda(), c["l32"].apply,
                                            c["l32"].inv_diag, tol=1e-8, mask32=c["l32"].mask, inner_tol=0.1,
                                            max_outer=40, inner_maxiter=4000, check_every=50)
    assert ok and r <= 1e-8
    assert outer == ka["outer_iterations"] and inner == ka["inner_iterations"]
    sol = hx.to_author(np.load(golden / "smoke-q4--pcg64-modal8--tol-00.npz")["solution"], c["shape"], True)
    assert rel(x, sol) < 5e-5


def test_device_pcg_reports_nonconvergence(smoke):
    from voxelhip import solvers
    c = smoke
    _, it, ok, r = solvers.pcg(c["l64"].apply, torch.from_numpy(c["force"]).cuda(), c["l64"].inv_diag, tol=1e-8,
                               maxiter=1, mask=c["l64"].mask)
    assert it == 1 and not ok and r > 1e-8


@pytest.mark.parametrize("shape,rmin", [((9, 7, 11), 1.5), ((12, 10, 8), 2.4), ((5, 4, 3), 1.0)])
def test_cone_filter_matches_the_oracle(shape, rmin):
    from voxelhip.simp import ConeFilter
    x = np.random.default_rng(0).random(int(np.prod(shape)))
    fo = sr.ConeFilter(shape, rmin)
    f = ConeFilter(shape, rmin)
    xd = torch.from_numpy(x).cuda()
    assert rel(f.forward(xd), fo.forward(x)) < 1e-13
    assert rel(f.transpose(xd), fo.transpose(x)) < 1e-13
    assert f.check_adjoint() < 1e-12
    # the filtered-volume weights: mean(H x / s) == sum x w
    assert float(xd @ f.w_vol) == pytest.approx(float(fo.forward(x).mean()), rel=1e-12)


def test_oc_update_matches_the_oracle():
    from voxelhip.simp import ConeFilter, oc_top88
    shape = (10, 6, 6)
    rng = np.random.default_rng(4)
    n = int(np.prod(shape))
    x = rng.uniform(0.2, 0.5, n)
    dc = -rng.random(n) * 10
    fo = sr.ConeFilter(shape, 1.5)
    dv = fo.transpose(np.ones(n))
    xo = sr.oc_top88(x, fo.transpose(dc), dv, 0.3, 0.2, density=fo.forward)
    f = ConeFilter(shape, 1.5)
    xd = oc_top88(torch.from_numpy(x).cuda(), f.transpose(torch.from_numpy(dc).cuda()),
                  f.transpose(torch.ones(n, dtype=torch.float64, device="cuda")), 0.3, 0.2, filt=f)
    assert rel(xd, xo) < 1e-12


def _oc_top88_one_probe_at_a_time(x, dc, dv, volfrac, move, w, l2=1e9, tol=1e-3):
    """The bisection of experiments/session_h/run_w1b.py:82-94 with one vh_oc_probe per midpoint
    (the volume of a probe reduced as oc_top88 reduces it); returns (x_new, midpoints)."""
    from voxelhip import ops
    n = x.numel()
    nparts = min(1024, max(1, (n + 255) // 256))
    partial = torch.zeros(nparts, dtype=torch.float64, device=x.device)
    xnew = torch.empty_like(x)
    l1, l2, mids = 0.0, float(l2), []
    while (l2 - l1) / (l1 + l2) > tol:
        lmid = 0.5 * (l2 + l1)
        ops.oc_probe(x, dc, dv, w, lmid, move, xnew, partial, nparts)
        vol = partial.view(1, nparts).sum(dim=1).item()
        vol = vol if w is not None else vol / n
        mids.append(lmid)
        if vol > volfrac:
            l1 = lmid
        else:
            l2 = lmid
    return xnew, mids


@pytest.mark.parametrize("dtype,filtered,scale", [(torch.float64, True, 1e-9), (torch.float64, False, 10.0),
                                                  (torch.float32, True, 1e-3)])
def test_oc_bisection_evaluated_several_probes_at_a_time(dtype, filtered, scale):
    """oc_top88 evaluates the reference loop's midpoints up to 16 per pass (vh_oc_probe_multi):
    the multi-probe partials are bitwise the single probe's, and the design is bitwise the one of
    the loop that probes one midpoint at a time (sensitivities scaled so that the multiplier lies
    far below or near the initial bracket's top: up to 60 halvings before the first volume above
    the target)."""
    from voxelhip import ops
    from voxelhip.simp import ConeFilter, oc_top88
    shape = (40, 24, 20)
    n = int(np.prod(shape))
    g = torch.Generator(device="cuda").manual_seed(11)
    x = (0.2 + 0.3 * torch.rand(n, generator=g, device="cuda", dtype=torch.float64)).to(dtype)
    dc = (-scale * torch.rand(n, generator=g, device="cuda", dtype=torch.float64)).to(dtype)
    f = ConeFilter(shape, 1.5, dtype=dtype) if filtered else None
    dv = f.transpose(torch.ones(n, dtype=dtype, device="cuda")) if filtered else None
    w = f.w_vol if filtered else None
    nparts = min(1024, max(1, (n + 255) // 256))
    lams = [10.0 ** (-k) for k in range(16)]
    multi = torch.zeros(16 * nparts, dtype=torch.float64, device="cuda")
    ops.oc_probe_multi(x, dc, dv, w, lams, 0.2, multi, nparts)
    one = torch.zeros(nparts, dtype=torch.float64, device="cuda")
    xn = torch.empty_like(x)
    for k, lm in enumerate(lams):
        ops.oc_probe(x, dc, dv, w, lm, 0.2, xn, one, nparts)
        assert torch.equal(multi[k * nparts:(k + 1) * nparts], one), k
    ref, mids = _oc_top88_one_probe_at_a_time(x, dc, dv, 0.3, 0.2, w)
    print("case oc bisection", dtype, filtered, scale, len(mids), "probes", flush=True)
    assert len(mids) > 20
    got = oc_top88(x, dc, dv, 0.3, 0.2, filt=f)
    assert torch.equal(got, ref)


def _problem(shape, ke, mask, force, precision="fp64", tol=1e-12):
    """The SIMP state problem on the GMG engine, solved far below the FD step's truncation."""
    from voxelhip.simp import SimpProblem
    return SimpProblem(shape, ke, mask, force, precision=precision, tol=tol, max_coarse_x=4, degree=3)


def test_compliance_adjoint_matches_finite_differences():
    from voxelhip.simp import Compliance
    shape = (8, 4, 4)
    ke = 2 * hx.quadrature([1, 1, 1], 0.3)
    rho, _, mask, force = sr.cantilever(shape, seed=2, rho_lo=0.3)
    pb = _problem(shape, ke, mask, force)
    r = torch.from_numpy(rho).cuda().requires_grad_(True)
    c = Compliance.apply(r, pb)
    c.backward()
    g = r.grad.cpu().numpy()
    # oracle sensitivity from the oracle's own dense solve
    K = hx.assemble_dense(shape, ke, sr.modulus(rho), mask)
    u = np.linalg.solve(K, force)
    assert float(c.detach()) == pytest.approx(float(force @ u), rel=1e-10)
    assert rel(g, sr.sensitivity(rho, hx.element_energy(shape, ke, u))) < 1e-9
    h = 1e-5
    for e in (0, 37, 101):
        rp, rm = rho.copy(), rho.copy()
        rp[e] += h
        rm[e] -= h
        cp = float(Compliance.apply(torch.from_numpy(rp).cuda(), pb))
        cm = float(Compliance.apply(torch.from_numpy(rm).cuda(), pb))
        assert (cp - cm) / (2 * h) == pytest.approx(g[e], rel=1e-5)


def test_simp_loop_runs_and_keeps_the_volume():
    from voxelhip.simp import simp_run
    shape = (16, 8, 8)
    ke = 2 * hx.quadrature([1, 1, 1], 0.3)
    _, _, mask, force = sr.cantilever(shape, seed=1)
    pb = _problem(shape, ke, mask, force)
    hist, rho = simp_run(pb, steps=5, volfrac=0.3, rmin=1.5, move=0.2)
    assert len(hist) == 5
    assert all(abs(h["volume"] - 0.3) < 2e-3 for h in hist)
    assert hist[-1]["compliance"] < hist[0]["compliance"]
    assert float(rho.min()) >= 0.0 and float(rho.max()) <= 1.0


_SIMP_RUN = r"""
import json, sys
sys.path[:0] = [sys.argv[1], sys.argv[1] + "/matrix-free-voxel-elasticity_amd"]
import numpy as np, torch
from oracle import hexref as hx, simpref as sr
from voxelhip.simp import SimpProblem, simp_run
shape = (96, 48, 48)
ke = 2 * hx.quadrature([1, 1, 1], 0.3)
_, _, mask, force = sr.cantilever(shape, seed=2)
pb = SimpProblem(shape, ke, mask, force, precision="mixed", tol=1e-8, max_coarse_x=12, degree=3, warm_start=True)
hist, rho = simp_run(pb, steps=8, volfrac=0.3, rmin=1.5, move=0.2)
print(json.dumps({"its": [h["iterations"] for h in hist], "c": [float(h["compliance"]).hex() for h in hist],
                  "rho": float(rho.double().square().sum()).hex()}))
"""


def test_simp_run_is_bitwise_reproducible_across_processes():
    """The warm-started mixed-precision SIMP loop (state solves, rebuilt hierarchies with continued
    spectrum estimates, dense coarsest factorisation, OC bisection) gives bit-identical iteration
    counts, compliances and designs in two separate processes: no step depends on timing, stream
    interleaving or memory left over from an earlier allocation (96 x 48 x 48, coarsest 12 x 6 x 6: the
    dense factorisation of the 384 x 192 x 192 SIMP run)."""
    import subprocess
    import sys
    from pathlib import Path
    root = str(Path(__file__).resolve().parents[1])
    outs = []
    for run in range(2):
        print("case process", run, flush=True)
        r = subprocess.run([sys.executable, "-c", _SIMP_RUN, root], capture_output=True, text=True, timeout=240)
        assert r.returncode == 0, r.stderr[-2000:]
        outs.append(json.loads(r.stdout.strip().splitlines()[-1]))
    assert outs[0] == outs[1]


def test_warm_started_simp_loop_follows_the_cold_one():
    """State solves started from the previous state (tolerance relative to |f| either way): the
    compliance and volume histories of the cold run within 1e-7, fewer iterations after the first
    step; a warm start that cannot converge falls back to the solve from zero."""
    from voxelhip.simp import SimpProblem, simp_run
    shape = (16, 8, 8)
    ke = 2 * hx.quadrature([1, 1, 1], 0.3)
    _, _, mask, force = sr.cantilever(shape, seed=1)
    runs = []
    # cold; warm start alone (fixed spectral bounds procedure); warm start with the bounds'
    # power iteration continued from the previous state (the SimpProblem default when warm)
    for warm, bw in ((False, None), (True, None), (True, "default")):
        print("case warm", warm, bw, flush=True)
        kw = {} if bw == "default" else dict(bound_warm_iterations=bw)
        pb = SimpProblem(shape, ke, mask, force, precision="mixed", tol=1e-8, max_coarse_x=4, degree=3,
                         warm_start=warm, **kw)
        runs.append((pb, simp_run(pb, steps=6, volfrac=0.3, rmin=1.5, move=0.2)[0]))
    (_, cold), (pb, warm), (pbw, warm_bw) = runs
    assert pbw.solver.bound_warm_iterations == 5
    for run in (warm, warm_bw):
        for a, b in zip(cold, run):
            assert b["compliance"] == pytest.approx(a["compliance"], rel=1e-7)
            assert b["volume"] == pytest.approx(a["volume"], abs=1e-7)
        assert run[0]["iterations"] == cold[0]["iterations"]
    assert sum(h["iterations"] for h in warm[1:]) < sum(h["iterations"] for h in cold[1:])
    # fail-closed: a warm start that misses the iteration limit is retried from zero (on the
    # problem whose rebuilds repeat the same bounds, so the retry is the cold solve itself)
    E = torch.full((int(np.prod(shape)),), 0.5, dtype=torch.float64, device="cuda")
    pb.warm_start = False
    u0 = pb.solve(E)
    n0 = pb.last_iterations
    assert not pb.last_retried
    pb.warm_start, pb.maxiter = True, n0 + 1
    pb._u = torch.full_like(u0, 1e6)          # far from the solution: needs many more than n0 + 1
    u = pb.solve(E)
    assert pb.last_retried and pb.last_iterations == n0 + 1 + n0 and pb.last_rel <= 1e-8
    assert float((u - u0).norm() / u0.norm()) < 1e-12    # the retry is the cold solve


def test_mixed_precision_state_solve_gives_the_fp64_sensitivities():
    """The SIMP step on the mixed solver (FP32 V-cycle, FP64 outer CG to 1e-8): compliance and
    dc/drho within 1e-6 of the FP64 solve at 1e-12."""
    from voxelhip.simp import Compliance
    shape = (16, 8, 8)
    ke = 2 * hx.quadrature([1, 1, 1], 0.3)
    rho, _, mask, force = sr.cantilever(shape, seed=3, rho_lo=0.3)
    out = []
    for prec, tol in (("fp64", 1e-12), ("mixed", 1e-8)):
        pb = _problem(shape, ke, mask, force, precision=prec, tol=tol)
        r = torch.from_numpy(rho).cuda().requires_grad_(True)
        c = Compliance.apply(r, pb)
        c.backward()
        out.append((float(c.detach()), r.grad.clone()))
    assert out[1][0] == pytest.approx(out[0][0], rel=1e-7)
    assert float((out[1][1] - out[0][1]).norm() / out[0][1].norm()) < 1e-6


# ------------------------------------------------------------------ projected SIMP driver
@pytest.mark.parametrize("beta", [0.0, 1.0, 16.0])
def test_heaviside_kernel_matches_the_oracle(beta):
    from voxelhip import ops
    rng = np.random.default_rng(41)
    x = rng.uniform(0.0, 1.0, 5000)
    xd = torch.from_numpy(x).cuda()
    r, d, E = (torch.empty_like(xd) for _ in range(3))
    partial = torch.zeros(20, dtype=torch.float64, device="cuda")
    ops.heaviside(xd, beta, 0.5, 3.5, 1.0, 1e-9, rho=r, drho=d, E=E, partial=partial)
    rp = sr.project(x, beta)
    assert np.abs(r.cpu().numpy() - rp).max() < 1e-14
    assert rel(d, sr.dproject(x, beta)) < 1e-13
    assert rel(E, sr.modulus(rp, 3.5)) < 1e-13
    assert float(partial.sum()) == pytest.approx(rp.sum(), rel=1e-13)


def test_projected_oc_update_matches_the_oracle():
    from voxelhip import simp
    shape = (10, 6, 6)
    rng = np.random.default_rng(7)
    n = int(np.prod(shape))
    rho = rng.uniform(0.05, 0.9, n)
    dc, dv = -rng.uniform(0.1, 2.0, n), rng.uniform(0.5, 1.5, n) / n
    filt, ofilt = simp.ConeFilter(shape, 1.35), sr.ConeFilter(shape, 1.35)
    for beta, hint in ((1.0, None), (8.0, 3.0)):
        print("case beta", beta, flush=True)
        want, vol_o, lam_o, _ = sr.oc_update(rho, dc, dv, 0.4, 0.1, ofilt, beta, 0.5, 1e-3, lam_hint=hint)
        got, vol, lam, _ = simp.oc_update(*(torch.from_numpy(a).cuda() for a in (rho, dc, dv)), 0.4, 0.1, filt, beta,
                                          0.5, 1e-3, lam_hint=hint)
        assert lam == pytest.approx(lam_o, rel=1e-7) and vol == pytest.approx(vol_o, abs=1e-9)
        assert np.abs(got.cpu().numpy() - want).max() < 1e-7


def test_warm_started_engine_solve():
    from voxelhip.engine import GMGSolver
    shape = (32, 16, 16)
    _, E, mask, force = sr.cantilever(shape, seed=3, rmin=1.5)
    f = torch.from_numpy(force).cuda()
    s = GMGSolver(shape, KE, mask, torch.from_numpy(E).cuda(), max_coarse_x=8, smoother="chebyshev", degree=3,
                  precision="mixed", graph=True)
    x, it, ok, r, _ = s.solve(f, tol=1e-10)
    cold = x.clone()
    assert ok
    x, it0, ok, r0, mv = s.solve(f, tol=1e-8, x0=cold)           # already converged: no iteration
    assert ok and it0 == 0 and mv == 1 and r0 <= 1e-8 and torch.equal(x, cold)
    x, itw, ok, rw, _ = s.solve(f, tol=1e-10, x0=0.9 * cold)     # a close start needs fewer iterations
    assert ok and rw <= 1e-10 and itw < it
    assert float((x - cold).norm() / cold.norm()) < 1e-8


def test_projected_simp_driver_matches_the_oracle():
    from voxelhip import simp
    shape = (8, 4, 4)
    _, _, mask, force = sr.cantilever(shape, seed=1)
    pb = simp.SimpProblem(shape, KE, mask, force, precision="fp64", max_coarse_x=2, maxiter=300)
    res = simp.run_simp(pb, fixed_budget=8, cg_tol=1e-11)
    hist_o, rho_o = sr.run_simp(shape, KE, mask, force, steps=8)
    c = [h["compliance"] for h in res["history"]]
    assert np.allclose(c, [h[0] for h in hist_o], rtol=1e-7)
    assert all(abs(h["vol_phys"] - 0.3) < 1e-6 for h in res["history"])
    assert np.abs(res["rho"].cpu().numpy() - rho_o).max() < 1e-6
    assert res["stop_reason"] == "fixed_budget" and res["warm_start_retries"] == 0
    # warm starts: later solves take fewer iterations than the first
    its = [s_["cg_iters"] for s_ in res["solve_log"][:-1]]
    assert min(its[3:]) < its[0]


def test_design_gradient_through_filter_and_projection():
    from voxelhip import simp
    shape = (10, 5, 5)
    _, _, mask, force = sr.cantilever(shape, seed=2)
    pb = simp.SimpProblem(shape, KE, mask, force, precision="fp64", tol=1e-12, max_coarse_x=4, maxiter=400,
                          penal=3.0)
    filt = simp.ConeFilter(shape, 1.5)
    rng = np.random.default_rng(5)
    x = torch.from_numpy(rng.uniform(0.2, 0.8, int(np.prod(shape)))).cuda().requires_grad_(True)
    c = simp.design_compliance(x, pb, filt, beta=4.0)
    c.backward()
    g = x.grad.clone()
    for k in range(3):
        d = torch.from_numpy(rng.standard_normal(x.numel())).cuda()
        h = 1e-5
        cp = float(simp.design_compliance((x + h * d).detach(), pb, filt, beta=4.0))
        cm = float(simp.design_compliance((x - h * d).detach(), pb, filt, beta=4.0))
        fd = (cp - cm) / (2 * h)
        assert float(g @ d) == pytest.approx(fd, rel=1e-5)
