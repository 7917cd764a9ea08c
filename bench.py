"""Benchmark of the voxel elasticity hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--cpu-baseline sample|full|none]
                    [--no-batch] [--no-reference] [--simp-steps N]

Workload (one "step" = one complete solve from x = 0 to the FP64 true relative residual
||b - A x|| / ||b|| <= 1e-8, setup excluded as in the reference's timing boundary):
  Hex8 voxel cantilever 256 x 128 x 128 elements (12,830,211 DOFs), E0 = 1, Emin = 1e-9,
  SIMP p = 3, densities iid U[0.01, 1] (seed 3) smoothed by the radius-1.5 cone filter, x = 0
  face clamped, unit downward line load on the x = nx, z = nz/2 edge; GMG-preconditioned
  flexible CG (5 levels, 2x coarsening, Chebyshev-Jacobi degree 3 on [lambda_max / 15, lambda_max],
  degree 2 on the levels from the third down as the donor's coarse_smooth_iters, dense coarsest), mixed
  precision: FP32 V-cycle inside the FP64 outer CG with FP64 true residual and FP64 dot
  accumulation.  The plain FP64 solve and the FP64 iterative refinement around the FP32 solve
  of the same system are timed beside it; each solution's error is stated against the FP64
  NumPy oracle (its constrained residual computed on the host).

Printed JSON line: value = mixed-precision solves/s; extra fields carry the FP64 and refinement
solves, the 128 x 64 x 64 FP64 GMG solve, a high-contrast solve with the full Galerkin hierarchy,
the 48-problem batch, the 384 x 192 x 192 SIMP run, the reference's own study solves (q64x1, q64x2,
q96x2) beside its published times, the product's GDOF/s under the automatic and the tuned launch
plan at three sizes, the roofline of the dominant kernel (the FP64 parity product at 256 x 128 x 128,
tuned plan, HIP events on its launch stream), and the CPU baseline (the NumPy oracle on the host
cores: a bounded sample of the same solve, or with --cpu-baseline full the products, the 32x16x16
Jacobi-PCG and the 128x64x64 GMG-FCG solves in two thread settings).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "matrix-free-voxel-elasticity_amd"))

import numpy as np  # noqa: E402

HBM_PEAK = 8.0e12           # MI355X HBM3E bytes/s
# HBM bytes per launch of the FP64 product at 256 x 128 x 128, from FETCH_SIZE / WRITE_SIZE
# counter runs of `bench.py --only matvec`, each scaled by the same counter read on a plain copy
TRAFFIC_FILE = ROOT / "profiles" / "parity_xmarch_fp64_256x128x128_traffic.json"


def measured_traffic():
    try:
        return int(json.loads(TRAFFIC_FILE.read_text())["traffic_bytes_per_launch"])
    except (OSError, ValueError, KeyError):
        return None


def problem(shape, seed, rmin=1.5):
    """Synthetic cantilever in the author layout (device tensors): rho, E, force, mask values."""
    import torch
    from voxelhip import ops
    from voxelhip.simp import ConeFilter, modulus
    nx, ny, nz = shape
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(seed)
    rho = torch.from_numpy(rng.uniform(0.01, 1.0, nx * ny * nz)).to(dev)
    if rmin:
        rho = ConeFilter(shape, rmin, device=dev).forward(rho)
    E = modulus(rho, 3.0, 1.0, 1e-9)
    f = torch.zeros((nx + 1, ny + 1, nz + 1, 3), dtype=torch.float64, device=dev)
    line = torch.full((ny + 1,), -1.0 / ny, dtype=torch.float64, device=dev)
    line[0] = line[-1] = -0.5 / ny
    f[nx, :, nz // 2, 2] = line
    m = torch.ones_like(f)
    m[0] = 0
    return rho, E, f.reshape(-1).contiguous(), m.reshape(-1), ops


def c4_problems(seed=4, n=48):
    from voxelhip.batch import random_cantilevers
    return random_cantilevers(seed=seed, n=n)


def timed_batch(problems, ke, steps, warmup, graph=True, sequential=True):
    """The batched solver on all problems (one launch per pass) and, for comparison, the same
    instances solved one after another with the single-problem engine."""
    import torch
    from voxelhip.batch import BatchedGMGSolver, coarse_dof_shapes
    from voxelhip.engine import GMGSolver
    t0 = time.perf_counter()
    bs = BatchedGMGSolver(problems, ke, precision="mixed", graph=graph, lower_frac=1.0 / 15.0)
    setup = time.perf_counter() - t0
    for _ in range(warmup):
        bs.solve(tol=1e-8, maxiter=200)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        _, its, ok, rel = bs.solve(tol=1e-8, maxiter=200)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    n_inst = bs.n
    out = {"problems": len(problems), "instances": n_inst,
           "dofs_total": int(sum(3 * (p.shape[0] + 1) * (p.shape[1] + 1) * (p.shape[2] + 1) * len(p.loads)
                                 for p in problems)),
           "grids": sorted({"x".join(map(str, p.shape)) for p in problems}),
           "setup_s": setup, "seconds_per_batch": dt, "instance_solves_per_s": n_inst / dt,
           "iterations_min_max": [int(its.min()), int(its.max())], "converged": bool(ok.all()),
           "max_rel_residual": float(rel.max())}
    del bs
    if sequential:
        solvers = []
        for p in problems:
            shapes = coarse_dof_shapes(p.shape)
            solvers.append(GMGSolver(p.shape, ke, p.mask, p.E, smoother="chebyshev", degree=3, precision="mixed",
                                     graph=graph, max_coarse_x=0, n_levels=len(shapes), lower_frac=1.0 / 15.0))
        loads = [[torch.from_numpy(f).cuda() for f in p.loads] for p in problems]
        for s, fs in zip(solvers, loads):
            for f in fs:
                s.solve(f)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            for s, fs in zip(solvers, loads):
                for f in fs:
                    s.solve(f)
        torch.cuda.synchronize()
        seq = (time.perf_counter() - t0) / steps
        out["sequential_seconds"] = seq
        out["batched_speedup_vs_sequential"] = seq / dt
    return out


def simp_leg(ke, shape=(384, 192, 192), steps=100, seed=5):
    """SIMP topology optimisation on the cantilever: volume fraction 0.3, cone filter r = 1.5, OC
    move 0.2, the state solved by the mixed-precision GMG-FCG to 1e-8 (relative to |f|) each step,
    started from the previous step's state (retried from zero if that fails), dc/drho from the
    adjoint element-energy kernel through torch.autograd."""
    from voxelhip.simp import SimpProblem, simp_run
    _, _, force, mask, _ = problem(shape, seed=seed, rmin=0)
    t0 = time.perf_counter()
    pb = SimpProblem(shape, ke, mask, force, precision="mixed", tol=1e-8, max_coarse_x=16, degree=3, warm_start=True)
    hist, rho = simp_run(pb, steps=steps, volfrac=0.3, rmin=1.5, move=0.2)
    total = time.perf_counter() - t0
    its = [h["iterations"] for h in hist]
    retried = [h["step"] for h in hist if h.get("retried")]
    return {"shape": list(shape), "ndof": int(force.numel()), "steps": len(hist), "warm_start": True,
            "continued_bound_steps": pb.solver.bound_warm_iterations,
            "warm_start_retries": len(retried), "retried_steps": retried,
            "seconds_total": total,
            "seconds_per_step": total / len(hist), "compliance_first": hist[0]["compliance"],
            "compliance_last": hist[-1]["compliance"], "volume_last": hist[-1]["volume"],
            "change_last": hist[-1]["change"], "solve_iterations_min_max_mean": [min(its), max(its),
                                                                                   float(np.mean(its))]}


# the reference's own study solves (inputs regenerated from its base designs in tests/golden),
# with its published times on other GPUs (the paper tables cited per case)
REFERENCE_WORKLOADS = {
    "q64x1": {"design": "optimized-q64 (session D)", "omega_seed": 2026092320, "iterations": 13,
              "published": {"fp64_s": 0.1037, "mixed_s": 0.0626, "gpu": "RTX 4090",
                            "source": "paper/session_d_tables.md:26,29"}},
    "q64x2": {"design": "optimized-q64 upsampled x2 (session I)", "omega_seed": 2026092621, "iterations": 14,
              "published": {"fp64_s": 0.340, "mixed_s": 0.236, "gpu": "H100 80GB HBM3",
                            "source": "paper/supplement/outcomes-scale-tables.md:217,219"}},
    "q96x2": {"design": "optimized-q96 upsampled x2 (session I)", "omega_seed": 2026092621, "iterations": 24,
              "published": {"fp64_s": 2.200, "mixed_s": 1.392, "gpu": "H100 80GB HBM3",
                            "source": "paper/supplement/outcomes-scale-tables.md:221,226"}},
}


def reference_workloads(steps, warmup, graph, tuner):
    """The reference's GMG-FCG study solves (voxel hierarchy halved to x <= 16, damped Jacobi
    V(4,4) with omega = 4/(3 lambda) from 30 seeded power steps, dense coarsest, FP64 flexible CG
    to a true relative residual of 1e-8 tested every iteration; Emin = 1e-6) on the engine, FP64
    and with the FP32 V-cycle, timed like the reference (setup excluded)."""
    import torch
    from voxelhip import harness
    from voxelhip.engine import GMGSolver
    out = {}
    for case, ref in REFERENCE_WORKLOADS.items():
        data = harness.load_case(ROOT / "tests" / "golden", case)
        shape = tuple(int(v) for v in data["shape"])
        emin = harness.PROTOCOL["material"]["Emin"]
        E = torch.from_numpy(emin + (1 - emin) * np.asarray(data["rho"]) ** 3).cuda()
        mask = torch.from_numpy(np.asarray(data["mask"], dtype=np.float64)).cuda()
        force = torch.from_numpy(np.asarray(data["force"])).cuda()
        row = {"design": ref["design"], "shape": list(shape), "ndof": int(force.numel()),
               "reference_iterations": ref["iterations"], "published": ref["published"]}
        for prec in ("fp64", "mixed"):
            s = GMGSolver(shape, data["ke"], mask, E, smoother="jacobi", pre=4, post=4, max_coarse_x=16,
                          precision=prec, graph=graph, tune=tuner, omega_seed=ref["omega_seed"])
            dt, st = timed_solves(s, force, steps, warmup)
            pub = ref["published"][f"{prec}_s"]
            row[prec] = {"seconds_per_solve": dt, "iterations": sorted({x[0] for x in st}),
                         "converged": all(x[1] for x in st), "max_rel_residual": max(x[2] for x in st),
                         "speedup_vs_published": pub / dt}
            del s
            torch.cuda.empty_cache()
        row["mixed_speedup_vs_fp64"] = row["fp64"]["seconds_per_solve"] / row["mixed"]["seconds_per_solve"]
        out[case] = row
    return out


def timed_solves(solver, b, steps, warmup):
    import torch
    for _ in range(warmup):
        solver.solve(b)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    for _ in range(steps):
        _, it, ok, rel, mv = solver.solve(b)
        stats.append((it, ok, rel, mv))
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    return dt, stats


def matvec_timing(shape, dtype, reps=20, plan=None):
    """Average duration of the parity product (raw K u), HIP events on the launch stream; plan =
    (tile, lx) of the launch, None for the library's automatic plan."""
    import torch
    import voxel_kernels as vk
    from voxelhip import ops
    ke = 2 * vk.quadrature([1, 1, 1], 0.3)
    blocks = vk.Modal(ke).blocks
    ne, nd = int(np.prod(shape)), 3 * ops.n_nodes(shape)
    g = torch.Generator(device="cuda").manual_seed(0)
    E = torch.rand(ne, generator=g, device="cuda", dtype=dtype)
    us = [torch.randn(nd, generator=g, device="cuda", dtype=dtype) for _ in range(2)]
    ys = [torch.empty(nd, device="cuda", dtype=dtype) for _ in range(2)]
    tile, lx = plan if plan is not None else (0, 0)
    for i in range(4):
        ops.matvec(blocks, E, us[i % 2], ys[i % 2], shape, tile=tile, lx=lx)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    stream.synchronize()
    e0.record(stream)
    for i in range(reps):
        ops.matvec(blocks, E, us[i % 2], ys[i % 2], shape, tile=tile, lx=lx)
    e1.record(stream)
    e1.synchronize()
    sec = e0.elapsed_time(e1) / 1e3 / reps
    word = 8 if dtype == torch.float64 else 4
    algo = word * (2 * nd + ne)
    nt, nw = ops.matvec_tiles(shape, lx=lx, tile=tile, dtype=dtype, blocks=blocks)
    return {"shape": list(shape), "dtype": str(dtype).replace("torch.", ""), "us": sec * 1e6,
            "gdof_s": nd / sec / 1e9, "algorithmic_bytes": algo, "achieved_GBs": algo / sec / 1e9,
            "frac_hbm": algo / sec / HBM_PEAK, "plan": {"tile": tile, "lx": lx, "workgroups": nw}}


TILE_NAMES = {0: "auto", 1: "16x32", 2: "8x32", 3: "4x8", 4: "wide"}


def slab_share(dtype, shape=(48, 192, 192), reps=20):
    """One GPU's share of the 384x192x192 grid split into x-slabs (48x192x192: eight GPUs): the
    product on the slab (tuned plan) and the two steps of the shared-plane exchange one GPU can
    show, timed with HIP events: staging the boundary node plane (contiguous in the author layout,
    3 (ny+1)(nz+1) values) into a send buffer, and adding a received plane into y."""
    import torch
    r = tuned_matvec(shape, dtype)
    nplane = 3 * (shape[1] + 1) * (shape[2] + 1)
    y = torch.randn((shape[0] + 1) * nplane, device="cuda", dtype=dtype)
    send, recv = torch.empty(nplane, device="cuda", dtype=dtype), torch.randn(nplane, device="cuda", dtype=dtype)
    out = {"product": r, "plane_values": nplane, "plane_bytes": nplane * y.element_size()}
    for name, fn in (("stage_plane_us", lambda: send.copy_(y[-nplane:])),
                     ("add_received_plane_us", lambda: y[:nplane].add_(recv))):
        for _ in range(3):
            fn()
        s = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.synchronize()
        e0.record(s)
        for _ in range(reps):
            fn()
        e1.record(s)
        e1.synchronize()
        out[name] = e0.elapsed_time(e1) / reps * 1e3
    # the transfer itself is not measurable on one GPU: at ~153 GB/s per xGMI link
    out["xgmi_transfer_us_estimate"] = out["plane_bytes"] / 153e9 * 1e6
    return out


def tuned_matvec(shape, dtype, reps=20):
    """The product under the launch plan the library's tuner would pick for it: every (tile, lx)
    candidate of voxelhip.tuner and the automatic plan timed briefly, the four fastest timed
    again with `reps` launches each, and the fastest of those kept (the short first round alone
    is at the mercy of run-to-run noise of a few per cent)."""
    from voxelhip import tuner
    first = []
    for plan in [(0, 0)] + list(tuner._candidates(shape[0])):
        first.append((matvec_timing(shape, dtype, reps=5, plan=plan)["us"], plan))
    first.sort()
    best = None
    for _, plan in first[:4]:
        r = matvec_timing(shape, dtype, reps=reps, plan=plan)
        if best is None or r["us"] < best["us"]:
            best = r
    best["plan"]["tile_name"] = TILE_NAMES.get(best["plan"]["tile"], str(best["plan"]["tile"]))
    return best


def oracle_residual(shape, ke, E, mask, force, x):
    """||f - A x|| / ||f|| of a device solution, A the constrained operator of the NumPy oracle
    (oracle/hexref.py NumpyProduct, FP64, host): the error statement of a GPU solution against
    the FP64 oracle."""
    sys.path.insert(0, str(ROOT))
    from oracle import hexref as hx
    Eh = E.detach().cpu().numpy().astype(np.float64)
    m = mask.detach().cpu().numpy().astype(np.float64)
    f = force.detach().cpu().numpy().astype(np.float64)
    u = x.detach().cpu().numpy().astype(np.float64)
    Au = hx.NumpyProduct(shape, ke).raw(u * m, Eh) * m + u * (1 - m)
    return float(np.linalg.norm(f - Au) / np.linalg.norm(f))


def cpu_baseline(shape, E, mask, force, ke, lambdas, iterations):
    """Oracle (NumPy, host cores) cost of the same FP64 GMG-FCG solve, from a bounded sample.

    The GPU solve runs one V-cycle before the loop and, per flexible-CG iteration, one V-cycle
    and two fine constrained products (A p and the true residual).  The sample times one oracle
    V-cycle and one oracle constrained product at the full size; the solve time is
    (iterations + 1) * V-cycle + (2 * iterations) * product."""
    sys.path.insert(0, str(ROOT))
    from oracle import hexref as hx
    from oracle import mgref as mgr
    E = E.cpu().numpy()
    mask = mask.cpu().numpy()
    force = force.cpu().numpy()
    shapes = mgr.hierarchy_shapes(shape, 16)
    prods = [hx.NumpyProduct(s, ke * 2 ** i) for i, s in enumerate(shapes)]
    Ec, mc = E, mask
    for i in range(1, len(shapes)):
        Ec, mc = mgr.coarsen_modulus(Ec, shapes[i - 1]), mgr.coarsen_mask(mc, shapes[i - 1])
    A = hx.assemble_dense(shapes[-1], ke * 2 ** (len(shapes) - 1), Ec, mc)
    M = mgr.Multigrid(shapes, prods, E, mask, np.linalg.cholesky(A), np.float64, pre=3, post=3,
                      smoother="chebyshev", lambdas=lambdas)
    fine = hx.NumpyProduct(shape, ke)
    t0 = time.perf_counter()
    z = M(force)
    t_v = time.perf_counter() - t0
    t0 = time.perf_counter()
    fine.raw(z * mask, E) * mask + z * (1 - mask)
    t_a = time.perf_counter() - t0
    solve_sec = (iterations + 1) * t_v + 2 * iterations * t_a
    cores = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count())
    return {"value": 1.0 / solve_sec, "unit": "solves/s", "cores": cores, "kind": "port",
            "sample": f"NumPy oracle, FP64 GMG-FCG on {shape[0]}x{shape[1]}x{shape[2]}: one V-cycle "
                      f"({t_v:.2f} s) and one constrained fine product ({t_a:.2f} s) timed, solve = "
                      f"(it+1) V-cycles + 2 it products at the GPU iteration count it = {iterations}; "
                      f"BLAS threads = {cores}, np.add.at single-threaded"}


def cpu_baseline_full():
    """The CPU-baseline protocol (oracle/cpu_baseline.py: products at 128x64x64 and 256x128x128,
    the 32x16x16 Jacobi-PCG solve, the 128x64x64 GMG-FCG solve; medians of 5 after a warm-up) in
    two thread settings: the process's own (OMP_NUM_THREADS as set, else every host core) and 1.
    Children run one after another, so each has the host to itself."""
    import subprocess
    runs = {}
    own = os.environ.get("OMP_NUM_THREADS") or str(os.cpu_count())
    for label, n in (("all", own), ("one", "1")):
        env = dict(os.environ, OMP_NUM_THREADS=n, OPENBLAS_NUM_THREADS=n, MKL_NUM_THREADS=n)
        r = subprocess.run([sys.executable, str(ROOT / "oracle" / "cpu_baseline.py")], env=env,
                           capture_output=True, text=True, timeout=1800)
        if r.returncode != 0:
            raise RuntimeError(f"oracle/cpu_baseline.py failed: {r.stderr[-2000:]}")
        runs[label] = json.loads(r.stdout.strip().splitlines()[-1])
        print(f"bench.py: CPU baseline protocol, {n} thread(s), done", file=sys.stderr, flush=True)
    return runs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--graph", type=int, default=1)
    ap.add_argument("--tune", type=int, default=1, help="empirical launch tuning of the product passes")
    ap.add_argument("--no-batch", action="store_true", help="skip the 48-problem batched leg")
    ap.add_argument("--simp-steps", type=int, default=100,
                    help="SIMP steps of the 384x192x192 optimisation leg (0 skips it)")
    ap.add_argument("--only", choices=["matvec", "solve", "batch"], default=None,
                    help="matvec: only the FP64 and FP32 products at 256x128x128 (plans from --mv-plan) "
                         "and a plain copy of known size, for counter collection; solve: only the "
                         "256x128x128 solves of --precision (for a per-kernel trace of one solve); batch: "
                         "only the 48-problem batched solve (for a trace of its launches)")
    ap.add_argument("--mv-plan", default="tuned",
                    help="launch plan of the --only matvec products: 'tuned', or 'tile,lx' for both "
                         "precisions, or 'tile,lx;tile,lx' for FP64;FP32 (tile 1 = 16x32, 2 = 8x32, 4 = wide)")
    ap.add_argument("--precision", choices=["mixed", "fp64", "both"], default="both",
                    help="precisions of the 256x128x128 solve legs")
    ap.add_argument("--no-reference", action="store_true",
                    help="skip the reference's own study solves (q64x1, q64x2, q96x2)")
    ap.add_argument("--cpu-baseline", choices=["sample", "full", "none"], default="sample",
                    help="sample: the bounded C3 sample (about 10 s); full: the CPU-baseline protocol "
                         "(oracle/cpu_baseline.py in two thread settings, a few minutes); none: skip")
    args = ap.parse_args()
    if args.no_cpu_baseline:
        args.cpu_baseline = "none"
    if args.gpus != 1:
        print(f"bench.py: the solver runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the multi-GPU slab decomposition is a design note only)", file=sys.stderr)
        sys.exit(2)
    import torch
    import voxel_kernels as vk
    from voxelhip.engine import GMGSolver
    from voxelhip.tuner import LaunchTuner
    torch.cuda.set_device(0)
    if args.only == "matvec":
        res = {}
        plans = [None, None]
        if args.mv_plan != "tuned":
            parts = [tuple(int(v) for v in q.split(",")) for q in args.mv_plan.split(";")]
            plans = [parts[0], parts[-1]]
        for dt_, plan in zip((torch.float64, torch.float32), plans):
            key = str(dt_).replace("torch.", "")
            res[key] = (tuned_matvec((256, 128, 128), dt_) if plan is None
                        else matvec_timing((256, 128, 128), dt_, plan=plan))
        nd = 3 * 257 * 129 * 129
        src = torch.ones(nd, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):          # plain copy of 2 x 102.6 MB: the counter calibration reference
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"only": "matvec", "matvec": res, "copy_bytes_each_way": nd * 8}), flush=True)
        return
    if args.only == "batch":
        ke = 2 * vk.quadrature([1, 1, 1], 0.3)
        r = timed_batch(c4_problems(), ke, args.steps, args.warmup, graph=bool(args.graph), sequential=False)
        print(json.dumps({"only": "batch", "c4_batched": r}), flush=True)
        return
    ke = 2 * vk.quadrature([1, 1, 1], 0.3)          # the donor's KE_UNIT_3D (ke_unit_3d.npy)
    t_start = time.perf_counter()

    def progress(leg):
        # one line per finished leg on stderr (the JSON result stays the only line on stdout)
        print(f"bench.py: {leg} done at {time.perf_counter() - t_start:.1f} s", file=sys.stderr, flush=True)

    # launch plans of every product pass, timed at set-up (outside the timed solves)
    tuner = LaunchTuner() if args.tune else False
    shape = (256, 128, 128)
    rho, E, force, mask, ops = problem(shape, seed=3)
    ndof = force.numel()
    out = {}
    solvers = {}
    precs = ("mixed", "fp64") if args.precision == "both" else (args.precision,)
    for prec in precs:
        s = GMGSolver(shape, ke, mask, E, smoother="chebyshev", degree=3, precision=prec, graph=bool(args.graph),
                      tune=tuner, lower_frac=1.0 / 15.0, coarse_degree=2, coarse_level=2)
        solvers[prec] = s
        dt, st = timed_solves(s, force, args.steps, args.warmup)
        its = sorted({x[0] for x in st})
        out[prec] = {"seconds_per_solve": dt, "solves_per_s": 1.0 / dt, "iterations": its,
                     "converged": all(x[1] for x in st), "max_rel_residual": max(x[2] for x in st)}
        progress(f"{prec} solve")
    if args.only == "solve" or args.precision != "both":
        print(json.dumps({"only": "solve", "solves": out}), flush=True)
        return
    # the precision scheme C3 names: FP64 iterative refinement around the FP32 GMG-FCG
    s_ir = GMGSolver(shape, ke, mask, E, smoother="chebyshev", degree=3, precision="ir", graph=bool(args.graph),
                     tune=tuner, lower_frac=1.0 / 15.0, coarse_degree=2, coarse_level=2)
    dt, st = timed_solves(s_ir, force, max(1, args.steps // 2), 1)
    out["ir"] = {"seconds_per_solve": dt, "solves_per_s": 1.0 / dt, "inner_iterations": sorted({x[0] for x in st}),
                 "outer_corrections": s_ir.last_outer, "converged": all(x[1] for x in st),
                 "max_rel_residual": max(x[2] for x in st)}
    solvers["ir"] = s_ir
    # BF16-storage ablation: the mixed solve with the finest level's smoothing and residual passes
    # reading the V-cycle iterate through a bfloat16 copy
    s_bf = GMGSolver(shape, ke, mask, E, smoother="chebyshev", degree=3, precision="mixed", graph=bool(args.graph),
                     tune=tuner, lower_frac=1.0 / 15.0, coarse_degree=2, coarse_level=2, vcycle_storage="bf16")
    dt, st = timed_solves(s_bf, force, max(1, args.steps // 2), 1)
    out["mixed_bf16_storage"] = {"seconds_per_solve": dt, "solves_per_s": 1.0 / dt,
                                 "iterations": sorted({x[0] for x in st}), "converged": all(x[1] for x in st),
                                 "max_rel_residual": max(x[2] for x in st)}
    solvers["mixed_bf16_storage"] = s_bf
    progress("ir and bf16-storage solves")
    # error of every solution against the FP64 oracle (the constrained residual of the NumPy
    # oracle's product on the host) and against the FP64 GPU solve
    x64 = solvers["fp64"].solve(force)[0].clone()
    for prec in ("mixed", "fp64", "ir", "mixed_bf16_storage"):
        x = solvers[prec].solve(force)[0].clone()
        out[prec]["oracle_rel_residual"] = oracle_residual(shape, ke, E, mask, force, x)
        if prec != "fp64":
            out[prec]["rel_diff_vs_fp64_solution"] = float((x - x64).norm() / x64.norm())
    lambdas = solvers["fp64"].M.lambdas
    progress("oracle residuals")
    iters64 = out["fp64"]["iterations"][0]
    del solvers, s_ir, s_bf, x64, x
    torch.cuda.empty_cache()
    # the FP64 GMG-FCG of the 128 x 64 x 64 configuration
    shape2 = (128, 64, 64)
    _, E2, f2, m2, _ = problem(shape2, seed=2)
    s2 = GMGSolver(shape2, ke, m2, E2, smoother="chebyshev", degree=3, precision="fp64", coarse="galerkin",
                   graph=bool(args.graph), tune=tuner, lower_frac=1.0 / 15.0)
    dt2, st2 = timed_solves(s2, f2, args.steps, args.warmup)
    out["c2_fp64"] = {"shape": list(shape2), "seconds_per_solve": dt2, "solves_per_s": 1 / dt2,
                      "iterations": sorted({x[0] for x in st2}), "converged": all(x[1] for x in st2),
                      "oracle_rel_residual": oracle_residual(shape2, ke, E2, m2, f2, s2.solve(f2)[0])}
    del s2
    # high-contrast variant of the same grid (void, E = 1e-9, where rho <= 0.5): the rediscretised
    # hierarchy against the full Galerkin hierarchy, mixed precision, 200 iterations at most
    rho2 = problem(shape2, seed=2)[0]
    E2v = torch.where(rho2 > 0.5, E2, torch.full_like(E2, 1e-9))
    gal = {}
    for coarse in ("rediscretize", "galerkin-all"):
        t0 = time.perf_counter()
        sv = GMGSolver(shape2, ke, m2, E2v, smoother="chebyshev", degree=3, precision="mixed", coarse=coarse,
                       graph=bool(args.graph), tune=tuner)
        torch.cuda.synchronize()
        setup = time.perf_counter() - t0
        dtv, stv = timed_solves(sv, f2, max(1, args.steps // 5), 1)
        gal[coarse] = {"setup_s": setup, "seconds_per_solve": dtv, "iterations": sorted({x[0] for x in stv}),
                       "converged": all(x[1] for x in stv), "max_rel_residual": max(x[2] for x in stv)}
        del sv
    out["c2_void_mixed"] = {"shape": list(shape2), "void": "E = 1e-9 where rho <= 0.5", **gal}
    progress("128x64x64 solves")
    if not args.no_batch:
        out["c4_batched"] = timed_batch(c4_problems(), ke, max(1, args.steps // 2), 1, graph=bool(args.graph))
        torch.cuda.empty_cache()
        progress("batched solves")
    if args.simp_steps > 0:
        out["c5_simp"] = simp_leg(ke, steps=args.simp_steps)
        torch.cuda.empty_cache()
        progress("SIMP run")
    if not args.no_reference:
        out["reference_workloads"] = reference_workloads(max(2, args.steps // 4), 1, bool(args.graph), tuner)
        torch.cuda.empty_cache()
        progress("reference workloads")
    # the product under the library's automatic plan and under the tuned plan (the one the
    # solvers launch); the roofline is that of the tuned FP64 product at 256 x 128 x 128
    mv = {}
    for sh in ((256, 128, 128), (128, 64, 64), (384, 192, 192)):
        for d in ("float64", "float32"):
            key = f"{d}@{'x'.join(map(str, sh))}"
            mv[key] = tuned_matvec(sh, getattr(torch, d))
            mv[key]["auto_plan_us"] = matvec_timing(sh, getattr(torch, d))["us"]
    dom = mv["float64@256x128x128"]
    for parts in (8, 4):
        mv[f"x_slab_share_of_384x192x192_over_{parts}"] = {
            d: slab_share(getattr(torch, d), shape=(384 // parts, 192, 192)) for d in ("float64", "float32")}
    progress("product table")
    cpu = None
    if args.cpu_baseline == "sample":
        cpu = cpu_baseline(shape, E, mask, force, ke, lambdas, iters64)
    elif args.cpu_baseline == "full":
        cpu = cpu_baseline(shape, E, mask, force, ke, lambdas, iters64)
        cpu["protocol"] = cpu_baseline_full()
    line = {
        "metric": "CG solves/s at FP64-equiv rel.res 1e-8 (mixed vs FP64); matvec GDOF/s, %HBM",
        "value": out["mixed"]["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": out["mixed"]["seconds_per_solve"] * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 V-cycle and search direction in f64 outer CG",
        "data": "synthetic (seeded densities; no dataset)",
        "config": {"workload": "cantilever 256x128x128 (12,830,211 DOF), GMG-FCG 5 levels Chebyshev deg 3 on "
                               "[lmax/15, lmax] (deg 2 on levels 2-3), FP32 V-cycle and search direction in FP64 outer CG "
                               "(FP64 iterate and true residual), rtol 1e-8 on the FP64 true residual",
                   "ndof": ndof, "parallelism": "single GPU"},
        "fp64_solves_per_s": out["fp64"]["solves_per_s"],
        "mixed_speedup_vs_fp64": out["fp64"]["seconds_per_solve"] / out["mixed"]["seconds_per_solve"],
        "solves": out,
        "matvec": mv,
        "roofline": {"bound": "hbm", "achieved": dom["achieved_GBs"], "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": dom["frac_hbm"], "traffic": measured_traffic(),
                     "kernel": "parity_xmarch<double> (FP64 product, 256x128x128, tuned plan "
                               f"{dom['plan']['tile_name']} tile, lx {dom['plan']['lx']})",
                     "algorithmic_bytes_per_launch": dom["algorithmic_bytes"], "avg_launch_us": dom["us"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
