"""CPU baseline of the hot path: the NumPy oracle timed on the host cores.

TEST INFRASTRUCTURE ONLY: run by bench.py (--cpu-baseline full), in one child process per thread
setting, because the BLAS thread count has to be fixed before NumPy loads:

    OMP_NUM_THREADS=<n> python oracle/cpu_baseline.py [--reps 5] [--c2-cap 90]

It times, each as the median of `reps` repetitions after one warm-up (wall clock,
time.perf_counter):
  1. one FP64 product (NumpyProduct.raw) at 128x64x64 and 256x128x128;
  2. the FP64 Jacobi-PCG to a relative residual of 1e-8 (true residual every 50 iterations) of the
     32x16x16 cantilever (Emin 1e-9, densities U[0.01, 1], seed 1);
  3. the FP64 GMG-FCG to 1e-8 of the 128x64x64 cantilever (seed 2, densities smoothed by the
     radius-1.5 cone filter; 4 levels, Chebyshev-Jacobi degree 3, rediscretised dense coarsest),
     one run if a solve exceeds `--c2-cap` seconds, the median otherwise.
np.add.at (the oracle product's scatter) is single-threaded in either setting; the thread count
reaches the BLAS product u[edof] @ ke.T.  Prints one JSON object.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))

from oracle import hexref as hx  # noqa: E402
from oracle import mgref as mgr  # noqa: E402
from oracle import simpref as sr  # noqa: E402
from oracle import solvers as sv  # noqa: E402


def timed(fn, reps, warmup=1):
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts, out


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--c2-cap", type=float, default=30.0)
    args = ap.parse_args()
    ke = 2 * hx.quadrature([1, 1, 1], 0.3)
    res = {"omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "cpu_model": cpu_model(),
           "nproc": os.cpu_count(), "reps": args.reps}
    for shape, seed in (((128, 64, 64), 2), ((256, 128, 128), 3)):
        _, E, mask, force = sr.cantilever(shape, seed=seed, rmin=1.5)
        prod = hx.NumpyProduct(shape, ke)
        u = np.random.default_rng(0).standard_normal(force.size)
        t, _, _ = timed(lambda: prod.raw(u, E), args.reps)
        res[f"matvec_{shape[0]}x{shape[1]}x{shape[2]}"] = {"median_s": t, "gdof_s": force.size / t / 1e9}
        print(f"matvec {shape}: {t:.3f} s", file=sys.stderr, flush=True)
    # C1: Jacobi-PCG
    shape = (32, 16, 16)
    _, E, mask, force = sr.cantilever(shape, seed=1)
    prod = hx.NumpyProduct(shape, ke)
    inv = 1.0 / np.where(mask != 0, prod.diagonal(E), 1.0)
    A = lambda v: prod.raw(v * mask, E) * mask + v * (1 - mask)  # noqa: E731
    t, _, (x, it, ok, rel) = timed(lambda: sv.pcg(A, force, inv, tol=1e-8, maxiter=20000, mask=mask,
                                                   check_every=50), args.reps)
    res["c1_jacobi_pcg"] = {"median_s": t, "solves_per_s": 1.0 / t, "iterations": int(it), "converged": bool(ok),
                            "rel_residual": float(rel)}
    print(f"C1 pcg: {t:.3f} s, {it} iterations", file=sys.stderr, flush=True)
    # C2: GMG-FCG
    shape = (128, 64, 64)
    _, E, mask, force = sr.cantilever(shape, seed=2, rmin=1.5)
    _, M = mgr.build_hierarchy(shape, ke, E, mask, max_coarse_x=16, smoother="chebyshev", degree=3)
    fine = hx.NumpyProduct(shape, ke)
    A = lambda v: fine.raw(v * mask, E) * mask + v * (1 - mask)  # noqa: E731
    solve = lambda: mgr.fcg(A, force, M, 1e-8, 200)  # noqa: E731
    t0 = time.perf_counter()
    x, it, ok, rel, _ = solve()
    first = time.perf_counter() - t0
    if first > args.c2_cap:
        t, runs = first, [first]
    else:
        t, runs, _ = timed(solve, args.reps, warmup=0)
    res["c2_gmg_fcg"] = {"median_s": t, "runs": len(runs), "solves_per_s": 1.0 / t, "iterations": int(it),
                         "converged": bool(ok), "rel_residual": float(rel), "cap_s": args.c2_cap}
    print(f"C2 fcg: {t:.3f} s, {it} iterations", file=sys.stderr, flush=True)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
