"""Study harness of the product and the solvers on one MI355X: inputs, correctness gates, paired
randomized timing with a bootstrap decision, the W2b solve protocol and result.json records.

It restates the reference's harness (code/hex_modal_cuda/common.py:57-117 and run_gpu.py,
experiments/session_i/make_inputs.py:31-49 and run_gpu.py:90-104,255-330) for the paths of this
library:

- inputs: `cantilever_boundary(shape)` (every DOF of the x = 0 face clamped; a consistent line load
  of total 1 in -z along y at x = nx, z = 0), `refine_design(rho, shape, factor)` (nearest-neighbour
  refinement of a base design), `load_case` (both at a study case's resolution), all in the author
  layout; `to_author` (natural (z, y, x[, 3]) -> author layout) for the reference's input files;
- correctness: `product_check` (relative and per-sqrt(DOF) absolute L2 limits), `exactness` (zero in,
  zero out; identity rows at fixed DOFs), `physical_check` (true residual recomputed by the dense
  node kernel, an arithmetic independent of the parity product, and the energy / force-work check);
- timing: `paired_rounds` (each round runs every path once, in a seeded random order, HIP events
  around a batch of calls) and `timing_decision` (per baseline the paired median ratio and the 5 %
  bootstrap lower bound of the median; a primary case passes at median >= 1.15 and bound > 1.05);
- `solve_case`: the W2b solve of the scale study through the drop-in voxel_mg (voxel hierarchy
  halved to x <= 16, damped Jacobi V(4,4) with omega_l = 4/(3 lambda_l) from 30 seeded power steps,
  dense Cholesky coarsest, FP64 flexible CG to 1e-8 with the true residual every iteration);
- `write_result`: result.json written atomically, NaN and infinities rejected; `check_record`
  validates a solve record against RESULT_SCHEMA.

    python -m voxelhip.harness --inputs DIR --case q64x2 --out result.json
"""
from __future__ import annotations

import argparse
import hashlib
import json
import math
import time
from pathlib import Path

import numpy as np
import torch

from . import ops

# the scale study's declared protocol (experiments/session_i/protocol.json: multigrid, solver,
# material, order_seed) and its input cases (make_inputs.py:23-29: base design, upsampling factor)
PROTOCOL = {
    "multigrid": {"max_coarse_x": 16, "pre": 4, "post": 4, "power_iterations": 30},
    "solver": {"requested_rtol": 1e-8, "true_residual_limit": 1e-7, "energy_force_relative_limit": 1e-6,
               "maximum_iterations": 200, "check_every": 1},
    "material": {"Emin": 1e-6, "Emax": 1.0, "penal": 3.0, "nu": 0.3},
    "order_seed": 2026092621,
    "correctness": {"relative_l2": 1e-11, "absolute_per_sqrt_dof": 1e-12},
    "timing": {"rounds": 5, "calls_per_batch": 10, "bootstrap_seed": 20260923, "bootstrap_resamples": 10000,
               "median_ratio_gate": 1.15, "lower_bound_gate": 1.05},
}
CASES = {"q4x2": ("smoke-q4.npz", 2), "q64x1": ("optimized-q64.npz", 1), "q96x1": ("optimized-q96.npz", 1),
         "q64x2": ("optimized-q64.npz", 2),
         "q96x2": ("optimized-q96.npz", 2), "q96x3": ("optimized-q96.npz", 3)}

# fields of a solve record and their types (result.json "solves" rows)
RESULT_SCHEMA = {
    "case": str, "shape": list, "ndof": int, "iterations": int, "native_converged": bool,
    "native_true_residual": float, "independent_true_residual": float, "compliance": float,
    "force_work": float, "energy_force_relative_error": float, "solution_norm": float,
    "solution_sha256": str, "omegas": list, "lambdas": list, "setup_seconds": float,
    "solver_seconds": float, "physical_passed": bool,
}


# ---------------------------------------------------------------------------------- inputs
def cantilever_boundary(shape):
    """Author-layout (mask, force) of the study cantilever at any resolution: every DOF of the
    node plane ix = 0 fixed; a load of total 1 in -z along the line ix = nx, iz = 0, shared
    trapezoidally over y (weight 1/ny per node, half at iy = 0 and iy = ny)."""
    nx, ny, nz = (int(n) for n in shape)
    plane = 3 * (ny + 1) * (nz + 1)                # DOFs of one x node plane
    mask = np.ones(3 * (nx + 1) * (ny + 1) * (nz + 1))
    mask[:plane] = 0.0
    force = np.zeros_like(mask)
    iy = np.arange(ny + 1)
    weight = np.where((iy == 0) | (iy == ny), 0.5, 1.0) / ny
    force[nx * plane + 3 * (nz + 1) * iy + 2] = -weight
    return mask, force


def refine_design(rho, shape, factor):
    """Author-layout densities of the design `rho` (natural (z, y, x) order, `shape` elements)
    refined `factor` times per axis: fine element (ex, ey, ez) takes the value of coarse element
    (ex // f, ey // f, ez // f)."""
    nx, ny, nz = (int(n) for n in shape)
    f = int(factor)
    base = np.asarray(rho).reshape(nz, ny, nx)
    ex = np.arange(nx * f)[:, None, None] // f
    ey = np.arange(ny * f)[None, :, None] // f
    ez = np.arange(nz * f)[None, None, :] // f
    return np.ascontiguousarray(base[ez, ey, ex]).ravel()


def to_author(a, shape, nodal=False):
    """Natural (z, y, x[, 3]) order -> author (x, y, z[, 3]) order."""
    x, y, z = (int(s) for s in shape)
    sizes = (z + 1, y + 1, x + 1, 3) if nodal else (z, y, x)
    axes = (2, 1, 0, 3) if nodal else (2, 1, 0)
    return np.ascontiguousarray(np.asarray(a).reshape(sizes).transpose(axes)).ravel()


def load_case(inputs, case):
    """The named study case built from its base design in directory `inputs`: the design refined
    by the case's factor and the boundary data of the refined grid, all in the author layout."""
    base, factor = CASES[case]
    with np.load(Path(inputs) / base) as f:
        shape = [int(v) for v in f["shape"]]
        rho, ke = np.asarray(f["rho"]), np.asarray(f["ke"])
    fine = np.array([n * factor for n in shape])
    mask, force = cantilever_boundary(fine)
    return {"shape": fine, "rho": refine_design(rho, shape, factor), "mask": mask, "force": force, "ke": ke}


# ----------------------------------------------------------------------------- correctness
def relative(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    b = b.detach().cpu().numpy() if isinstance(b, torch.Tensor) else np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def product_check(actual, expected, limits=None):
    """L2 error against relative_l2 * |expected| + absolute_per_sqrt_dof * sqrt(n)."""
    c = limits or PROTOCOL["correctness"]
    a = actual.detach().cpu().numpy() if isinstance(actual, torch.Tensor) else np.asarray(actual)
    e = expected.detach().cpu().numpy() if isinstance(expected, torch.Tensor) else np.asarray(expected)
    err = float(np.linalg.norm(a - e))
    limit = c["relative_l2"] * float(np.linalg.norm(e)) + c["absolute_per_sqrt_dof"] * math.sqrt(e.size)
    return {"relative_l2": relative(a, e), "absolute_l2": err, "limit_l2": limit,
            "passed": bool(np.isfinite(a).all() and err <= limit)}


def exactness(apply, u, mask):
    """Zero input gives an exactly zero output; fixed DOFs return the input exactly."""
    z = apply(torch.zeros_like(u))
    y = apply(u)
    fixed = mask == 0
    return {"zero_exact": bool(torch.count_nonzero(z) == 0), "fixed_identity_exact": bool(torch.equal(y[fixed], u[fixed]))}


def constrained_dense(ke, E, shape, mask):
    """A u = M K M u + (I - M) u with the dense 24x24 node kernel (the independent arithmetic)."""
    def apply(u):
        y = torch.empty_like(u)
        ops.dense_product(ke, E, u * mask, y, shape)
        return y * mask + u * (1 - mask)
    return apply


def physical_check(shape, ke, E, mask, force, u, native_ok, native_rel, limits=None):
    """True residual by the dense kernel, compliance sum E_e u_e^T ke u_e against the force work."""
    s = limits or PROTOCOL["solver"]
    import voxel_kernels as vk
    blocks = vk.Modal(ke).blocks
    independent = relative(constrained_dense(ke, E, shape, mask)(u), force)
    ce = torch.empty_like(E)
    ops.element_energy(blocks, u * mask, ce, shape)
    compliance = float(E @ ce)
    work = float(force @ u)
    err = abs(compliance / work - 1.0) if work != 0 else float("inf")
    return {"native_converged": bool(native_ok), "native_true_residual": float(native_rel),
            "independent_true_residual": independent, "compliance": compliance, "force_work": work,
            "energy_force_relative_error": err,
            "physical_passed": bool(native_ok and native_rel <= s["requested_rtol"]
                                    and independent <= s["true_residual_limit"]
                                    and err <= s["energy_force_relative_limit"])}


# ---------------------------------------------------------------------------------- timing
def paired_rounds(paths, rounds=None, calls=None, seed=None, case=""):
    """paths: name -> zero-argument callable launching one product.  Every round runs every path
    once, in an order drawn from the seed; each entry times `calls` launches between HIP events."""
    t = PROTOCOL["timing"]
    rounds = t["rounds"] if rounds is None else rounds
    calls = t["calls_per_batch"] if calls is None else calls
    rng = np.random.default_rng(PROTOCOL["order_seed"] if seed is None else seed)
    names = list(paths)
    for fn in paths.values():          # warm-up
        fn()
    torch.cuda.synchronize()
    rows = []
    for r in range(rounds):
        for pos, i in enumerate(rng.permutation(len(names))):
            name = names[int(i)]
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            e0.record()
            for _ in range(calls):
                paths[name]()
            e1.record()
            e1.synchronize()
            rows.append({"id": case, "round": r, "position": pos, "path": name, "calls": calls,
                         "gpu_ms_per_call": e0.elapsed_time(e1) / calls})
    return rows


def timing_decision(timings, candidate, primary=True, seed=None, resamples=None):
    """Per baseline path: the median over rounds of baseline / candidate time and the 5 % quantile
    of the bootstrap distribution of that median (rounds resampled with replacement)."""
    t = PROTOCOL["timing"]
    rng = np.random.default_rng(t["bootstrap_seed"] if seed is None else seed)
    resamples = t["bootstrap_resamples"] if resamples is None else resamples
    by_path = {}
    for row in sorted(timings, key=lambda r: r["round"]):
        by_path.setdefault(row["path"], []).append(row["gpu_ms_per_call"])
    cand = np.asarray(by_path[candidate])
    if not (cand > 0).all() or not np.isfinite(cand).all():
        raise ValueError("candidate timings must be positive and finite")
    out = []
    for name, vals in by_path.items():
        if name == candidate:
            continue
        base = np.asarray(vals)
        if base.shape != cand.shape:
            raise ValueError("every path needs one timing per round")
        ratios = base / cand
        idx = rng.integers(0, len(ratios), size=(resamples, len(ratios)))
        medians = np.median(ratios[idx], axis=1)
        med, lower = float(np.median(ratios)), float(np.quantile(medians, 0.05))
        out.append({"baseline": name, "candidate_median_ms": float(np.median(cand)),
                    "baseline_median_ms": float(np.median(base)), "paired_median_ratio": med,
                    "paired_bootstrap_lower_95": lower, "primary": bool(primary),
                    "passed": med >= t["median_ratio_gate"] and lower > t["lower_bound_gate"]})
    return {"comparisons": out, "timing_gate_passed": all(c["passed"] for c in out if c["primary"])}


# ---------------------------------------------------------------------------------- solves
def solve_case(data, case="", device=None, vcycle_dtype=np.float64, seed=None):
    """The W2b multigrid solve of one case (author-layout data of load_case) through the drop-in
    voxel_mg: FP64 flexible CG around a V-cycle in `vcycle_dtype` (np.float32: the reference's
    mg32 arms, FP32 products and smoothing with an FP64 coarsest factor cast to FP32, the omegas
    from the FP64 hierarchy; `seed` of the power iteration, default the study's order seed);
    returns the record and the solution."""
    import voxel_mg as mg
    import voxelhip
    p = PROTOCOL
    dev = device or torch.device("cuda", torch.cuda.current_device())
    shape = tuple(int(s) for s in data["shape"])
    ke = np.asarray(data["ke"], dtype=np.float64)
    emin = p["material"]["Emin"]
    t0 = time.perf_counter()
    E = torch.from_numpy(emin + (1 - emin) * np.asarray(data["rho"], dtype=np.float64) ** p["material"]["penal"]).to(dev)
    mask = torch.from_numpy(np.asarray(data["mask"], dtype=np.float64)).to(dev)
    force = torch.from_numpy(np.asarray(data["force"], dtype=np.float64)).to(dev)
    shapes = mg.hierarchy_shapes(shape, p["multigrid"]["max_coarse_x"])
    prods = [mg.HipProduct(s, ke * 2 ** i, np.float64) for i, s in enumerate(shapes)]
    # coarsest operator: moduli averaged and mask injected level by level, assembled and factored
    # on the device (the reference assembles on the host and calls np.linalg.cholesky)
    Ec, mc = E, mask
    for i in range(1, len(shapes)):
        Ec = mg.coarsen_modulus(None, Ec, shapes[i - 1])
        mc = mg.coarsen_mask(None, mc, shapes[i - 1])
    kc = torch.from_numpy(np.ascontiguousarray(ke * 2 ** (len(shapes) - 1))).to(dev)
    factor = torch.linalg.cholesky(ops.assemble_dense(kc, Ec, shapes[-1], mask=ops.Mask.from_values(mc, shapes[-1])))
    xp = voxelhip.xp
    M = mg.Multigrid(xp, shapes, prods, E, mask, factor, np.float64, pre=p["multigrid"]["pre"],
                     post=p["multigrid"]["post"])
    omegas, lambdas = mg.spectral_omegas(xp, M.levels, p["multigrid"]["power_iterations"],
                                         p["order_seed"] if seed is None else seed)
    M.omegas = omegas
    if np.dtype(vcycle_dtype) != np.float64:
        prods = [mg.HipProduct(s, ke * 2 ** i, vcycle_dtype) for i, s in enumerate(shapes)]
        M = mg.Multigrid(xp, shapes, prods, E, mask, factor, vcycle_dtype, pre=p["multigrid"]["pre"],
                         post=p["multigrid"]["post"], omegas=omegas)
        fine64 = mg.Level(xp, shape, mg.HipProduct(shape, ke, np.float64), E, mask, np.float64)
    else:
        fine64 = M.levels[0]
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    t1 = time.perf_counter()
    x, it, ok, rel, _ = mg.fcg(xp, fine64.apply, force, M, p["solver"]["requested_rtol"],
                               p["solver"]["maximum_iterations"], check_every=p["solver"]["check_every"])
    torch.cuda.synchronize()
    seconds = time.perf_counter() - t1
    rec = {"case": case, "shape": list(shape), "ndof": int(force.numel()), "iterations": int(it),
           "omegas": [float(v) for v in omegas], "lambdas": [float(v) for v in lambdas],
           "setup_seconds": setup, "solver_seconds": seconds,
           "solution_norm": float(x.norm()),
           "solution_sha256": hashlib.sha256(x.detach().cpu().numpy().tobytes()).hexdigest()}
    rec.update(physical_check(shape, ke, E, mask, force, x, ok, rel))
    return rec, x


# ---------------------------------------------------------------------------------- records
def check_record(rec):
    """Raises if a solve record misses a field of RESULT_SCHEMA or has the wrong type."""
    for key, typ in RESULT_SCHEMA.items():
        if key not in rec:
            raise ValueError(f"record misses {key}")
        if typ is float and isinstance(rec[key], (int, float)) and not isinstance(rec[key], bool):
            continue
        if not isinstance(rec[key], typ):
            raise ValueError(f"{key}: {type(rec[key]).__name__}, expected {typ.__name__}")
    return True


def write_result(path, result):
    """Atomic JSON write (a temporary file renamed over the target); NaN and infinities raise."""
    path = Path(path)
    tmp = path.with_name(path.name + ".tmp")
    tmp.write_text(json.dumps(result, indent=2, allow_nan=False) + "\n")
    tmp.replace(path)
    return path


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--inputs", type=Path, required=True, help="directory holding the base designs (npz)")
    ap.add_argument("--case", choices=sorted(CASES), required=True)
    ap.add_argument("--out", type=Path, required=True)
    args = ap.parse_args(argv)
    data = load_case(args.inputs, args.case)
    rec, _ = solve_case(data, args.case)
    check_record(rec)
    write_result(args.out, {"status": "completed", "protocol": PROTOCOL, "solves": [rec]})
    print(json.dumps({k: rec[k] for k in ("case", "iterations", "native_true_residual", "compliance",
                                          "physical_passed")}), flush=True)


if __name__ == "__main__":
    main()
