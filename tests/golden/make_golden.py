"""Writes the golden fixtures under tests/golden/ from the reference checkout at /root/reference.

Run only where that checkout exists (the tests read the committed fixtures):  python tests/golden/make_golden.py

1. Data files the reference itself holds (copied byte-for-byte as data, no code):
   smoke-q4.npz, optimized-q64.npz            experiments/session_a/inputs/
   ke_unit_3d.npy                             experiments/session_f/
   smoke-q4 solution vectors of GPU runs      results/session-a-20260923/remote/run-01/,
                                              results/session-d-20260923/remote/run-01/
2. Known answers extracted from the reference's result.json records -> known_answers.json.
3. Input/output vectors of the functions that no closed form pins bit-for-bit
   (spectral_omegas, one Multigrid V-cycle, the fcg iterate), computed by importing the
   reference's src/voxel_mg.py in this container -> mg_golden.npz.
"""
from __future__ import annotations

import importlib.util
import json
import shutil
import sys
from pathlib import Path

import numpy as np

REF = Path("/root/reference")
HERE = Path(__file__).resolve().parent

DATA = {
    "smoke-q4.npz": "experiments/session_a/inputs/smoke-q4.npz",
    "optimized-q64.npz": "experiments/session_a/inputs/optimized-q64.npz",
    "ke_unit_3d.npy": "experiments/session_f/ke_unit_3d.npy",
    "smoke-q4--pcg64-modal8--tol-00.npz":
        "results/session-a-20260923/remote/run-01/smoke-q4--pcg64-modal8--tol-00.npz",
    "smoke-q4--modal8--fixed-00.npz":
        "results/session-a-20260923/remote/run-01/smoke-q4--modal8--fixed-00.npz",
    "smoke-q4--ir-modal8-modal8_32--tol-00.npz":
        "results/session-a-20260923/remote/run-01/smoke-q4--ir-modal8-modal8_32--tol-00.npz",
    "smoke-q4--mg64-modal8--tol-00.npz":
        "results/session-d-20260923/remote/run-01/smoke-q4--mg64-modal8--tol-00.npz",
    "smoke-q4--mg32-modal8-modal8_32--tol-00.npz":
        "results/session-d-20260923/remote/run-01/smoke-q4--mg32-modal8-modal8_32--tol-00.npz",
}


def copy_data():
    for name, rel in DATA.items():
        shutil.copyfile(REF / rel, HERE / name)


def known_answers():
    out = {}
    a = json.loads((REF / "results/session-a-20260923/remote/run-01/result.json").read_text())
    for row in a["to_tolerance"]:
        if row["case"] == "smoke-q4" and row["rep"] == 0:
            keep = {k: row[k] for k in ("iterations", "outer_iterations", "inner_iterations",
                                        "native_true_residual", "compliance", "force_work")
                    if k in row}
            out["session_a/" + row["method"]] = keep
    for row in a["fixed_work"]:
        if row["case"] == "smoke-q4" and row["path"] == "modal8" and row["round"] == 0:
            out["session_a/fixed/modal8"] = {k: row[k] for k in
                                             ("iterations", "matvec_calls", "native_true_residual")}
    for row in a["product_checks"]:
        if row["case"] == "smoke-q4":
            out.setdefault("session_a/product_checks", {})[row["path"]] = row["relative_l2"]
    d = json.loads((REF / "results/session-d-20260923/remote/run-01/result.json").read_text())
    for row in d["to_tolerance"]:
        if row["case"] == "smoke-q4" and row["rep"] == 0:
            out["session_d/" + row["arm"]] = {k: row[k] for k in
                                              ("iterations", "outer_matvecs", "native_true_residual",
                                               "compliance", "force_work") if k in row}
    # Published MG-FCG known answers on the reference's large states (paper/session_d_tables.md,
    # paper/supplement/scale-solves.csv), kept as numbers for the parity-anchor runs.
    out["published"] = {
        "optimized-q64/mg64/iterations": 13,
        "optimized-q64/omegas": [0.441, 0.437, 0.43],
        "uniform-q64/mg64/iterations": 9,
        "smoke-q4/mg64/iterations": 1,
        "q64x2/mg64/iterations": 14, "q64x2/mg64/exit_rel": 2.06235e-9,
        "q96x2/mg64/iterations": 24, "q96x2/mg64/exit_rel": 4.4318e-9,
        "jacobi_pcg_iterations": {"smoke-q4": 100, "optimized-q64": 4600,
                                  "optimized-q96": 6900, "uniform-q64": 1050},
    }
    (HERE / "known_answers.json").write_text(json.dumps(out, indent=1, sort_keys=True) + "\n")


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def mg_golden():
    mg = _load("ref_voxel_mg", REF / "src/voxel_mg.py")
    common = _load("ref_common", REF / "experiments/session_a/dependencies/common.py")
    d = np.load(REF / "experiments/session_a/inputs/smoke-q4.npz")
    KE = d["ke"]
    out = {}

    def case(q, seed):
        if seed is None:
            shape = tuple(int(s) for s in d["shape"])
            rho, mask, force = d["rho"], d["mask"], d["force"]
        else:
            shape = (2 * q, q, q)
            nx, ny, nz = shape
            rng = np.random.default_rng(seed)
            rho = 0.05 + 0.95 * rng.random(nx * ny * nz) ** 2
            m = np.ones((nz + 1, ny + 1, nx + 1, 3)); m[:, :, 0, :] = 0
            f = np.zeros((nz + 1, ny + 1, nx + 1, 3)); f[0, :, nx, 2] = -1.0 / (ny + 1)
            mask, force = m.ravel(), f.ravel()
        E = common.to_author(1e-6 + (1 - 1e-6) * rho ** 3, shape)
        return shape, E, common.to_author(mask, shape, True), common.to_author(force, shape, True)

    def build(shape, E, mask, dtype=np.float64, max_coarse_x=2):
        shapes = mg.hierarchy_shapes(shape, max_coarse_x)
        products = [mg.NumpyProduct(s, KE * 2 ** i, dtype) for i, s in enumerate(shapes)]
        Ec, mc = E, mask
        for i in range(1, len(shapes)):
            Ec, mc = mg.coarsen_modulus(np, Ec, shapes[i - 1]), mg.coarsen_mask(np, mc, shapes[i - 1])
        A = mg.assemble_dense(shapes[-1], KE * 2 ** (len(shapes) - 1), Ec, mc)
        M = mg.Multigrid(np, shapes, products, E, mask, np.linalg.cholesky(A), dtype)
        M.omegas, lams = mg.spectral_omegas(np, M.levels)
        return shapes, M, lams

    for tag, (q, seed) in {"smoke": (4, None), "q8s3": (8, 3)}.items():
        shape, E, mask, force = case(q, seed)
        shapes, M, lams = build(shape, E, mask)
        rng = np.random.default_rng(11)
        r = rng.standard_normal(mask.size) * mask
        out[f"{tag}/E"], out[f"{tag}/mask"], out[f"{tag}/force"] = E, mask, force
        out[f"{tag}/shape"] = np.array(shape)
        out[f"{tag}/lambdas"] = np.array(lams)
        out[f"{tag}/vcycle_in"] = r
        out[f"{tag}/vcycle_out"] = M(r)
        prod = mg.NumpyProduct(shape, KE)
        apply = lambda u: prod.raw(u * mask, E) * mask + u * (1 - mask)  # noqa: E731
        x, it, ok, rel, mv = mg.fcg(np, apply, force, M, 1e-8, 300)
        out[f"{tag}/fcg_x"], out[f"{tag}/fcg_stats"] = x, np.array([it, int(ok), rel, mv])
        if tag == "q8s3":
            shapes32, M32, lams32 = build(shape, E, mask, dtype=np.float32)
            out[f"{tag}/lambdas32"] = np.array(lams32)
            out[f"{tag}/vcycle32_out"] = M32(r)
    np.savez_compressed(HERE / "mg_golden.npz", **out)


def main():
    if not REF.exists():
        sys.exit("needs the reference checkout at /root/reference")
    copy_data()
    known_answers()
    mg_golden()
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
