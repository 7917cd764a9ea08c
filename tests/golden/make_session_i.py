"""Writes tests/golden/session_i_known_answers.json and copies optimized-q96.npz from the reference
checkout at /root/reference (only where that checkout exists; the tests read the committed fixtures).

    python tests/golden/make_session_i.py

Source: the reference's scale study (experiments/session_i/protocol.json, its W2b arms) and the
result.json records it left for three GPUs (results/session-i-{h100,a100,rtx4090}-20260926/remote/
run-01/result.json).  Per case (q96x1, q64x2, q96x2, q96x3: the optimized q96 / q64 cantilevers
upsampled by 1, 2, 2, 3 with the boundary data regenerated) it keeps the per-level omegas and
lambdas, and for the all-FP64 multigrid arms (mg64-*: voxel hierarchy, damped Jacobi V(4,4), FP64
flexible CG to 1e-8 with the true residual every iteration) the iteration counts and the ranges of
the true residual, compliance, force work and solution norm over arms, repetitions and GPUs.
Data only: numbers and the reference's own input file, no code.
"""
from __future__ import annotations

import json
import shutil
from pathlib import Path

REF = Path("/root/reference")
HERE = Path(__file__).resolve().parent
RUNS = [REF / f"results/session-i-{d}-20260926/remote/run-01/result.json" for d in ("h100", "a100", "rtx4090")]


def main():
    shutil.copyfile(REF / "experiments/session_i/inputs/optimized-q96.npz", HERE / "optimized-q96.npz")
    proto = json.loads((REF / "experiments/session_i/protocol.json").read_text())
    cases = {}
    for path in RUNS:
        r = json.loads(path.read_text())
        for c in r["cases"]:
            row = cases.setdefault(c["case"], {"shapes": c["shapes"], "omegas": c["omegas"], "lambdas": c["lambdas"],
                                              "iterations": set(), "runs": 0})
            assert row["shapes"] == c["shapes"]
        for t in r["to_tolerance"]:
            if not t["arm"].startswith("mg64"):
                continue
            row = cases[t["case"]]
            row["iterations"].add(int(t["iterations"]))
            row["runs"] += 1
            for key in ("native_true_residual", "compliance", "force_work", "solution_norm"):
                lo, hi = row.get(key, (float("inf"), -float("inf")))
                row[key] = (min(lo, t[key]), max(hi, t[key]))
    out = {"protocol": {"multigrid": proto["multigrid"], "solver": proto["solver"], "material": proto["material"],
                        "order_seed": proto["order_seed"],
                        "cases": {c["id"]: c for c in proto["cases"]}},
           "inputs": {"q96x1": ["optimized-q96.npz", 1], "q64x2": ["optimized-q64.npz", 2],
                      "q96x2": ["optimized-q96.npz", 2], "q96x3": ["optimized-q96.npz", 3]},
           "cases": {}}
    for cid, row in cases.items():
        out["cases"][cid] = {k: (sorted(v) if isinstance(v, set) else v) for k, v in row.items()}
    (HERE / "session_i_known_answers.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
