"""Writes tests/golden/sessions_ad_known_answers.json and copies uniform-q64.npz from the reference
checkout at /root/reference (only where that checkout exists; the tests read the committed fixtures).

    python tests/golden/make_sessions_ad.py

Sources: the reference's session A (Jacobi-PCG and FP32-in-FP64 iterative refinement to a true
relative residual of 1e-8, check every 50 iterations, Emin 1e-6;
experiments/session_a/protocol.json) and session D (flexible CG around the voxel GMG: damped
Jacobi V(4,4), omegas from 30 power steps seeded with the protocol's order seed, FP64 or FP32
V-cycle; experiments/session_d/protocol.json) result.json records
(results/session-{a,d}-20260923/remote/run-01/result.json).  Per case and method/arm it keeps the
set of iteration counts (outer and inner for refinement) and the ranges of the native true
residual, the compliance sum E_e u_e^T ke u_e and the force work f.u over repetitions; per
session-D case the hierarchy, omegas and lambdas.  Data only: numbers and the reference's own
input file, no code.
"""
from __future__ import annotations

import json
import shutil
from pathlib import Path

REF = Path("/root/reference")
HERE = Path(__file__).resolve().parent


def _agg(rows, key):
    out = {}
    for t in rows:
        if t["case"] == "smoke-q4":
            continue
        r = out.setdefault(f"{t['case']}/{t[key]}", {"iterations": set(), "outer_iterations": set(),
                                                      "native_true_residual": [], "compliance": [],
                                                      "force_work": []})
        for k in ("iterations", "outer_iterations"):
            if t.get(k) is not None:
                r[k].add(int(t[k]))
        for k in ("native_true_residual", "compliance", "force_work"):
            r[k].append(float(t[k]))
    for r in out.values():
        for k in ("iterations", "outer_iterations"):
            r[k] = sorted(r[k])
        for k in ("native_true_residual", "compliance", "force_work"):
            r[k] = [min(r[k]), max(r[k])]
    return out


def main():
    shutil.copyfile(REF / "experiments/session_a/inputs/uniform-q64.npz", HERE / "uniform-q64.npz")
    a = json.loads((REF / "results/session-a-20260923/remote/run-01/result.json").read_text())
    d = json.loads((REF / "results/session-d-20260923/remote/run-01/result.json").read_text())
    pd = json.loads((REF / "experiments/session_d/protocol.json").read_text())
    pa = json.loads((REF / "experiments/session_a/protocol.json").read_text())
    out = {"session_a": {"solver": pa["solver"], "material": pa["material"], "methods": _agg(a["to_tolerance"], "method")},
           "session_d": {"solver": pd["solver"], "material": pd["material"], "order_seed": pd["order_seed"],
                         "cases": {c["case"]: {"shapes": c["shapes"], "omegas": c["omegas"], "lambdas": c["lambdas"]}
                                   for c in d["cases"] if c["case"] != "smoke-q4"},
                         "arms": _agg(d["to_tolerance"], "arm")}}
    (HERE / "sessions_ad_known_answers.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
