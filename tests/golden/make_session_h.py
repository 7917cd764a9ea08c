"""Writes tests/golden/session_h_known_answers.json and copies the final designs of the
cantilever FP64 runs from the reference checkout at /root/reference (only where that checkout
exists; the tests read the committed fixtures).

    python tests/golden/make_session_h.py

Source: the reference's session-H SIMP study (experiments/session_h/protocol.json: the donor's
cantilever_gpu_large / cantilever_gpu_xlarge presets (donor_yang_gmg/src/gpu_fem/presets.py:
146-161), 120x60x30 / 160x80x40 elements on 2.0 x 1.0 x 0.5, x = 0 face clamped, unit point load
-y at (2.0, 0.5, 0.25); density filter r = 1.5, top88 OC with move 0.15, penal 3, E0 = 1,
Emin = 1e-9, 100 fixed steps from x = volfrac = 0.3) and the result.json records of its two hosts
(results/session-h-host{1,2}-20260926/remote/run-01/result.json).  For the all-FP64 schedule it
keeps, per case, arm and repetition, the per-step compliance, filtered volume, largest design
change and the solver's iteration count, plus the final physical design of host 1's stock arm
(designs/<case>-fp64-stock.npz).  Data only: numbers and the reference's own output files, no
code.
"""
from __future__ import annotations

import json
import shutil
from pathlib import Path

REF = Path("/root/reference")
HERE = Path(__file__).resolve().parent
RUNS = {h: REF / f"results/session-h-{h}-20260926/remote/run-01" for h in ("host1", "host2")}
CASES = {"cantilever-216k": ("cantilever_gpu_large", [120, 60, 30]),
         "cantilever-512k": ("cantilever_gpu_xlarge", [160, 80, 40])}


def main():
    proto = json.loads((REF / "experiments/session_h/protocol.json").read_text())
    out = {"simp": proto["simp"], "solver_tolerance": 1e-5, "volfrac": 0.3, "cases": {}}
    for case, (preset, shape) in CASES.items():
        row = {"preset": preset, "shape": shape, "lengths": [2.0, 1.0, 0.5], "point": [2.0, 0.5, 0.25],
               "load": [0.0, -1.0, 0.0], "runs": []}
        for host, d in RUNS.items():
            r = json.loads((d / "result.json").read_text())
            for run in r["runs"]:
                if run["case"] != case or run["schedule"] != "fp64":
                    continue
                h = run["history"]
                row["runs"].append({"host": host, "arm": run["arm"], "repetition": run.get("repetition"),
                                    "compliance": [x["compliance"] for x in h], "volume": [x["volume"] for x in h],
                                    "change": [x["change"] for x in h],
                                    "outer_iterations": [x["outer_iterations"] for x in h],
                                    "final_compliance": run["final_compliance"]})
        out["cases"][case] = row
        shutil.copyfile(RUNS["host1"] / "designs" / f"{case}-fp64-stock.npz", HERE / f"session_h-{case}-fp64-stock.npz")
        print(case, len(row["runs"]), "runs")
    (HERE / "session_h_known_answers.json").write_text(json.dumps(out) + "\n")


if __name__ == "__main__":
    main()
