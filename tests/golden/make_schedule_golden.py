"""Writes tests/golden/schedule_golden.json by importing the donor's SIMP driver module from the
reference checkout at /root/reference in this container (its module level needs only the standard
library; the GPU functions import CuPy inside and are not called).

    python tests/golden/make_schedule_golden.py

It records schedule_at(it) (experiments/session_a/donor/src/gpu_fem/simp_r2.py:420-433: the
continuation state (penal, beta, move, rmin, in_last_phase) at outer iteration it) for
it = 1 .. 130 with the default table and with a two-phase custom table.  Nothing else pins the
continuation numbers, so they are taken from the reference itself.  Data only.
"""
from __future__ import annotations

import importlib.util
import json
from pathlib import Path

SRC = Path("/root/reference/experiments/session_a/donor/src/gpu_fem/simp_r2.py")
HERE = Path(__file__).resolve().parent
CUSTOM = [(5, 2.0, 2.0, 0.3, 2.0), (12, 3.0, 8.0, 0.1, 1.1)]


def main():
    spec = importlib.util.spec_from_file_location("donor_simp_r2", SRC)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    out = {"default": [list(mod.schedule_at(it)) for it in range(1, 131)],
           "custom_table": [list(r) for r in CUSTOM],
           "custom": [list(mod.schedule_at(it, CUSTOM)) for it in range(1, 21)]}
    (HERE / "schedule_golden.json").write_text(json.dumps(out) + "\n")


if __name__ == "__main__":
    main()
