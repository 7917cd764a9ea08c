"""CPU checks of the study harness (voxelhip.harness): input regeneration and upsampling against the
reference's own base designs (experiments/session_i/test_session_i.py:16-31 restated on the fixtures),
the paired bootstrap timing decision and the result records."""
import json
import math

import numpy as np
import pytest


def _harness():
    from voxelhip import harness
    return harness


@pytest.mark.parametrize("name", ["smoke-q4.npz", "optimized-q64.npz", "optimized-q96.npz"])
def test_boundary_data_is_the_fixtures_at_factor_one(golden, name):
    h = _harness()
    d = dict(np.load(golden / name))
    shape = [int(s) for s in d["shape"]]
    mask, force = h.cantilever_boundary(shape)
    assert np.array_equal(mask, h.to_author(d["mask"], shape, True))
    assert np.array_equal(force, h.to_author(d["force"], shape, True))
    assert np.array_equal(h.refine_design(d["rho"], shape, 1), h.to_author(d["rho"], shape))


@pytest.mark.parametrize("case", ["q4x2", "q64x2", "q96x3"])
def test_study_inputs_equal_the_oracle_regeneration_bitwise(golden, case):
    """load_case (author layout, built directly) against the oracle's natural-layout restatement of
    experiments/session_i/make_inputs.py:31-50 transposed to the author layout."""
    from oracle import simpref as sr
    h = _harness()
    base, factor = h.CASES[case]
    d = dict(np.load(golden / base))
    ref = sr.upsample(d, factor)
    got = h.load_case(golden, case)
    shape = [int(s) for s in ref["shape"]]
    assert list(got["shape"]) == shape
    assert np.array_equal(got["rho"], h.to_author(ref["rho"], shape))
    assert np.array_equal(got["mask"], h.to_author(ref["mask"], shape, True))
    assert np.array_equal(got["force"], h.to_author(ref["force"], shape, True))


def test_refinement_preserves_the_design(golden):
    h = _harness()
    d = dict(np.load(golden / "optimized-q64.npz"))
    up = h.load_case(golden, "q64x2")
    nx, ny, nz = (int(s) for s in d["shape"])
    r = up["rho"].reshape(2 * nx, 2 * ny, 2 * nz)
    assert np.array_equal(r[1::2, ::2, 1::2], h.to_author(d["rho"], (nx, ny, nz)).reshape(nx, ny, nz))
    assert abs(up["force"].sum() + 1) < 1e-12 and up["mask"].size == 3 * (2 * nx + 1) * (2 * ny + 1) * (2 * nz + 1)
    assert list(up["shape"]) == [2 * nx, 2 * ny, 2 * nz]


def test_to_author_matches_the_oracle_layout():
    from oracle import hexref as hx
    h = _harness()
    shape = (3, 2, 4)
    rng = np.random.default_rng(0)
    e, n = rng.standard_normal(24), rng.standard_normal(3 * 4 * 3 * 5)
    assert np.array_equal(h.to_author(e, shape), hx.to_author(e, shape))
    assert np.array_equal(h.to_author(n, shape, True), hx.to_author(n, shape, True))


def _rounds(ratio, rounds=7, noise=0.02, seed=1):
    rng = np.random.default_rng(seed)
    rows = []
    for r in range(rounds):
        c = 1.0 * (1 + noise * rng.standard_normal())
        rows.append({"round": r, "path": "cand", "gpu_ms_per_call": c})
        rows.append({"round": r, "path": "base", "gpu_ms_per_call": ratio * c * (1 + noise * rng.standard_normal())})
    return rows


def test_timing_decision_gates():
    h = _harness()
    fast = h.timing_decision(_rounds(1.4), "cand", resamples=2000)
    same = h.timing_decision(_rounds(1.0), "cand", resamples=2000)
    c = fast["comparisons"][0]
    assert fast["timing_gate_passed"] and c["passed"] and c["baseline"] == "base"
    assert c["paired_bootstrap_lower_95"] <= c["paired_median_ratio"]
    assert abs(c["paired_median_ratio"] - 1.4) < 0.1
    assert not same["timing_gate_passed"]
    # a non-primary comparison does not gate
    assert h.timing_decision(_rounds(1.0), "cand", primary=False, resamples=500)["timing_gate_passed"]
    bad = _rounds(1.2)[:-1]
    with pytest.raises(ValueError):
        h.timing_decision(bad, "cand", resamples=100)


def test_records_and_atomic_write(tmp_path):
    h = _harness()
    rec = {"case": "q4x2", "shape": [16, 8, 8], "ndof": 2187, "iterations": 3, "native_converged": True,
           "native_true_residual": 1e-9, "independent_true_residual": 1e-9, "compliance": 1.0,
           "force_work": 1.0, "energy_force_relative_error": 0.0, "solution_norm": 2.0,
           "solution_sha256": "0" * 64, "omegas": [0.4], "lambdas": [3.0], "setup_seconds": 0.1,
           "solver_seconds": 0.01, "physical_passed": True}
    assert h.check_record(rec)
    bad = dict(rec)
    del bad["compliance"]
    with pytest.raises(ValueError):
        h.check_record(bad)
    with pytest.raises(ValueError):
        h.check_record(dict(rec, iterations="3"))
    out = h.write_result(tmp_path / "result.json", {"solves": [rec]})
    assert json.loads(out.read_text())["solves"][0]["case"] == "q4x2"
    with pytest.raises(ValueError):
        h.write_result(tmp_path / "nan.json", {"x": math.nan})
    assert not (tmp_path / "nan.json").exists()
