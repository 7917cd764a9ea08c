"""GPU checks of the study harness: the reference's scale-study solves (experiments/session_i, W2b
all-FP64 multigrid arms) reproduced through the drop-in voxel_mg on the reference's own base designs,
against the numbers its result.json records hold for three GPUs (tests/golden/session_i_known_answers.json),
and the product gates and paired timing on a small case."""
import json

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def known(golden):
    return json.loads((golden / "session_i_known_answers.json").read_text())


def _within(value, lo, hi, rel):
    return lo - rel * abs(lo) <= value <= hi + rel * abs(hi)


@pytest.mark.parametrize("case", ["q96x1", "q64x2", "q96x2", "q96x3"])
def test_study_case_reproduces_the_reference(golden, known, case):
    from voxelhip import harness
    ref = known["cases"][case]
    data = harness.load_case(golden, case)
    print("case", case, [int(s) for s in data["shape"]], flush=True)
    rec, x = harness.solve_case(data, case)
    harness.check_record(rec)
    assert [list(s) for s in __import__("voxel_mg").hierarchy_shapes(tuple(rec["shape"]))] == ref["shapes"]
    assert np.allclose(rec["lambdas"], ref["lambdas"], rtol=1e-9, atol=0), (rec["lambdas"], ref["lambdas"])
    assert np.allclose(rec["omegas"], ref["omegas"], rtol=1e-9, atol=0)
    assert rec["iterations"] in ref["iterations"], (rec["iterations"], ref["iterations"])
    lo, hi = ref["native_true_residual"]
    assert rec["native_true_residual"] <= 1e-8 and _within(rec["native_true_residual"], lo, hi, 1e-3)
    assert _within(rec["compliance"], *ref["compliance"], 1e-9), (rec["compliance"], ref["compliance"])
    assert _within(rec["force_work"], *ref["force_work"], 1e-9), (rec["force_work"], ref["force_work"])
    assert _within(rec["solution_norm"], *ref["solution_norm"], 1e-9), (rec["solution_norm"], ref["solution_norm"])
    assert rec["physical_passed"] and rec["independent_true_residual"] <= 1e-8


def test_product_gates_and_paired_timing(golden):
    import voxel_kernels as vk
    from voxelhip import harness, ops
    data = harness.load_case(golden, "q4x2")
    shape = tuple(int(s) for s in data["shape"])
    ke = data["ke"]
    emin = harness.PROTOCOL["material"]["Emin"]
    E = torch.from_numpy(emin + (1 - emin) * data["rho"] ** 3).cuda()
    mask = torch.from_numpy(data["mask"]).cuda()
    m = ops.Mask.from_values(mask, shape)
    blocks = vk.Modal(ke).blocks
    u = torch.from_numpy(np.random.default_rng(3).standard_normal(mask.numel())).cuda()

    def parity(v):
        return ops.matvec(blocks, E, v, torch.empty_like(v), shape, mask=m)

    dense = harness.constrained_dense(ke, E, shape, mask)
    chk = harness.product_check(parity(u), dense(u))
    assert chk["passed"] and chk["relative_l2"] < 1e-13
    ex = harness.exactness(parity, u, mask)
    assert ex["zero_exact"] and ex["fixed_identity_exact"]
    y = torch.empty_like(u)
    rows = harness.paired_rounds({"parity": lambda: ops.matvec(blocks, E, u, y, shape, mask=m),
                                  "dense": lambda: dense(u)}, rounds=3, calls=2, case="q4x2")
    assert len(rows) == 6 and all(r["gpu_ms_per_call"] > 0 for r in rows)
    assert sorted({r["round"] for r in rows}) == [0, 1, 2]
    dec = harness.timing_decision(rows, "parity", resamples=500)
    assert len(dec["comparisons"]) == 1 and np.isfinite(dec["comparisons"][0]["paired_median_ratio"])
