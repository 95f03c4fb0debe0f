"""The header-only C++ drop-in (include/sparsesolv_amd/sparsesolv.hpp) used
like reference client code: tests/cpp/dropin_test (built by build()) runs
solve_iccg, an explicit CGSolver + ICPreconditioner, and solve_cocr on the
GPU and reports its results."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "dropin_test")


def test_cpp_dropin_client():
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/dropin_test is not built (run __graft_entry__.build())")
    out = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rows = {r["case"]: r for r in (json.loads(l) for l in out.stdout.splitlines() if l.startswith("{"))}
    it = rows["iccg"]
    assert it["converged"] and it["true_residual"] < 1e-9 and it["history"] == it["iterations"] + 1
    assert rows["cg_object"]["converged"] and rows["cg_object"]["max_diff"] == 0.0
    assert rows["cg_object"]["iterations"] == it["iterations"]
    assert rows["apply_before_setup"]["threw"] == 1
    assert rows["cocr"]["converged"] and rows["cocr"]["true_residual"] < 1e-9
    # a preconditioner set up on another matrix is used as given: A's solve
    # converges with M's factor, in another number of iterations
    o = rows["cocr_other_pre"]
    assert o["converged"] and o["true_residual"] < 1e-9
    assert o["own_pre_iterations"] == rows["cocr"]["iterations"] and o["max_diff"] > 0.0
