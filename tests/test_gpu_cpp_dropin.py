"""The header-only C++ drop-in (include/sparsesolv_amd/sparsesolv.hpp) used
like reference client code: tests/cpp/dropin_test (built by build()) runs
solve_iccg, CGSolver / ICPreconditioner through the reference's base classes
(IterativeSolver, Preconditioner), apply_fused_dot, the SparseMatrixView
accessors and product, solve_cocr, COCRSolver with a preconditioner set up
on another matrix and on a matrix whose values then changed in place.  It
writes the matrices, right-hand sides and solutions it used; each is checked
here against the oracle (the reference's algorithm: iterations within 2%,
solutions within 1e-9, the preconditioner apply within 1e-12)."""
import json
import math
import os
import subprocess

import numpy as np
import pytest

from sparsesolv_amd.problems import CSR

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "dropin_test")


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def it_tol(n):
    return max(1, math.ceil(0.02 * n))


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/dropin_test is not built (run __graft_entry__.build())")
    d = tmp_path_factory.mktemp("dropin")
    out = subprocess.run([BIN, str(d)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    rows = {r["case"]: r for r in (json.loads(l) for l in out.stdout.splitlines() if l.startswith("{"))}

    def load(name, dt):
        return np.fromfile(os.path.join(d, name + ".bin"), dtype=dt)
    return rows, load


def mat(load, prefix, dt, values=None):
    rp = load(prefix + "_rp", np.int32)
    ci = load(prefix + "_ci", np.int32)
    v = load(values or prefix + "_v", dt)
    return CSR(len(rp) - 1, rp, ci, v)


def test_iccg_and_base_classes(run, oracle):
    rows, load = run
    A = mat(load, "real", np.float64)
    b = load("real_b", np.float64)
    it = rows["iccg"]
    xo, ro = oracle.solve_iccg(A, b, oracle.config(use_abmc=1, save_residual_history=1))
    assert it["converged"] and ro["converged"]
    assert abs(it["iterations"] - ro["iterations"]) <= it_tol(ro["iterations"])
    assert rel(load("iccg_x", np.float64), xo) < 1e-9
    hist = load("iccg_hist", np.float64)
    assert len(hist) == it["history"] == it["iterations"] + 1
    np.testing.assert_allclose(hist[:3], ro["residual_history"][:3], rtol=1e-9)
    cg = rows["cg_base_classes"]
    assert cg["solver"] == "CG" and cg["precond"] == "IC" and cg["ready"] == 1
    assert cg["converged"] and cg["iterations"] == it["iterations"] and cg["max_diff"] == 0.0


def test_apply_fused_dot_matches_oracle_apply(run, oracle):
    rows, load = run
    A = mat(load, "real", np.float64)
    b = load("real_b", np.float64)
    yo = oracle.ic(A, oracle.config(use_abmc=1)).apply(b)
    y = load("fused_y", np.float64)
    assert rel(y, yo) < 1e-12
    assert abs(rows["fused_dot"]["dot"] - float(b @ yo)) <= 1e-12 * abs(float(b @ yo))


def test_errors_and_host_preconditioner(run):
    rows, _ = run
    assert rows["apply_before_setup"]["threw"] == 1
    assert rows["apply_before_setup"]["solve_threw"] == 1
    h = rows["host_preconditioner"]
    assert h["refused"] == 1 and h["dot_rel_err"] < 1e-14


def test_view_accessors_and_product(run):
    rows, load = run
    v = rows["view"]
    assert v["bad"] == 0 and v["multiply_max_diff"] == 0.0
    assert v["nnz"] == len(load("real_ci", np.int32))


def test_cocr_against_oracle(run, oracle):
    rows, load = run
    A = mat(load, "cplx", np.complex128)
    b = load("cplx_b", np.complex128)
    cfg = oracle.config(use_abmc=1, abmc_num_colors=8)
    xo, ro = oracle.solve_cocr(A, b, cfg)
    r = rows["cocr"]
    assert r["converged"] and ro["converged"]
    assert abs(r["iterations"] - ro["iterations"]) <= it_tol(ro["iterations"])
    assert rel(load("cocr_x", np.complex128), xo) < 1e-9


def test_cocr_with_another_matrix_preconditioner(run, oracle):
    # COCRSolver::solve(A, b, x, n, &pc) with pc.setup(M): the oracle's
    # solve_cocr_pre runs the reference's COCR with that pairing
    rows, load = run
    A = mat(load, "cplx", np.complex128)
    M = mat(load, "cplx", np.complex128, values="m_v")
    b = load("cplx_b", np.complex128)
    cfg = oracle.config(use_abmc=1, abmc_num_colors=8)
    xo, ro = oracle.solve_cocr_pre(A, M, b, cfg)
    r = rows["cocr_other_pre"]
    assert r["converged"] and ro["converged"]
    assert abs(r["iterations"] - ro["iterations"]) <= it_tol(ro["iterations"])
    assert rel(load("cocr_other_pre_x", np.complex128), xo) < 1e-9
    assert r["own_pre_iterations"] == rows["cocr"]["iterations"]
    assert r["iterations"] != r["own_pre_iterations"]


def test_cocr_after_values_changed_in_place(run, oracle):
    # the preconditioner was set up on A's old values; the reference's solver
    # multiplies by the live A: COCR on the new A with the old factor
    rows, load = run
    A_old = mat(load, "cplx", np.complex128)
    A_new = mat(load, "cplx", np.complex128, values="cplx_v_new")
    b = load("cplx_b", np.complex128)
    cfg = oracle.config(use_abmc=1, abmc_num_colors=8)
    xo, ro = oracle.solve_cocr_pre(A_new, A_old, b, cfg)
    r = rows["cocr_changed_values"]
    assert r["converged"] and ro["converged"] and r["true_residual"] < 1e-9
    assert abs(r["iterations"] - ro["iterations"]) <= it_tol(ro["iterations"])
    assert rel(load("cocr_changed_x", np.complex128), xo) < 1e-9


def test_sgs_preconditioner_and_sgsmrtr(run, oracle):
    # SGSPreconditioner through the base class (apply, CGSolver with it) and
    # solve_sgsmrtr / SGSMRTRSolver (which ignores a preconditioner)
    rows, load = run
    A = mat(load, "real", np.float64)
    b = load("real_b", np.float64)
    assert rel(load("sgs_y", np.float64), oracle.sgs_apply(A, b)) < 1e-12
    r = rows["sgs_cg"]
    xo, ro = oracle.solve_sgs(A, b, oracle.config(), "CG")
    assert r["precond"] == "SGS" and r["converged"] and ro["converged"]
    assert abs(r["iterations"] - ro["iterations"]) <= it_tol(ro["iterations"])
    assert rel(load("sgs_cg_x", np.float64), xo) < 1e-9
    m = rows["sgsmrtr"]
    xo, ro = oracle.solve_sgsmrtr(A, b, oracle.config(save_residual_history=1))
    assert m["solver"] == "SGS-MRTR" and m["converged"] and ro["converged"]
    assert abs(m["iterations"] - ro["iterations"]) <= it_tol(ro["iterations"])
    assert rel(load("sgsmrtr_x", np.float64), xo) < 1e-9
    hist = load("sgsmrtr_hist", np.float64)
    assert len(hist) == m["iterations"] + 1
    np.testing.assert_allclose(hist[:5], ro["residual_history"][:5], rtol=1e-9)
    assert m["with_precond_iterations"] == m["iterations"] and m["with_precond_max_diff"] == 0.0


def test_complex_sgs_cocr_and_sgsmrtr(run, oracle):
    # on the complex operator after its in-place change; its diagonal is
    # complex, where the reference's SGS-MRTR split formula is not exact and
    # its iterate is followed all the same
    rows, load = run
    A = mat(load, "cplx", np.complex128, values="cplx_v_new")
    b = load("cplx_b", np.complex128)
    cfg = oracle.config(use_abmc=1, abmc_num_colors=8)
    r = rows["cplx_sgs"]
    xo, ro = oracle.solve_sgs(A, b, cfg, "COCR")
    assert r["cocr_converged"] and ro["converged"] and r["cocr_true_residual"] < 1e-9
    assert abs(r["cocr_iterations"] - ro["iterations"]) <= it_tol(ro["iterations"])
    assert rel(load("cplx_sgs_cocr_x", np.complex128), xo) < 1e-9
    xo, ro = oracle.solve_sgsmrtr(A, b, cfg)
    assert bool(r["mrtr_converged"]) == bool(ro["converged"])
    assert abs(r["mrtr_iterations"] - ro["iterations"]) <= it_tol(ro["iterations"])
    assert rel(load("cplx_sgsmrtr_x", np.complex128), xo) < 1e-9
