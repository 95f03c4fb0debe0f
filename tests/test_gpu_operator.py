"""ssa_ic_set_operator: solves on one matrix preconditioned by the IC
factor of another (COCRSolverNGS(mat, pre), ngsolve/sparsesolv_solvers.hpp:
256-396; IterativeSolver::solve(A, b, x, n, precond),
solvers/iterative_solver.hpp:58-80) against the oracle's COCR with the
preconditioner set up on the other matrix (pinned against the reference
build in tests/test_oracle_cocr_pre.py)."""
import math

import numpy as np
import pytest
import scipy.sparse as sp

from sparsesolv_amd import _lib
from sparsesolv_amd import problems as P
from test_gpu_parity import complex_shifted_laplace

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def shifted(A, s):
    return P.CSR.from_scipy((A.to_scipy() + sp.identity(A.n) * s).tocsr())


@pytest.mark.parametrize("use_abmc", [1, 0])
def test_cocr_on_another_operator(oracle, use_abmc):
    A, b = complex_shifted_laplace(12)
    M = shifted(A, 0.5)
    kw = dict(use_abmc=use_abmc, abmc_num_colors=8)
    cfg = _lib.default_config(**kw)
    h = _lib.IcHandle(M, cfg)
    x0 = P.uniform(9, A.n, -1.0, 1.0) + 0j
    y_pre = h.apply(x0)
    h.set_operator(A)
    assert h.info()["nnz"] == int(A.row_ptr[-1])
    # the SpMV is the new operator's, the preconditioner is unchanged
    assert rel(h.spmv(x0), A.to_scipy() @ x0) <= 1e-14
    assert np.array_equal(h.apply(x0), y_pre)
    x, res, _ = h.solve(_lib.METHOD_ICCOCR, b, None, cfg)
    xo, ro = oracle.solve_cocr_pre(A, M, b, oracle.config(**kw))
    assert res.converged and ro["converged"]
    assert abs(res.iterations - ro["iterations"]) <= max(1, math.ceil(0.02 * ro["iterations"]))
    assert rel(x, xo) < 1e-9
    assert np.linalg.norm(A.to_scipy() @ x - b) <= 1e-9 * np.linalg.norm(b)
    # back to the handle's own matrix: the same solve as a fresh handle's
    h.set_operator(M)
    x1, r1, _ = h.solve(_lib.METHOD_ICCOCR, b, None, cfg)
    x2, r2, _ = _lib.IcHandle(M, cfg).solve(_lib.METHOD_ICCOCR, b, None, cfg)
    assert r1.iterations == r2.iterations and np.array_equal(x1, x2)


def test_cocr_solver_front_end_with_another_matrix(oracle):
    import sparsesolv_amd as ss
    A, b = complex_shifted_laplace(10)
    M = shifted(A, 0.3)
    pre = ss.ICPreconditioner(M)
    pre.use_abmc = True
    pre.abmc_num_colors = 8
    solver = ss.COCRSolver(A, pre, maxiter=500, tol=1e-10)
    x = solver * b
    xo, ro = oracle.solve_cocr_pre(A, M, b, oracle.config(use_abmc=1, abmc_num_colors=8,
                                                          tolerance=1e-10, max_iterations=500,
                                                          save_best_result=0))
    assert ro["converged"]
    assert abs(solver.GetIterations() - ro["iterations"]) <= max(1, math.ceil(0.02 * ro["iterations"]))
    assert rel(x, xo) < 1e-9


def test_set_operator_errors():
    A = P.laplace7(6)
    h = _lib.IcHandle(A, _lib.default_config(use_abmc=1))
    with pytest.raises(_lib.SsaError) as e:
        h.set_operator(P.laplace7(5))
    assert e.value.code == _lib.SSA_ERR_ARGUMENT
    with pytest.raises(TypeError):
        h.set_operator(complex_shifted_laplace(6)[0])
    hb = _lib.BatchHandle([P.laplace7(4), P.laplace7(5)], _lib.default_config(use_abmc=1))
    with pytest.raises(_lib.SsaError) as e:
        _lib.IcHandle.set_operator(hb, P.laplace7(4))
    assert e.value.code == _lib.SSA_ERR_UNSUPPORTED


def test_cocr_solver_sees_values_changed_in_place(oracle):
    # the reference's COCRSolverNGS multiplies by the live matrix and applies
    # its preconditioner as set up: after an in-place change of A's values
    # the solve is COCR on the new A with the old factor
    import sparsesolv_amd as ss
    A, b = complex_shifted_laplace(9)
    A_old = P.CSR(A.n, A.row_ptr.copy(), A.col_idx.copy(), A.values.copy())
    pre = ss.ICPreconditioner(A)
    pre.use_abmc = True
    pre.abmc_num_colors = 8
    solver = ss.COCRSolver(A, pre, maxiter=500, tol=1e-10)
    ocfg = oracle.config(use_abmc=1, abmc_num_colors=8, tolerance=1e-10, max_iterations=500,
                         save_best_result=0)
    x0 = solver * b
    xo, _ = oracle.solve_cocr_pre(A, A, b, ocfg)
    assert rel(x0, xo) < 1e-9
    diag = A.col_idx == np.repeat(np.arange(A.n), np.diff(A.row_ptr))
    A.values[diag] += 0.4 + 0.3j
    x1 = solver * b
    xo1, ro1 = oracle.solve_cocr_pre(A, A_old, b, ocfg)
    assert ro1["converged"]
    assert abs(solver.GetIterations() - ro1["iterations"]) <= max(1, math.ceil(0.02 * ro1["iterations"]))
    assert rel(x1, xo1) < 1e-9
    # a second solver on another matrix shares pre: each solve uses its own
    M2 = shifted(A_old, 0.5)
    s2 = ss.COCRSolver(M2, pre, maxiter=500, tol=1e-10)
    x2 = s2 * b
    assert rel(x2, oracle.solve_cocr_pre(M2, A_old, b, ocfg)[0]) < 1e-9
    assert rel(solver * b, xo1) < 1e-9


def test_cocr_solver_projects_with_its_own_freedofs():
    import sparsesolv_amd as ss
    A, b = complex_shifted_laplace(7)
    free = np.ones(A.n, dtype=bool)
    free[::5] = False
    pre = ss.ICPreconditioner(A, freedofs=free)
    pre.use_abmc = True
    with pytest.raises(ValueError):
        ss.COCRSolver(A, pre) * b          # the preconditioner's freedofs only
    x = ss.COCRSolver(A, pre, freedofs=free, tol=1e-10) * b
    assert np.all(x[~free] == 0)
    r = np.where(free, b - A.to_scipy() @ x, 0)
    assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(np.where(free, b, 0))
