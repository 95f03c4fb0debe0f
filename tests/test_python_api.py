"""Python drop-in surface: factory names, keyword defaults, properties,
typed real/complex classes, error behaviour (tests/test_sparsesolv.py of the
reference: :243-295, :718-750).  No GPU work here."""
import numpy as np
import pytest
import scipy.sparse as sp

import sparsesolv_amd as ss
from sparsesolv_amd import _lib
from sparsesolv_amd import problems as P
from sparsesolv_amd.solvers import _with_freedofs


def test_properties_and_defaults():
    A = P.laplace7(3)
    s = ss.SparseSolvSolver(A, method="ICCG")
    assert isinstance(s, ss.SparseSolvSolverD)
    assert s.method == "ICCG" and abs(s.tol - 1e-10) < 1e-15 and s.maxiter == 1000
    assert abs(s.shift - 1.05) < 1e-10 and s.save_best_result is True
    assert s.save_residual_history is False and s.auto_shift is False
    assert s.diagonal_scaling is False and s.divergence_check is False
    assert abs(s.divergence_threshold - 1000.0) < 1e-10 and s.divergence_count == 100
    assert s.use_abmc is False and s.abmc_block_size == 4 and s.abmc_num_colors == 4
    assert s.abmc_reorder_spmv is False and s.abmc_use_rcm is False and s.conjugate is False
    s.method = "COCR"; s.tol = 1e-8; s.maxiter = 500; s.auto_shift = True
    s.use_abmc = True; s.abmc_block_size = 8; s.abmc_num_colors = 6
    assert (s.method, s.tol, s.maxiter, s.auto_shift) == ("COCR", 1e-8, 500, True)
    assert (s.use_abmc, s.abmc_block_size, s.abmc_num_colors) == (True, 8, 6)


def test_factory_parameters():
    A = P.laplace7(3)
    s = ss.SparseSolvSolver(A, method="ICCG", use_abmc=True, abmc_block_size=8,
                            abmc_num_colors=6, conjugate=True)
    assert s.use_abmc and s.abmc_block_size == 8 and s.abmc_num_colors == 6 and s.conjugate


def test_complex_dispatch():
    E = P.eddy(3).A
    assert isinstance(ss.SparseSolvSolver(E), ss.SparseSolvSolverC)
    assert isinstance(ss.ICPreconditioner(E), ss.ICPreconditionerC)
    pre = ss.ICPreconditioner(E)
    assert isinstance(ss.COCRSolver(E, pre), ss.COCRSolverC)
    assert ss.COCRSolver(E, pre).iterations == 0


def test_ic_preconditioner_properties():
    pre = ss.ICPreconditioner(P.laplace7(3), shift=1.2)
    assert pre.shift == 1.2 and pre.use_abmc is False
    assert pre.abmc_block_size == 4 and pre.abmc_num_colors == 4 and pre.diagonal_scaling is False
    pre.use_abmc = True
    pre.abmc_num_colors = 8
    assert pre.use_abmc and pre.abmc_num_colors == 8


def test_invalid_method_raises_runtime_error():
    A = P.laplace7(3)
    s = ss.SparseSolvSolver(A, method="INVALID")
    with pytest.raises(RuntimeError, match="Unknown method"):
        s.Solve(np.ones(A.n), np.zeros(A.n))


def test_non_sparse_matrix_is_type_error():
    with pytest.raises(TypeError):
        ss.SparseSolvSolver(np.eye(3))


def test_result_repr():
    r = ss.SparseSolvResult(True, 12, 3.5e-11, [1.0, 0.1])
    assert repr(r) == "SparseSolvResult(converged=True, iterations=12, residual=0.000000)"


def test_freedofs_identity_rows():
    # BuildSparseMatrixView: constrained rows -> identity, constrained columns
    # of free rows -> 0, pattern kept (ngsolve/sparsesolv_precond.hpp:60-74)
    A = P.laplace7(3)
    free = np.ones(A.n, dtype=bool)
    free[[0, 13, 26]] = False
    M = _with_freedofs(A, free)
    assert np.array_equal(M.row_ptr, A.row_ptr) and np.array_equal(M.col_idx, A.col_idx)
    D = M.to_scipy().toarray()
    Ad = A.to_scipy().toarray()
    for i in range(A.n):
        for j in range(A.n):
            if Ad[i, j] == 0:
                continue
            if not free[i]:
                assert D[i, j] == (1.0 if i == j else 0.0)
            elif not free[j]:
                assert D[i, j] == 0.0
            else:
                assert D[i, j] == Ad[i, j]


def test_mixed_precision_switch_and_host_setup():
    s = P.poisson(8)
    solver = ss.SparseSolvSolver(s.A, method="ICCG")
    assert not solver.mixed_precision
    solver.mixed_precision = True
    assert solver.precision == _lib.PRECISION_MIXED
    # a host-only mixed handle runs the same FP64 setup as an FP64 handle
    f = []
    for prec in (_lib.PRECISION_FP64, _lib.PRECISION_MIXED):
        h = _lib.IcHandle(s.A, _lib.default_config(use_abmc=1, precision=prec), host_only=True)
        f.append(h.factor())
    for k in f[0]:
        assert np.array_equal(f[0][k], f[1][k]), k


def test_solve_block_checks_its_arrays():
    # no GPU needed: the shapes are checked before any handle is made
    A = P.laplace7(4)
    s = ss.SparseSolvSolver(A, method="ICCG")
    B = np.ones((3, A.n))
    with pytest.raises(ValueError):
        s.SolveBlock(B, np.zeros(A.n))              # sol must be (k, n)
    with pytest.raises(ValueError):
        s.SolveBlock(B, np.zeros((2, A.n)))         # one row per right-hand side
    with pytest.raises(ValueError):
        s.SolveBlock(np.ones((3, A.n + 1)), np.zeros((3, A.n + 1)))
