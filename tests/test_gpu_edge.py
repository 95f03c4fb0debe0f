"""Degenerate shapes on the GPU: one-row systems, diagonal matrices (rows
without off-diagonal entries; on the ABMC path every row is its own
aggregate, so the stragglers form one oversized block walked in a single
level), systems smaller than a wavefront, and the 1-D tridiagonal matrix,
whose natural-order IC(0) with shift 1.0 is exact (M = A: the pivots are
d_k = 2 - 1/d_{k-1}, (k+1)/k, and L has no dropped fill), so PCG on the
level path converges in one iteration.  The layouts of these shapes are
checked on the host by tests/test_layout_cpp.py's checker as well."""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from sparsesolv_amd import _lib
from sparsesolv_amd import problems as P

pytestmark = pytest.mark.gpu


def csr(M):
    return P.CSR.from_scipy(sp.csr_matrix(M))


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def tridiagonal(n):
    return csr(sp.diags([-np.ones(n - 1), 2.0 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1]))


@pytest.mark.parametrize("use_abmc", [1, 0])
def test_one_row_systems(use_abmc):
    cfg = _lib.default_config(use_abmc=use_abmc)
    x, res, _ = _lib.IcHandle(csr(np.array([[4.0]])), cfg).solve(_lib.METHOD_ICCG, np.array([2.0]),
                                                                  None, cfg)
    assert res.converged and res.iterations == 1 and abs(x[0] - 0.5) <= 1e-15
    a = 4.0 + 1.0j
    x, res, _ = _lib.IcHandle(csr(np.array([[a]])), cfg).solve(_lib.METHOD_ICCOCR,
                                                               np.array([2.0 + 0.0j]), None, cfg)
    assert res.converged and res.iterations == 1
    assert abs(x[0] - 2.0 / a) <= 1e-15 * abs(2.0 / a)


@pytest.mark.parametrize("use_abmc", [1, 0])
def test_diagonal_matrix(oracle, use_abmc):
    d = np.linspace(1.0, 5.0, 100)
    A = csr(sp.diags(d))
    b = P.uniform(4, 100, -1.0, 1.0)
    cfg = _lib.default_config(use_abmc=use_abmc)
    h = _lib.IcHandle(A, cfg)
    if use_abmc:
        assert h.info()["big_blocks"] == 1     # 75 stragglers + the last seed
    x, res, _ = h.solve(_lib.METHOD_ICCG, b, None, cfg)
    assert res.converged and res.iterations == 1
    assert np.max(np.abs(x - b / d)) <= 1e-14
    # the apply is (1.05 D)^-1
    y = h.apply(b)
    yo = oracle.ic(A, oracle.config(use_abmc=use_abmc)).apply(b)
    assert rel(y, yo) <= 1e-15 and rel(y, b / (1.05 * d)) <= 1e-15


def test_tridiagonal_level_path_shift_one_is_exact():
    n = 300
    A = tridiagonal(n)
    b = P.uniform(3, n, -1.0, 1.0)
    cfg = _lib.default_config(use_abmc=0, shift_parameter=1.0)
    h = _lib.IcHandle(A, cfg)
    x, res, _ = h.solve(_lib.METHOD_ICCG, b, None, cfg)
    assert res.converged and res.iterations == 1
    # cond(A) ~ 3.7e4: the exact preconditioner leaves only rounding
    assert rel(x, spla.spsolve(A.to_scipy().tocsc(), b)) <= 1e-11
    # M = A: the apply is the solve
    assert rel(A.to_scipy() @ h.apply(b), b) <= 1e-11


@pytest.mark.parametrize("shift", [1.0, 1.05])
@pytest.mark.parametrize("use_abmc", [1, 0])
def test_tridiagonal_matches_oracle(oracle, use_abmc, shift):
    n = 300
    A = tridiagonal(n)
    b = P.uniform(3, n, -1.0, 1.0)
    kw = dict(use_abmc=use_abmc, shift_parameter=shift)
    cfg = _lib.default_config(**kw)
    x, res, _ = _lib.IcHandle(A, cfg).solve(_lib.METHOD_ICCG, b, None, cfg)
    xo, ro = oracle.solve_iccg(A, b, oracle.config(**kw))
    assert res.converged and ro["converged"]
    assert abs(res.iterations - ro["iterations"]) <= max(1, math.ceil(0.02 * ro["iterations"]))
    assert rel(x, xo) < 1e-9


@pytest.mark.parametrize("use_abmc", [1, 0])
def test_system_smaller_than_a_wavefront(oracle, use_abmc):
    M = np.array([[4, -1, 0, 0, -1], [-1, 4, -1, 0, 0], [0, -1, 4, -1, 0], [0, 0, -1, 4, -1],
                  [-1, 0, 0, -1, 4]], dtype=float)
    A = csr(M)
    b = np.array([1.0, -2.0, 0.5, 3.0, -1.0])
    cfg = _lib.default_config(use_abmc=use_abmc)
    h = _lib.IcHandle(A, cfg)
    x, res, _ = h.solve(_lib.METHOD_ICCG, b, None, cfg)
    xo, ro = oracle.solve_iccg(A, b, oracle.config(use_abmc=use_abmc))
    assert res.converged and abs(res.iterations - ro["iterations"]) <= 1
    assert rel(x, np.linalg.solve(M, b)) < 1e-9 and rel(x, xo) < 1e-9
    assert rel(h.spmv(b), M @ b) <= 1e-15


@pytest.mark.parametrize("shape", ["diag100", "tri300", "n5"])
def test_blocks_of_right_hand_sides_on_degenerate_shapes(shape):
    if shape == "diag100":
        A = csr(sp.diags(np.linspace(1.0, 5.0, 100)))
    elif shape == "tri300":
        A = tridiagonal(300)
    else:
        A = csr(np.eye(5) * 4.0 - np.eye(5, k=1) - np.eye(5, k=-1))
    B = np.stack([P.uniform(20 + k, A.n, -1.0, 1.0) for k in range(5)])
    cfg = _lib.default_config(use_abmc=1, max_iterations=600)
    h = _lib.IcHandle(A, cfg)
    X, results = h.solve_multi(_lib.METHOD_ICCG, B, None, cfg)
    for k in range(B.shape[0]):
        xk, rk, _ = h.solve(_lib.METHOD_ICCG, B[k], None, cfg)
        assert np.array_equal(X[k], xk) and results[k].iterations == rk.iterations
