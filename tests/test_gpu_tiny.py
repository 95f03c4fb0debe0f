"""Tiny and ragged systems on the GPU against the oracle (itself pinned bit
for bit against the reference build on the same systems in
test_oracle_tiny.py): one and two rows, and row counts on either side of
the 64-row slice and the 256-row workgroup, through every entry point that
takes a system -- solves on the ABMC and level paths, mixed precision,
COCR, SGS-MRTR, the SpMV, blocks of right-hand sides and a batch of such
systems (each equal to its own solve bit for bit), and the preconditioner
with clamped pivots.  An empty system solves
as the oracle's does (converged after the one breakdown check, x empty);
the reference itself does not accept one."""
import math

import numpy as np
import pytest

from sparsesolv_amd import _lib
from sparsesolv_amd import problems as P
from tiny_systems import GRIDS, grid_laplacian, rhs

pytestmark = pytest.mark.gpu

CASES = sorted(GRIDS, key=lambda k: int(k[1:]))


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def it_tol(n):
    return max(1, math.ceil(0.02 * n))


def close_to_oracle(x, res, xo, ro, bar=1e-9):
    assert res.converged == ro["converged"]
    assert abs(res.iterations - ro["iterations"]) <= it_tol(ro["iterations"])
    assert rel(x, xo) < bar


@pytest.mark.parametrize("name", CASES)
def test_tiny_iccg_and_cocr_match_oracle(oracle, name):
    print(f"case {name}", flush=True)
    a, b_ = GRIDS[name]
    A = grid_laplacian(a, b_)
    b = rhs(A.n, 5)
    for kw in (dict(use_abmc=1), dict(use_abmc=0), dict(use_abmc=1, abmc_block_size=8, abmc_num_colors=6)):
        cfg = _lib.default_config(**kw)
        x, res, _ = _lib.IcHandle(A, cfg).solve(_lib.METHOD_ICCG, b, None, cfg)
        xo, ro = oracle.solve_iccg(A, b, oracle.config(**kw))
        close_to_oracle(x, res, xo, ro)
    # mixed precision: the FP32 factor inside FP64 CG with refinement
    cfg = _lib.default_config(use_abmc=1, precision=_lib.PRECISION_MIXED)
    x, res, _ = _lib.IcHandle(A, cfg).solve(_lib.METHOD_ICCG, b, None, cfg)
    xo, ro = oracle.solve_iccg(A, b, oracle.config(use_abmc=1))
    assert res.converged and rel(x, xo) < 1e-8
    # COCR on the complex-symmetric variant
    Z = grid_laplacian(a, b_, complex_=True)
    z = rhs(Z.n, 6, complex_=True)
    cfg = _lib.default_config(use_abmc=1, abmc_num_colors=8)
    x, res, _ = _lib.IcHandle(Z, cfg).solve(_lib.METHOD_ICCOCR, z, None, cfg)
    xo, ro = oracle.solve_cocr(Z, z, oracle.config(use_abmc=1, abmc_num_colors=8))
    close_to_oracle(x, res, xo, ro)
    # the operator
    h = _lib.IcHandle(Z, cfg)
    assert rel(h.spmv(z), Z.to_scipy() @ z) < 1e-13


@pytest.mark.parametrize("name", ["n1", "n2", "n65", "n257"])
def test_tiny_sgsmrtr_matches_oracle(oracle, name):
    print(f"case {name}", flush=True)
    a, b_ = GRIDS[name]
    A = grid_laplacian(a, b_)
    b = rhs(A.n, 5)
    cfg = _lib.default_config(preconditioner=_lib.PRECOND_SGSMRTR, save_residual_history=1)
    x, res, hist = _lib.IcHandle(A, cfg).solve(_lib.METHOD_SGSMRTR, b, None, cfg)
    xo, ro = oracle.solve_sgsmrtr(A, b, oracle.config(save_residual_history=1))
    close_to_oracle(x, res, xo, ro)


@pytest.mark.parametrize("name", ["n1", "n63", "n65", "n257"])
def test_tiny_blocks_match_single_solves(name):
    print(f"case {name}", flush=True)
    a, b_ = GRIDS[name]
    A = grid_laplacian(a, b_)
    B = np.stack([rhs(A.n, 20 + k) for k in range(5)])
    cfg = _lib.default_config(use_abmc=1)
    h = _lib.IcHandle(A, cfg)
    X, results = h.solve_multi(_lib.METHOD_ICCG, B, None, cfg)
    for k in range(B.shape[0]):
        xk, rk, _ = h.solve(_lib.METHOD_ICCG, B[k], None, cfg)
        assert np.array_equal(X[k], xk), k
        assert results[k].iterations == rk.iterations and results[k].converged == rk.converged, k


def test_tiny_batch_matches_single_solves(oracle):
    names = ["n1", "n2", "n63", "n64", "n65", "n257"]
    mats = [grid_laplacian(*GRIDS[k]) for k in names]
    bs = [rhs(A.n, 40 + k) for k, A in enumerate(mats)]
    cfg = _lib.default_config(use_abmc=1)
    hb = _lib.BatchHandle(mats, cfg)
    x, results = hb.solve_batch(_lib.METHOD_ICCG, np.concatenate(bs), None, cfg)
    for k, (A, bk) in enumerate(zip(mats, bs)):
        xk, rk, _ = _lib.IcHandle(A, cfg).solve(_lib.METHOD_ICCG, bk, None, cfg)
        assert np.array_equal(hb.split(x)[k], xk), names[k]
        assert results[k].iterations == rk.iterations, names[k]
        assert results[k].final_residual == rk.final_residual, names[k]
        xo, ro = oracle.solve_iccg(A, bk, oracle.config(use_abmc=1))
        close_to_oracle(xk, rk, xo, ro)


def test_empty_system_solves_as_the_oracle(oracle):
    import scipy.sparse as sp
    A = P.CSR.from_scipy(sp.csr_matrix((0, 0)))
    cfg = _lib.default_config(use_abmc=1)
    x, res, _ = _lib.IcHandle(A, cfg).solve(_lib.METHOD_ICCG, np.zeros(0), None, cfg)
    xo, ro = oracle.solve_iccg(A, np.zeros(0), oracle.config(use_abmc=1))
    assert x.shape == (0,)
    assert res.converged == ro["converged"] and res.iterations == ro["iterations"]


@pytest.mark.parametrize("use_abmc", [1, 0])
def test_clamped_pivots_apply_matches_oracle(oracle, use_abmc):
    # every second pivot of uncoupled [[1, 1], [1, 1]] blocks at shift 1.0
    # vanishes and is clamped to 1e-10 (inverse 1e10)
    import scipy.sparse as sp
    A = P.CSR.from_scipy(sp.kron(sp.identity(3000), np.array([[1.0, 1.0], [1.0, 1.0]])).tocsr())
    cfg = _lib.default_config(use_abmc=use_abmc, shift_parameter=1.0)
    x = rhs(A.n, 7)
    y = _lib.IcHandle(A, cfg).apply(x)
    want = oracle.ic(A, oracle.config(use_abmc=use_abmc, shift_parameter=1.0)).apply(x)
    assert np.max(np.abs(want)) > 1e9
    assert rel(y, want) < 1e-12
