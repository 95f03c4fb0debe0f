"""diagonal_scaling on the GPU against the oracle.

The factor is that of S A S, S = diag(1/sqrt|a_ii|) (bit-exact on the host:
tests/test_host_setup.py).  The reference applies S M^-1 S on the level path
and in ICCG's reordered-space apply (Path 1, abmc_reorder_spmv), but its
ABMC apply (the preconditioner's own, and ICCG Path 3 / COCR) runs M^-1
unscaled: the composite scaling vector that apply multiplies by is built
before the scaling factors (ic_preconditioner.hpp:125 vs :162).  IC(0)
commutes with diagonal scaling (the factor of S A S is S L), so the scaled
modes equal the unscaled preconditioner up to rounding, while the unscaled
ABMC apply of the scaled factor is another preconditioner: on a
variable-coefficient Laplacian it costs Path 3 ~600 iterations against 40.
Each mode is compared with the oracle's restatement, and with golden
vectors the reference itself wrote (tests/golden/*scaled*: Path 3, Path 1,
the level path, COCR)."""
import math

import numpy as np
import pytest

from sparsesolv_amd import _lib
from sparsesolv_amd import problems as P
from conftest import load_golden
from test_gpu_parity import complex_shifted_laplace

import make_golden as G

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def varcoef(m, seed=21):
    return P.laplace7_varcoef(m, 10.0 ** P.uniform(seed, m ** 3, -2.0, 2.0))


@pytest.mark.parametrize("use_abmc", [1, 0])
def test_apply_with_diagonal_scaling(oracle, use_abmc):
    A = varcoef(10)
    kw = dict(use_abmc=use_abmc, diagonal_scaling=1)
    h = _lib.IcHandle(A, _lib.default_config(**kw))
    x = P.uniform(5, A.n, -1.0, 1.0)
    y = h.apply(x)
    yo = oracle.ic(A, oracle.config(**kw)).apply(x)
    assert rel(y, yo) <= 1e-12
    y_plain = _lib.IcHandle(A, _lib.default_config(use_abmc=use_abmc)).apply(x)
    if use_abmc:
        # M_s^-1 applied without S is another preconditioner
        assert rel(y, y_plain) > 1e-3
    else:
        # IC(0) commutes with diagonal scaling (the factor of S A S is S L),
        # so the scaled apply S M_s^-1 S is M^-1 up to rounding
        assert rel(y, y_plain) <= 1e-12


@pytest.mark.parametrize("use_abmc,path1", [(1, 0), (1, 1), (0, 0)])
def test_iccg_with_diagonal_scaling(oracle, use_abmc, path1):
    A = varcoef(10)
    b = P.uniform(3, A.n, -1.0, 1.0)
    kw = dict(use_abmc=use_abmc, abmc_reorder_spmv=path1, diagonal_scaling=1, max_iterations=2000)
    cfg = _lib.default_config(**kw)
    x, res, _ = _lib.IcHandle(A, cfg).solve(_lib.METHOD_ICCG, b, None, cfg)
    xo, ro = oracle.solve_iccg(A, b, oracle.config(**kw))
    assert res.converged and ro["converged"]
    assert abs(res.iterations - ro["iterations"]) <= max(1, math.ceil(0.02 * ro["iterations"]))
    assert rel(x, xo) < 1e-9


def test_cocr_with_diagonal_scaling(oracle):
    A, b = complex_shifted_laplace(10)
    kw = dict(use_abmc=1, abmc_num_colors=8, diagonal_scaling=1, max_iterations=2000)
    cfg = _lib.default_config(**kw)
    x, res, _ = _lib.IcHandle(A, cfg).solve(_lib.METHOD_ICCOCR, b, None, cfg)
    xo, ro = oracle.solve_cocr(A, b, oracle.config(**kw))
    assert res.converged and ro["converged"]
    assert abs(res.iterations - ro["iterations"]) <= max(1, math.ceil(0.02 * ro["iterations"]))
    assert rel(x, xo) < 1e-9


def test_mixed_precision_with_diagonal_scaling(oracle):
    A = varcoef(10)
    b = P.uniform(3, A.n, -1.0, 1.0)
    kw = dict(use_abmc=0, diagonal_scaling=1)
    cfg = _lib.default_config(precision=_lib.PRECISION_MIXED, **kw)
    x, res, _ = _lib.IcHandle(A, cfg).solve(_lib.METHOD_ICCG, b, None, cfg)
    xo, ro = oracle.solve_iccg(A, b, oracle.config(**kw))
    assert res.converged and rel(x, xo) < 1e-9


@pytest.mark.parametrize("use_abmc,path1", [(1, 1), (0, 0)])
def test_block_solves_with_diagonal_scaling(use_abmc, path1):
    A = varcoef(10, seed=4)
    B = np.stack([P.uniform(30 + k, A.n, -1.0, 1.0) for k in range(5)])
    cfg = _lib.default_config(use_abmc=use_abmc, abmc_reorder_spmv=path1, diagonal_scaling=1)
    h = _lib.IcHandle(A, cfg)
    X, results = h.solve_multi(_lib.METHOD_ICCG, B, None, cfg)
    for k in range(B.shape[0]):
        xk, rk, _ = h.solve(_lib.METHOD_ICCG, B[k], None, cfg)
        assert np.array_equal(X[k], xk) and results[k].iterations == rk.iterations
    Y = h.apply_multi(B)
    for k in range(B.shape[0]):
        assert np.array_equal(Y[k], h.apply(B[k]))


def test_python_property_diagonal_scaling(oracle):
    import sparsesolv_amd as ss
    A = varcoef(8)
    pre = ss.ICPreconditioner(A)
    pre.use_abmc = False
    pre.diagonal_scaling = True
    x = P.uniform(7, A.n, -1.0, 1.0)
    yo = oracle.ic(A, oracle.config(use_abmc=0, diagonal_scaling=1)).apply(x)
    assert rel(pre * x, yo) <= 1e-12


@pytest.mark.parametrize("use_abmc,path1", [(1, 1), (1, 0), (0, 0)])
def test_batch_with_diagonal_scaling(use_abmc, path1):
    # each system scaled by its own S: the batched solve gives each system's
    # one-system result bit for bit (scaled applies in the Path 1 and level
    # modes, the unscaled ABMC apply of the scaled factor in Path 3)
    mats = [varcoef(9, seed=5), varcoef(7, seed=6), P.laplace7(8)]
    bs = [P.uniform(70 + k, A.n, -1.0, 1.0) for k, A in enumerate(mats)]
    cfg = _lib.default_config(use_abmc=use_abmc, abmc_reorder_spmv=path1, diagonal_scaling=1,
                              max_iterations=2000)
    hb = _lib.BatchHandle(mats, cfg)
    x, results = hb.solve_batch(_lib.METHOD_ICCG, np.concatenate(bs), None, cfg)
    for k, (A, b, xk) in enumerate(zip(mats, bs, hb.split(x))):
        xs, rs, _ = _lib.IcHandle(A, cfg).solve(_lib.METHOD_ICCG, b, None, cfg)
        assert rs.converged and results[k].iterations == rs.iterations
        assert np.array_equal(xk, xs)


FACTOR_CASES = [n for n, c in sorted(G.CASES.items()) if c[2] in ("iccg", "cocr")]


@pytest.mark.parametrize("name", FACTOR_CASES)
def test_apply_matches_reference_golden(name):
    # the preconditioner's own apply (ICPreconditioner::apply) against the
    # reference's on the fixture's vector: unscaled on the ABMC path,
    # S M^-1 S on the level path
    g = load_golden(name)
    build, kw, _ = G.CASES[name]
    A = build().A
    h = _lib.IcHandle(A, _lib.default_config(**kw))
    assert rel(h.apply(g["apply_x"]), g["apply_y"]) < 1e-12


@pytest.mark.parametrize("name", [n for n in FACTOR_CASES if G.CASES[n][1].get("use_abmc", 0) == 0])
def test_level_solve_matches_reference_golden(name):
    g = load_golden(name)
    build, kw, solver = G.CASES[name]
    s = build()
    cfg = _lib.default_config(save_residual_history=1, **kw)
    code = _lib.METHOD_ICCG if solver == "iccg" else _lib.METHOD_ICCOCR
    x, res, hist = _lib.IcHandle(s.A, cfg).solve(code, s.b, None, cfg)
    n_ref = int(g["iterations"])
    assert bool(res.converged) == bool(g["converged"])
    assert abs(res.iterations - n_ref) <= max(1, math.ceil(0.02 * n_ref))
    assert rel(x, g["x"]) < 1e-9
    np.testing.assert_allclose(hist[:5], g["history"][:5], rtol=1e-9)
