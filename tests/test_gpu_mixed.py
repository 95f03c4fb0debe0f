"""Mixed precision: the IC(0) factor (computed in FP64) stored and applied in
float / complex<float>, SpMV, Krylov vectors and dot products in FP64, with
an FP64 true-residual check and restart after convergence.

The reference has no mixed form; the bar is the FP64 oracle: the mixed solve
reaches the same tolerance on the true residual, needs about the same
iterations (COCR: refinement restarts add some), and its solution is as close
to the oracle's FP64 solution as the tolerance allows on these
well-conditioned systems."""
import numpy as np
import pytest

from sparsesolv_amd import _lib
from sparsesolv_amd import problems as P
from test_gpu_parity import complex_shifted_laplace, rel

pytestmark = pytest.mark.gpu


def mixed(A, **kw):
    return _lib.IcHandle(A, _lib.default_config(use_abmc=1, precision=_lib.PRECISION_MIXED, **kw))


def varcoef(m, seed=11):
    coef = 10.0 ** P.uniform(seed, m ** 3, -1.0, 1.0)
    A = P.laplace7_varcoef(m, coef)
    return A, P.uniform(seed + 1, m ** 3, -1.0, 1.0)


@pytest.mark.parametrize("case", ["lap20", "varcoef16", "zlap12"])
def test_mixed_solve_matches_fp64_oracle(oracle, case):
    if case == "lap20":
        s = P.poisson(20)
        A, b, method, kw = s.A, s.b, "iccg", {}
    elif case == "varcoef16":
        A, b = varcoef(16)
        method, kw = "iccg", dict(abmc_num_colors=6)
    else:
        A, b = complex_shifted_laplace(12)
        method, kw = "cocr", dict(abmc_num_colors=8)
    cfg = _lib.default_config(use_abmc=1, precision=_lib.PRECISION_MIXED, tolerance=1e-10,
                              max_iterations=2000, **kw)
    h = _lib.IcHandle(A, cfg)
    code = _lib.METHOD_ICCG if method == "iccg" else _lib.METHOD_ICCOCR
    x, res, _ = h.solve(code, b, None, cfg)
    ocfg = oracle.config(use_abmc=1, tolerance=1e-10, max_iterations=2000, **kw)
    xo, ro = (oracle.solve_iccg if method == "iccg" else oracle.solve_cocr)(A, b, ocfg)
    assert res.converged and ro["converged"]
    true_rel = np.linalg.norm(b - A.to_scipy() @ x) / np.linalg.norm(b)
    assert true_rel < 1e-10
    assert abs(res.final_residual - true_rel) <= 1e-3 * true_rel + 1e-16
    # a float preconditioner changes CG little on these systems; COCR
    # restarts once (iterative refinement) and loses some of its subspace
    slack = 1.1 if method == "iccg" else 1.6
    assert res.iterations <= int(slack * ro["iterations"]) + 3
    assert rel(x, xo) < 1e-8


def test_mixed_apply_is_single_precision_accurate(oracle):
    s = P.poisson(16)
    h = mixed(s.A)
    x = P.uniform(5, s.A.n, -1.0, 1.0)
    y = h.apply(x)
    yo = oracle.ic(s.A, oracle.config(use_abmc=1)).apply(x)
    e = rel(y, yo)
    assert 1e-12 < e < 1e-5      # float arithmetic, but no worse than that


def test_mixed_rejects_plain_methods():
    s = P.poisson(8)
    cfg = _lib.default_config(use_abmc=1, precision=_lib.PRECISION_MIXED)
    h = _lib.IcHandle(s.A, cfg)
    with pytest.raises(_lib.SsaError):
        h.solve(_lib.METHOD_CG, s.b, None, cfg)


def test_unknown_precision_rejected():
    s = P.poisson(6)
    with pytest.raises(_lib.SsaError):
        _lib.IcHandle(s.A, _lib.default_config(use_abmc=1, precision=7))


def test_mixed_graph_and_eager_agree_bitwise():
    s = P.poisson(14)
    out = []
    for graph in (1, 0):
        cfg = _lib.default_config(use_abmc=1, precision=_lib.PRECISION_MIXED, use_graph=graph)
        x, res, _ = _lib.IcHandle(s.A, cfg).solve(_lib.METHOD_ICCG, s.b, None, cfg)
        out.append((x, res.iterations))
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][0], out[1][0])


@pytest.mark.parametrize("case", ["cc8-iccg", "eddy8-cocr"])
def test_mixed_unconverged_reports_its_own_residual(case):
    # near-singular systems that do not reach the tolerance in the budget:
    # the reported residual is the true residual of the returned x, and x is
    # never worse than the initial guess (x0 = 0, relative residual 1)
    name, method = case.split("-")
    s = P.curlcurl(8) if name == "cc8" else P.eddy(8)
    cfg = _lib.default_config(use_abmc=1, precision=_lib.PRECISION_MIXED, tolerance=1e-14,
                              max_iterations=60, auto_shift=1, abmc_num_colors=8)
    h = _lib.IcHandle(s.A, cfg)
    code = _lib.METHOD_ICCG if method == "iccg" else _lib.METHOD_ICCOCR
    x, res, _ = h.solve(code, s.b, None, cfg)
    true_rel = np.linalg.norm(s.b - s.A.to_scipy() @ x) / np.linalg.norm(s.b)
    assert not res.converged
    assert res.iterations <= 60
    assert abs(res.final_residual - true_rel) <= 1e-6 * true_rel
    assert true_rel <= 1.0
