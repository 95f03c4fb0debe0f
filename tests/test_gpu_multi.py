"""Several right-hand sides on one matrix (ssa_solve_multi and friends).

The block kernels load each matrix and factor entry once per lane and apply
it to every right-hand side; the vector kernels, reductions and control steps
run per right-hand side with the launch shape of a one-vector solve, and the
single-wavefront parts (oversized-block walk, small-colour tail) run one
workgroup per right-hand side.  So each right-hand side of a block solve must
reproduce its own single solve bit for bit -- solution, iterations, status
and residual -- while the right-hand sides stop at different iterations."""
import ctypes as C

import numpy as np
import pytest

from sparsesolv_amd import _lib
from sparsesolv_amd import problems as P
from test_gpu_parity import complex_shifted_laplace

pytestmark = pytest.mark.gpu


def rhs_block(n, count, seed, complex_=False):
    B = np.stack([P.uniform(seed + k, n, -1.0, 1.0) * (1.0 + k) for k in range(count)])
    if complex_:
        B = B + 1j * np.stack([P.uniform(seed + 100 + k, n, -1.0, 1.0) for k in range(count)])
    return B


def varcoef(m, seed=21):
    return P.laplace7_varcoef(m, 10.0 ** P.uniform(seed, m ** 3, -2.0, 2.0))


def check_against_singles(A, B, cfg, method, X0=None):
    h = _lib.IcHandle(A, cfg)
    X, results = h.solve_multi(method, B, X0, cfg)
    iters = []
    for k in range(B.shape[0]):
        xk, rk, _ = h.solve(method, B[k], None if X0 is None else X0[k], cfg)
        assert np.array_equal(X[k], xk), k
        assert results[k].iterations == rk.iterations, k
        assert results[k].converged == rk.converged, k
        assert results[k].status == rk.status, k
        assert results[k].final_residual == rk.final_residual, k
        iters.append(rk.iterations)
    return iters


@pytest.mark.parametrize("nrhs", [1, 3, 8, 16, 20])
def test_multi_iccg_matches_single_solves(nrhs):
    # 3: a block of 4 with one padding column; 20: blocks of 16 and 4
    A = varcoef(14)
    cfg = _lib.default_config(use_abmc=1, max_iterations=800)
    iters = check_against_singles(A, rhs_block(A.n, nrhs, 40), cfg, _lib.METHOD_ICCG)
    if nrhs >= 8:
        assert len(set(iters)) > 1          # the right-hand sides stop apart


def test_multi_initial_guess_and_plain_cg():
    A = varcoef(10, seed=5)
    B = rhs_block(A.n, 4, 50)
    X0 = rhs_block(A.n, 4, 60) * 0.1
    cfg = _lib.default_config(use_abmc=1, max_iterations=800)
    check_against_singles(A, B, cfg, _lib.METHOD_ICCG, X0)
    check_against_singles(A, B, cfg, _lib.METHOD_CG)


def test_multi_cocr_complex():
    s = P.eddy(6)
    cfg = _lib.default_config(use_abmc=1, abmc_num_colors=8, max_iterations=600)
    check_against_singles(s.A, rhs_block(s.A.n, 8, 70, complex_=True), cfg, _lib.METHOD_ICCOCR)


def test_multi_auto_shift_curlcurl_block16():
    s = P.curlcurl(6)
    cfg = _lib.default_config(use_abmc=1, auto_shift=1, abmc_block_size=16, max_iterations=400)
    check_against_singles(s.A, rhs_block(s.A.n, 5, 80), cfg, _lib.METHOD_ICCG)


@pytest.mark.parametrize("case", ["curl6_b16_w16", "eddy6_b8_w16_mixed"])
def test_multi_split_sweep_block_sizes(case):
    # 16 right-hand sides: four lanes a row, 16 rows a wavefront, so 16-row
    # blocks fill a wavefront and 8-row blocks pair up within one
    if case == "curl6_b16_w16":
        s = P.curlcurl(6)
        cfg = _lib.default_config(use_abmc=1, auto_shift=1, abmc_block_size=16, max_iterations=400)
        check_against_singles(s.A, rhs_block(s.A.n, 16, 85), cfg, _lib.METHOD_ICCG)
    else:
        s = P.eddy(6)
        cfg = _lib.default_config(use_abmc=1, abmc_block_size=8, abmc_num_colors=6, max_iterations=600,
                                  precision=_lib.PRECISION_MIXED)
        check_against_singles(s.A, rhs_block(s.A.n, 16, 86, complex_=True), cfg, _lib.METHOD_ICCOCR)


def test_multi_walk_of_oversized_block():
    # 48^3 Laplacian, B=4: one 366-row straggler block walked per right-hand side
    s = P.poisson(48)
    cfg = _lib.default_config(use_abmc=1, max_iterations=400)
    h = _lib.IcHandle(s.A, cfg)
    assert h.info()["big_blocks"] == 1
    check_against_singles(s.A, rhs_block(s.A.n, 4, 90), cfg, _lib.METHOD_ICCG)


def test_multi_level_path():
    A = varcoef(9, seed=7)
    cfg = _lib.default_config(use_abmc=0, max_iterations=800)
    check_against_singles(A, rhs_block(A.n, 4, 100), cfg, _lib.METHOD_ICCG)


@pytest.mark.parametrize("case", ["varcoef-iccg", "zlap-cocr", "eddy-cocr"])
def test_multi_mixed_precision(case):
    # per right-hand side iterative refinement: restarts, undo of a pass
    # that does not help, and the stop all follow that right-hand side alone
    kw = {}
    if case == "varcoef-iccg":
        A, method = varcoef(12, seed=9), _lib.METHOD_ICCG
        B = rhs_block(A.n, 6, 110)
    elif case == "zlap-cocr":
        A, _ = complex_shifted_laplace(10)
        method, kw = _lib.METHOD_ICCOCR, dict(abmc_num_colors=8)
        B = rhs_block(A.n, 5, 120, complex_=True)
    else:
        # near-singular: a float factor stalls COCR above the tolerance, so
        # the passes end unconverged at different points per right-hand side
        A, method, kw = P.eddy(6).A, _lib.METHOD_ICCOCR, dict(abmc_num_colors=8)
        B = rhs_block(A.n, 4, 125, complex_=True)
    cfg = _lib.default_config(use_abmc=1, precision=_lib.PRECISION_MIXED, max_iterations=1500, **kw)
    check_against_singles(A, B, cfg, method)
    X, res = _lib.IcHandle(A, cfg).solve_multi(method, B, None, cfg)
    As = A.to_scipy()
    for k in range(B.shape[0]):
        true_rel = np.linalg.norm(B[k] - As @ X[k]) / np.linalg.norm(B[k])
        # two FP64 evaluations of b - Ax differ by up to the rounding floor
        # ~ eps * || |A| |x| || (large on the near-singular system)
        floor = 4e-15 * np.linalg.norm(abs(As) @ abs(X[k])) / np.linalg.norm(B[k])
        assert abs(res[k].final_residual - true_rel) <= 1e-3 * true_rel + floor
        if case != "eddy-cocr":
            assert res[k].converged and true_rel < 1e-10


def test_multi_graph_and_eager_agree():
    A = varcoef(10, seed=3)
    B = rhs_block(A.n, 8, 130)
    out = []
    for graph in (1, 0):
        cfg = _lib.default_config(use_abmc=1, use_graph=graph)
        X, res = _lib.IcHandle(A, cfg).solve_multi(_lib.METHOD_ICCG, B, None, cfg)
        out.append((X, [r.iterations for r in res]))
    assert np.array_equal(out[0][0], out[1][0]) and out[0][1] == out[1][1]


@pytest.mark.parametrize("nrhs", [2, 8, 16])
def test_multi_apply_and_spmv_match_single(oracle, nrhs):
    s = P.eddy(6)
    h = _lib.IcHandle(s.A, _lib.default_config(use_abmc=1, abmc_num_colors=8))
    X = rhs_block(s.A.n, nrhs, 140, complex_=True)
    Y = h.apply_multi(X)
    AX = h.spmv_multi(X)
    for k in range(nrhs):
        assert np.array_equal(Y[k], h.apply(X[k])), k
        assert np.array_equal(AX[k], h.spmv(X[k])), k
    yo = oracle.ic(s.A, oracle.config(use_abmc=1, abmc_num_colors=8)).apply(X[-1])
    assert np.linalg.norm(Y[-1] - yo) <= 1e-12 * np.linalg.norm(yo)


def test_multi_profile_kinds():
    s = P.poisson(20)
    h = _lib.IcHandle(s.A, _lib.default_config(use_abmc=1))
    for what in (_lib.PROFILE_SPMV, _lib.PROFILE_FORWARD, _lib.PROFILE_BACKWARD, _lib.PROFILE_APPLY):
        us1, by1 = h.profile_multi(what, 1, reps=3)
        us16, by16 = h.profile_multi(what, 16, reps=3)
        assert us1 > 0 and us16 > 0 and by16 > by1 > 0
    with pytest.raises(_lib.SsaError):
        h.profile_multi(_lib.PROFILE_SPMV, 17)


def test_multi_rejections():
    A = varcoef(6)
    cfg = _lib.default_config(use_abmc=1)
    h = _lib.IcHandle(A, cfg)
    B = rhs_block(A.n, 2, 150)
    X = np.zeros_like(B)
    res = (_lib.Result * 2)()
    with pytest.raises(_lib.SsaError):      # COCRSolverNGS runs one right-hand side
        h.solve_multi(_lib.METHOD_NGS_ICCOCR, B, None, cfg)
    with pytest.raises(_lib.SsaError):
        _lib.check(h.lib.ssa_solve_multi(h.h, _lib.METHOD_ICCG, 0, B.ctypes.data, X.ctypes.data, 0,
                                         C.byref(cfg), res))
    hb = _lib.BatchHandle([A, varcoef(5)], cfg)
    with pytest.raises(_lib.SsaError):
        hb.solve_multi(_lib.METHOD_ICCG, np.zeros((2, hb.n)), None, cfg)


def test_python_front_end_block_solve():
    # SparseSolvSolver.SolveBlock / Mult on a (k, n) array and the
    # preconditioner's Mult on a block: each row as the one-vector call gives it
    import sparsesolv_amd as ss
    s = P.eddy(5)
    B = rhs_block(s.A.n, 3, 160, complex_=True)
    solver = ss.SparseSolvSolver(s.A, method="COCR", tol=1e-10, maxiter=400)
    solver.use_abmc = True
    solver.abmc_num_colors = 8
    X = np.zeros_like(B)
    results = solver.SolveBlock(B, X)
    assert len(results) == 3
    for k in range(3):
        xk = np.zeros(s.A.n, dtype=complex)
        rk = solver.Solve(B[k], xk)
        assert np.array_equal(X[k], xk)
        assert results[k].iterations == rk.iterations
    assert np.array_equal(solver * B, X)
    pre = ss.ICPreconditioner(s.A)
    pre.use_abmc = True
    Y = pre * B
    for k in range(3):
        assert np.array_equal(Y[k], pre * B[k])
