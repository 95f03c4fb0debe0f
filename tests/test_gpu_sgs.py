"""The SGS preconditioner and SGS-MRTR on the GPU.

SGSPreconditioner (preconditioners/sgs_preconditioner.hpp:47-203) is the IC
sweep machinery with A's lower triangle as the factor and the pivots
diag(A), on the level schedule in natural order; CG / COCR run with it as
with IC(0).  SGS-MRTR (solvers/sgs_mrtr_solver.hpp:53-365) iterates in the
scaled space of D A D with the split forward / backward sweeps of
L = lower(D A D).  Each is compared with golden vectors the reference
itself wrote (tests/golden/*sgs*), and on larger systems with the oracle's
restatement (pinned bit for bit against the reference build on the CPU).
The device sums dot products in another order and multiplies by 1/L_ii
where the reference divides by L_ii, so solutions agree to 1e-9 and
iteration counts within 2% (plus one)."""
import math

import numpy as np
import pytest

from sparsesolv_amd import _lib
from sparsesolv_amd import problems as P
from conftest import load_golden

import make_golden as G

pytestmark = pytest.mark.gpu

SGS_CASES = [k for k, (_, _, s) in G.CASES.items() if s.startswith("sgs")]
MRTR_CASES = [k for k, (_, _, s) in G.CASES.items() if s == "sgsmrtr"]


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def it_tol(n):
    return max(1, math.ceil(0.02 * n))


def handle(A, kind, **kw):
    cfg = _lib.default_config(preconditioner=kind, **kw)
    return _lib.IcHandle(A, cfg), cfg


@pytest.mark.parametrize("name", SGS_CASES)
def test_sgs_apply_matches_reference_golden(name):
    g = load_golden(name)
    sysm = G.CASES[name][0]()
    h, _ = handle(sysm.A, _lib.PRECOND_SGS)
    assert h.info()["actual_shift"] == 0.0
    assert rel(h.apply(g["apply_x"]), g["apply_y"]) < 1e-12


@pytest.mark.parametrize("name", ["vc6_sgs_cg", "eddy5_sgs_cocr"])
def test_sgs_krylov_matches_reference_golden(name):
    build, kw, solver = G.CASES[name]
    g = load_golden(name)
    sysm = build()
    h, cfg = handle(sysm.A, _lib.PRECOND_SGS, save_residual_history=1, **kw)
    method = _lib.METHOD_ICCOCR if solver == "sgs_cocr" else _lib.METHOD_ICCG
    x, res, hist = h.solve(method, sysm.b, None, cfg)
    assert bool(res.converged) == bool(g["converged"])
    assert abs(res.iterations - int(g["iterations"])) <= it_tol(int(g["iterations"]))
    assert rel(x, g["x"]) < 1e-9
    np.testing.assert_allclose(hist[:5], g["history"][:5], rtol=1e-9)


@pytest.mark.parametrize("name", MRTR_CASES)
def test_sgsmrtr_matches_reference_golden(name):
    # zvc6: complex diagonals, where the reference's split formula is exact
    # only for diag(D A D) = I; its iterate is followed all the same
    build, kw, _ = G.CASES[name]
    g = load_golden(name)
    sysm = build()
    h, cfg = handle(sysm.A, _lib.PRECOND_SGSMRTR, save_residual_history=1, **kw)
    x, res, hist = h.solve(_lib.METHOD_SGSMRTR, sysm.b, None, cfg)
    assert bool(res.converged) == bool(g["converged"])
    assert rel(x, g["x"]) < 1e-9
    assert res.history_len == res.iterations + 1
    np.testing.assert_allclose(hist[:5], g["history"][:5], rtol=1e-9)
    tol = kw.get("tolerance", 1e-10)
    ref_hist = g["history"]
    if np.count_nonzero((ref_hist >= tol) & (ref_hist < 10 * tol)) >= 5:
        # eddy5: the reference's residual sits at its attainable accuracy,
        # 3-9e-10, for some 20 iterations before one falls below 1e-10; which
        # iteration that is depends on the order of rounding.  The histories
        # agree down to 1000x the tolerance.
        k = int(np.argmax(ref_hist < 1e3 * tol))
        np.testing.assert_allclose(hist[:k], ref_hist[:k], rtol=1e-6)
    else:
        assert abs(res.iterations - int(g["iterations"])) <= it_tol(int(g["iterations"]))
    if g["converged"]:
        assert res.final_residual < kw.get("tolerance", 1e-10)
    else:
        # not converged: the best iterate and its residual
        assert res.final_residual == pytest.approx(float(g["final_residual"]), rel=1e-9)


def varcoef(m, seed=21):
    return P.laplace7_varcoef(m, 10.0 ** P.uniform(seed, m ** 3, -2.0, 2.0))


def systems():
    A = varcoef(16)
    yield "varcoef16", A, P.uniform(3, A.n, -1.0, 1.0), {}
    e = P.eddy(10)
    yield "eddy10", e.A, e.b, dict(max_iterations=3000)
    A = varcoef(12).to_scipy().astype(np.complex128)
    import scipy.sparse as sp
    Z = P.CSR.from_scipy((A + sp.identity(A.shape[0]) * 0.3j).tocsr())
    z = P.gaussian(9, 2 * Z.n)
    yield "zvarcoef12-conjugate", Z, z[0::2] + 1j * z[1::2], dict(conjugate=1)


@pytest.mark.parametrize("case", list(systems()), ids=lambda c: c[0])
def test_sgs_against_oracle(oracle, case):
    _, A, b, kw = case
    h, cfg = handle(A, _lib.PRECOND_SGS, **kw)
    x = P.uniform(11, A.n, -1.0, 1.0).astype(A.values.dtype)
    assert rel(h.apply(x), oracle.sgs_apply(A, x)) < 1e-12
    method = "COCR" if A.is_complex and not kw.get("conjugate") else "CG"
    xg, res, _ = h.solve(_lib.METHOD_ICCOCR if method == "COCR" else _lib.METHOD_ICCG, b, None, cfg)
    xo, ro = oracle.solve_sgs(A, b, oracle.config(**kw), method)
    assert bool(res.converged) == bool(ro["converged"])
    assert abs(res.iterations - ro["iterations"]) <= it_tol(ro["iterations"])
    assert rel(xg, xo) < 1e-9


@pytest.mark.parametrize("case", list(systems()), ids=lambda c: c[0])
def test_sgsmrtr_against_oracle(oracle, case):
    _, A, b, kw = case
    h, cfg = handle(A, _lib.PRECOND_SGSMRTR, save_residual_history=1, **kw)
    x, res, hist = h.solve(_lib.METHOD_SGSMRTR, b, None, cfg)
    xo, ro = oracle.solve_sgsmrtr(A, b, oracle.config(save_residual_history=1, **kw))
    assert bool(res.converged) == bool(ro["converged"])
    assert abs(res.iterations - ro["iterations"]) <= it_tol(ro["iterations"])
    assert rel(x, xo) < 1e-9
    np.testing.assert_allclose(hist[:5], ro["residual_history"][:5], rtol=1e-9)
    if not A.is_complex:
        # a real SPD system: the solution itself
        r = b - A.to_scipy() @ x
        assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-8


def test_sgsmrtr_initial_guess_and_early_exit(oracle):
    A = varcoef(8)
    b = P.uniform(4, A.n, -1.0, 1.0)
    h, cfg = handle(A, _lib.PRECOND_SGSMRTR)
    x0 = P.uniform(6, A.n, -1.0, 1.0)
    x, res, _ = h.solve(_lib.METHOD_SGSMRTR, b, x0, cfg)
    xo, ro = oracle.solve_sgsmrtr(A, b, oracle.config(), x0)
    assert res.converged and ro["converged"]
    assert abs(res.iterations - ro["iterations"]) <= it_tol(ro["iterations"])
    assert rel(x, xo) < 1e-9
    # started at the solution: converged with no iteration (|r| / |D b|
    # below tolerance / 10)
    x2, res2, _ = h.solve(_lib.METHOD_SGSMRTR, b, x, _lib.default_config(preconditioner=_lib.PRECOND_SGSMRTR, tolerance=1e-6))
    xo2, ro2 = oracle.solve_sgsmrtr(A, b, oracle.config(tolerance=1e-6), x)
    assert res2.converged and res2.iterations == 0 and ro2["iterations"] == 0
    # x comes back as D (D^-1 x), as in the reference
    assert rel(x2, xo2) < 1e-15 and rel(x2, x) < 1e-15


def test_split_handles_refuse_what_they_cannot_do():
    A = varcoef(6)
    b = np.ones(A.n)
    h, cfg = handle(A, _lib.PRECOND_SGSMRTR)
    with pytest.raises(_lib.SsaError):
        h.apply(b)
    with pytest.raises(_lib.SsaError):
        h.solve(_lib.METHOD_ICCG, b, None, cfg)
    g, gcfg = handle(A, _lib.PRECOND_SGS)
    with pytest.raises(_lib.SsaError):
        g.solve(_lib.METHOD_SGSMRTR, b, None, gcfg)
    with pytest.raises(_lib.SsaError):
        _lib.IcHandle(A, _lib.default_config(preconditioner=_lib.PRECOND_SGSMRTR,
                                             precision=_lib.PRECISION_MIXED))


# ---------------------------------------------------------------------------
# Python drop-in: SGSPreconditioner and SparseSolvSolver(method="SGSMRTR"),
# the cases of the reference's tests/test_sparsesolv.py (SGSMRTR with
# freedofs, SGS with an outside CG, complex auto-dispatch, Hermitian systems
# with conjugate=True, printrates)
# ---------------------------------------------------------------------------
def test_sparsesolv_solver_sgsmrtr_with_freedofs(oracle):
    import sparsesolv_amd as ss
    from sparsesolv_amd.solvers import _with_freedofs
    s = P.poisson(10)
    free = np.ones(s.A.n, dtype=bool)
    free[::7] = False
    solver = ss.SparseSolvSolver(s.A, method="SGSMRTR", freedofs=free, tol=1e-10, maxiter=2000)
    assert solver.method == "SGSMRTR"
    sol = np.full(s.A.n, 3.0)
    res = solver.Solve(s.b, sol)
    assert res.converged and solver.last_result.iterations == res.iterations
    assert np.all(sol[~free] == 3.0)          # constrained entries untouched
    M = _with_freedofs(s.A, free)
    xo, ro = oracle.solve_sgsmrtr(M, np.where(free, s.b, 0), oracle.config(max_iterations=2000),
                                  x0=np.where(free, 3.0, 0.0))
    assert abs(res.iterations - ro["iterations"]) <= it_tol(ro["iterations"])
    assert rel(sol[free], xo[free]) < 1e-9
    # Mult: zero initial guess, the solution on the free entries
    y = solver * s.b
    r = (s.A.to_scipy() @ y - s.b)[free]
    assert np.linalg.norm(r) / np.linalg.norm(s.b[free]) < 1e-8


def test_sgs_preconditioner_with_an_outside_cg():
    # the reference hands its SGSPreconditioner to NGSolve's CGSolver; here
    # scipy's CG calls Mult through a LinearOperator
    import scipy.sparse.linalg as sla
    import sparsesolv_amd as ss
    A = varcoef(12)
    b = P.uniform(2, A.n, -1.0, 1.0)
    pre = ss.SGSPreconditioner(A)
    assert isinstance(pre, ss.SGSPreconditionerD)
    M = sla.LinearOperator((A.n, A.n), matvec=lambda v: pre.Mult(np.ascontiguousarray(v)))
    its = []
    x, info = sla.cg(A.to_scipy(), b, M=M, rtol=1e-10, maxiter=500, callback=lambda _: its.append(1))
    assert info == 0
    assert np.linalg.norm(b - A.to_scipy() @ x) / np.linalg.norm(b) < 1e-9
    # the same preconditioner in the GPU's own CG: the same count
    h, cfg = handle(A, _lib.PRECOND_SGS, tolerance=1e-10)
    _, res, _ = h.solve(_lib.METHOD_ICCG, b, None, cfg)
    assert res.converged and abs(res.iterations - len(its)) <= 2


def test_sgs_factory_complex_dispatch_and_cocr_solver():
    import sparsesolv_amd as ss
    A, b = complex_laplace(10)
    pre = ss.SGSPreconditioner(A)
    assert isinstance(pre, ss.SGSPreconditionerC)
    x = pre * b
    assert x.dtype == np.complex128 and np.linalg.norm(x) > 0
    inv = ss.COCRSolver(A, pre, maxiter=2000, tol=1e-10)
    y = inv * b
    r = b - A.to_scipy() @ y
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-9


def complex_laplace(m):
    """7-point Laplacian + 0.3j I: complex symmetric, real-dominant diagonal."""
    import scipy.sparse as sp
    L = P.laplace7(m).to_scipy().astype(np.complex128)
    A = P.CSR.from_scipy((L + 0.3j * sp.identity(L.shape[0])).tocsr())
    z = P.gaussian(7, 2 * A.n)
    return A, z[0::2] + 1j * z[1::2]


def hermitian_laplace(m):
    """A Hermitian system in a complex space: real coefficients, complex RHS."""
    L = P.laplace7(m).to_scipy().astype(np.complex128)
    A = P.CSR.from_scipy(L.tocsr())
    z = P.gaussian(8, 2 * A.n)
    return A, z[0::2] + 1j * z[1::2]


@pytest.mark.parametrize("method", ["ICCG", "SGSMRTR"])
def test_hermitian_system_with_conjugate(oracle, method):
    import sparsesolv_amd as ss
    A, b = hermitian_laplace(10)
    solver = ss.SparseSolvSolver(A, method=method, tol=1e-10, maxiter=2000, conjugate=True)
    assert isinstance(solver, ss.SparseSolvSolverC)
    x = solver * b
    res = solver.last_result
    assert res.converged
    assert np.linalg.norm(b - A.to_scipy() @ x) / np.linalg.norm(b) < 1e-8
    if method == "SGSMRTR":
        xo, ro = oracle.solve_sgsmrtr(A, b, oracle.config(conjugate=1, max_iterations=2000))
        assert abs(res.iterations - ro["iterations"]) <= it_tol(ro["iterations"])
        assert rel(x, xo) < 1e-9


def test_sgsmrtr_printrates_and_refusals(capsys):
    import sparsesolv_amd as ss
    g = load_golden("vc6_sgsmrtr")
    s = G.varcoef6()
    solver = ss.SparseSolvSolver(s.A, method="SGSMRTR", printrates=True)
    x = solver * s.b
    res = solver.last_result
    assert res.converged and res.iterations == int(g["iterations"])
    assert rel(x, g["x"]) < 1e-9
    lines = capsys.readouterr().out.splitlines()
    assert len(lines) == len(res.residual_history) + 1
    assert lines[0] == f"SGSMRTR iteration 0, residual = {g['history'][0]:g}"
    assert lines[-1] == (f"SGSMRTR converged in {res.iterations} iterations, "
                         f"residual = {res.final_residual:g}")
    # SGS-MRTR runs in FP64 and one vector at a time
    solver.mixed_precision = True
    with pytest.raises(_lib.SsaError):
        solver * s.b
    solver.mixed_precision = False
    with pytest.raises(RuntimeError):
        solver.SolveBlock(np.stack([s.b, s.b]), np.zeros((2, s.A.n)))


def test_sgs_blocks_and_batches_match_single_solves():
    # the SGS factor through the block (several right-hand sides) and batch
    # paths: each right-hand side / system is its one-vector solve, bit for bit
    A = varcoef(10)
    B = np.stack([P.uniform(60 + k, A.n, -1.0, 1.0) for k in range(6)])
    h, cfg = handle(A, _lib.PRECOND_SGS, max_iterations=1000)
    X, res = h.solve_multi(_lib.METHOD_ICCG, B, None, cfg)
    Y = h.apply_multi(B)
    for k in range(B.shape[0]):
        x1, r1, _ = h.solve(_lib.METHOD_ICCG, B[k], None, cfg)
        np.testing.assert_array_equal(X[k], x1)
        assert res[k].iterations == r1.iterations and res[k].converged
        np.testing.assert_array_equal(Y[k], h.apply(B[k]))
    mats = [varcoef(m, 70 + m) for m in (7, 9, 6)]
    bs = [P.uniform(80 + k, M.n, -1.0, 1.0) for k, M in enumerate(mats)]
    hb = _lib.BatchHandle(mats, cfg)
    xb, rb = hb.solve_batch(_lib.METHOD_ICCG, np.concatenate(bs), None, cfg)
    for k, (M, bk) in enumerate(zip(mats, bs)):
        xk, rk, _ = _lib.IcHandle(M, cfg).solve(_lib.METHOD_ICCG, bk, None, cfg)
        np.testing.assert_array_equal(hb.split(xb)[k], xk)
        assert rb[k].iterations == rk.iterations
    with pytest.raises(_lib.SsaError):      # SGS-MRTR: one system per handle
        _lib.BatchHandle(mats, _lib.default_config(preconditioner=_lib.PRECOND_SGSMRTR))


def test_sgs_mixed_precision_against_fp64_oracle(oracle):
    # the SGS factor and sweeps in float inside the FP64 CG / COCR with
    # refinement: the FP64 oracle's solution to 1e-8
    A = varcoef(12)
    b = P.uniform(5, A.n, -1.0, 1.0)
    h, cfg = handle(A, _lib.PRECOND_SGS, precision=_lib.PRECISION_MIXED, max_iterations=2000)
    x, res, _ = h.solve(_lib.METHOD_ICCG, b, None, cfg)
    xo, ro = oracle.solve_sgs(A, b, oracle.config(max_iterations=2000), "CG")
    assert res.converged and ro["converged"]
    assert np.linalg.norm(b - A.to_scipy() @ x) / np.linalg.norm(b) < 1e-10
    assert rel(x, xo) < 1e-8
    assert res.iterations <= int(1.6 * ro["iterations"]) + 3
    Z, zb = complex_laplace(8)
    hz, cz = handle(Z, _lib.PRECOND_SGS, precision=_lib.PRECISION_MIXED, max_iterations=2000)
    xz, rz, _ = hz.solve(_lib.METHOD_ICCOCR, zb, None, cz)
    xzo, rzo = oracle.solve_sgs(Z, zb, oracle.config(max_iterations=2000), "COCR")
    assert rz.converged and rzo["converged"] and rel(xz, xzo) < 1e-8
