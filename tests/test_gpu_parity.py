"""GPU parity: the HIP path (through the C ABI) against the oracle and the
reference's golden outputs.

Bars (north star): ABMC arrays and real factors bit-exact (checked on the
host handle in test_host_setup.py); operator / preconditioner applications
within 1e-12 relative (FMA and summation order differ from the CPU);
iteration counts within 2% (at least +-1); solutions within 1e-9 relative of
the oracle's at tol 1e-10.
"""
import math

import numpy as np
import pytest

from conftest import load_golden
from sparsesolv_amd import _lib
from sparsesolv_amd import problems as P

import make_golden as G

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def it_tol(n_ref):
    return max(1, math.ceil(0.02 * n_ref))


def handle(A, **kw):
    return _lib.IcHandle(A, _lib.default_config(use_abmc=1, **kw))


SYSTEMS = {
    "lap12": lambda: P.poisson(12),
    "lap20": lambda: P.poisson(20),
    "cc8": lambda: P.curlcurl(8),
    "eddy8": lambda: P.eddy(8),
}


@pytest.mark.parametrize("name", sorted(SYSTEMS))
def test_spmv_matches_scipy(name):
    s = SYSTEMS[name]()
    h = handle(s.A)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(s.A.n)
    if s.A.is_complex:
        x = x + 1j * rng.standard_normal(s.A.n)
    y = h.spmv(x)
    want = s.A.to_scipy() @ x
    assert rel(y, want) < 1e-13


@pytest.mark.parametrize("name,kw", [("lap12", {}), ("lap20", dict(abmc_block_size=8)),
                                     ("lap20", dict(abmc_block_size=16, abmc_num_colors=16)),
                                     ("lap12", dict(abmc_block_size=3, abmc_num_colors=5)),
                                     ("cc8", dict(auto_shift=1)), ("eddy8", dict(abmc_num_colors=8)),
                                     ("lap20", dict(abmc_block_size=2))])
def test_ic_apply_matches_oracle(oracle, name, kw):
    s = SYSTEMS[name]()
    h = handle(s.A, **kw)
    ic = oracle.ic(s.A, oracle.config(use_abmc=1, **kw))
    rng = np.random.default_rng(2)
    x = rng.standard_normal(s.A.n)
    if s.A.is_complex:
        x = x + 1j * rng.standard_normal(s.A.n)
    assert rel(h.apply(x), ic.apply(x)) < 1e-12


def test_ic_apply_with_big_block(oracle):
    # 48^3 Laplacian, B=4: one 366-row straggler block (27 dependency levels)
    s = P.poisson(48)
    h = handle(s.A)
    info = h.info()
    assert info["big_blocks"] == 1 and info["max_big_rows"] == 366
    ic = oracle.ic(s.A, oracle.config(use_abmc=1))
    x = P.uniform(4, s.A.n, -1, 1)
    assert rel(h.apply(x), ic.apply(x)) < 1e-12


@pytest.mark.parametrize("name", [n for n, c in sorted(G.CASES.items())
                                  if c[1].get("use_abmc", 0) == 1 or c[2].startswith("plain")])
def test_solver_matches_reference_golden(name):
    g = load_golden(name)
    build, kw, solver = G.CASES[name]
    s = build()
    kw = dict(kw)
    kw["use_abmc"] = 1
    kw.setdefault("abmc_reorder_spmv", 0)
    cfg = _lib.default_config(save_residual_history=1, **kw)
    h = _lib.IcHandle(s.A, cfg)
    method = {"iccg": _lib.METHOD_ICCG, "cocr": _lib.METHOD_ICCOCR,
              "plain_cg": _lib.METHOD_CG, "plain_cocr": _lib.METHOD_COCR}[solver]
    x, res, hist = h.solve(method, s.b, None, cfg)
    n_ref = int(g["iterations"])
    assert bool(res.converged) == bool(g["converged"])
    assert abs(res.iterations - n_ref) <= it_tol(n_ref)
    if g["converged"]:
        assert rel(x, g["x"]) < 1e-9
        np.testing.assert_allclose(hist[:5], g["history"][:5], rtol=1e-9)
    else:
        # not converged: best iterate restored, residual of the returned x
        Ax = s.A.to_scipy() @ x
        r = np.linalg.norm(s.b - Ax) / np.linalg.norm(s.b)
        assert abs(r - res.final_residual) <= 1e-8 * max(r, 1e-30) + 1e-14
        # the early history follows the reference; a long non-converging run
        # (an indefinite preconditioner) amplifies rounding differences, so
        # its best residual is compared only for short runs, and on the
        # near-singular curl-curl / eddy systems (whose iterates summation
        # order moves by a few per cent within tens of iterations) within 5%
        np.testing.assert_allclose(hist[:5], g["history"][:5], rtol=1e-8)
        if n_ref <= 50:
            tol = 5e-2 if name.startswith(("cc", "eddy")) else 1e-8
            np.testing.assert_allclose(res.final_residual, float(g["final_residual"]), rtol=tol)


def complex_shifted_laplace(m, shift=0.3, seed=6):
    """A well-conditioned complex-symmetric system: 7-point Laplacian + j*shift*I."""
    import scipy.sparse as sp
    from sparsesolv_amd.problems import CSR
    A = P.laplace7(m).to_scipy().astype(complex) + sp.identity(m ** 3) * (1j * shift)
    z = P.gaussian(seed, 2 * m ** 3)
    return CSR.from_scipy(A), z[0::2] + 1j * z[1::2]


# Well-conditioned systems: the full bar (x within 1e-9, history within 1e-8).
@pytest.mark.parametrize("name,kw,method", [
    ("lap20", dict(), "iccg"),
    ("lap20", dict(abmc_block_size=8, abmc_num_colors=6), "iccg"),
    ("zlap12", dict(abmc_num_colors=8), "cocr"),
    ("zlap12", dict(), "iccg"),
])
def test_solve_matches_oracle(oracle, name, kw, method):
    if name == "zlap12":
        A, b = complex_shifted_laplace(12)
    else:
        s = SYSTEMS[name]()
        A, b = s.A, s.b
    cfg = _lib.default_config(use_abmc=1, save_residual_history=1, **kw)
    h = _lib.IcHandle(A, cfg)
    code = _lib.METHOD_ICCG if method == "iccg" else _lib.METHOD_ICCOCR
    x, res, hist = h.solve(code, b, None, cfg)
    fn = oracle.solve_iccg if method == "iccg" else oracle.solve_cocr
    xo, ro = fn(A, b, oracle.config(use_abmc=1, save_residual_history=1, **kw))
    assert res.converged and ro["converged"]
    assert abs(res.iterations - ro["iterations"]) <= it_tol(ro["iterations"])
    assert rel(x, xo) < 1e-9
    k = min(len(hist), len(ro["residual_history"]), 10)
    np.testing.assert_allclose(hist[:k], ro["residual_history"][:k], rtol=1e-8)


# Near-singular curl-curl / eddy systems (eps = 1e-6 gradient null space):
# the solution is determined only to about tol * cond(A).  The reference's own
# Path 1 and Path 3, which differ only in SpMV summation order, differ here by
# 1.4% in iterations and 1.6e-6 in x (cc8); relative perturbations of b of
# 1e-15 move the reference algorithm's iteration count by up to 9.3% (eddy8
# COCR), 6.3% (eddy8 COCG) and 3.2% (cc8).  The bar is therefore: the
# iteration count inside the oracle's own envelope over such perturbations
# (widened by 2%), both solutions satisfy the tolerance (their difference is
# a residual-level vector), and the solution difference within the
# reference's own path-to-path spread.
@pytest.mark.parametrize("name,kw,method", [
    ("cc8", dict(auto_shift=1, max_iterations=5000), "iccg"),
    ("eddy8", dict(abmc_num_colors=8, max_iterations=5000), "cocr"),
    ("eddy8", dict(max_iterations=5000), "iccg"),
])
def test_solve_matches_oracle_near_singular(oracle, name, kw, method):
    s = SYSTEMS[name]()
    cfg = _lib.default_config(use_abmc=1, save_residual_history=1, **kw)
    h = _lib.IcHandle(s.A, cfg)
    code = _lib.METHOD_ICCG if method == "iccg" else _lib.METHOD_ICCOCR
    x, res, hist = h.solve(code, s.b, None, cfg)
    fn = oracle.solve_iccg if method == "iccg" else oracle.solve_cocr
    xo, ro = fn(s.A, s.b, oracle.config(use_abmc=1, save_residual_history=1, **kw))
    assert res.converged and ro["converged"]
    counts = [ro["iterations"]]
    ocfg = oracle.config(use_abmc=1, **kw)
    for seed in range(1, 7):
        bp = s.b * (1.0 + 1e-15 * np.random.default_rng(seed).standard_normal(s.A.n))
        counts.append(fn(s.A, bp, ocfg)[1]["iterations"])
    lo, hi = min(counts), max(counts)
    assert lo - it_tol(lo) <= res.iterations <= hi + it_tol(hi), (res.iterations, counts)
    # The eddy RHS excites the eps-scale gradient modes, so ||x|| ~ 1e6 ||b||
    # and the true residual b - Ax is floored by rounding (~1e-9) for any
    # implementation; the GPU's must match the oracle's, not 1e-10.
    Asp = s.A.to_scipy()
    nb = np.linalg.norm(s.b)
    true_gpu = np.linalg.norm(s.b - Asp @ x) / nb
    true_orc = np.linalg.norm(s.b - Asp @ xo) / nb
    floor = max(1e-10, 3.0 * true_orc)
    assert true_gpu < floor, (true_gpu, true_orc)
    assert np.linalg.norm(Asp @ (x - xo)) / nb < 2.0 * floor
    assert rel(x, xo) < 2e-6
    np.testing.assert_allclose(hist[:3], ro["residual_history"][:3], rtol=1e-9)


def test_graph_and_eager_agree_bitwise():
    s = P.poisson(16)
    base = dict(use_abmc=1, max_iterations=300)
    xs = []
    for use_graph in (1, 0, 1):
        cfg = _lib.default_config(use_graph=use_graph, **base)
        h = _lib.IcHandle(s.A, cfg)
        x, res, _ = h.solve(_lib.METHOD_ICCG, s.b, None, cfg)
        xs.append((x, res.iterations))
    for x, it in xs[1:]:
        assert it == xs[0][1]
        np.testing.assert_array_equal(x, xs[0][0])


def test_best_result_and_history_when_not_converged(oracle):
    s = P.poisson(10)
    cfg = _lib.default_config(use_abmc=1, max_iterations=6, save_residual_history=1)
    h = _lib.IcHandle(s.A, cfg)
    x, res, hist = h.solve(_lib.METHOD_ICCG, s.b, None, cfg)
    xo, ro = oracle.solve_iccg(s.A, s.b, oracle.config(use_abmc=1, max_iterations=6,
                                                       save_residual_history=1))
    assert not res.converged and res.iterations == 6 == ro["iterations"]
    assert len(hist) == len(ro["residual_history"]) == 7
    np.testing.assert_allclose(hist, ro["residual_history"], rtol=1e-9)
    np.testing.assert_allclose(res.final_residual, ro["final_residual"], rtol=1e-9)
    assert rel(x, xo) < 1e-9


def test_initial_guess_and_zero_rhs():
    s = P.poisson(8)
    cfg = _lib.default_config(use_abmc=1)
    h = _lib.IcHandle(s.A, cfg)
    xe = np.linalg.solve(s.A.to_scipy().toarray(), s.b)
    x, res, _ = h.solve(_lib.METHOD_ICCG, s.b, xe, cfg)
    assert res.converged and res.iterations <= 2
    x0, r0, _ = h.solve(_lib.METHOD_ICCG, np.zeros(s.A.n), None, cfg)
    assert r0.converged and np.all(x0 == 0)


@pytest.mark.parametrize("norm_type,custom", [(0, 1.0), (1, 1.0), (2, 3.5), (2, 1e-3)])
@pytest.mark.parametrize("case", ["lap", "zlap"])
def test_norm_types_with_initial_guess_match_oracle(oracle, case, norm_type, custom):
    # NormType RHS / InitialResidual / Custom (core/types.hpp:42-46,
    # iterative_solver.hpp:155-160): the normaliser of the convergence test
    # and of the residual history; the oracle is pinned to the reference
    # build on the same cases (test_oracle_golden.py)
    if case == "lap":
        s = P.poisson(9)
    else:
        A, b = complex_shifted_laplace(9)
        s = P.System(A, b, "zlap9", {})
    method = _lib.METHOD_ICCG if case == "lap" else _lib.METHOD_ICCOCR
    kw = dict(use_abmc=1, abmc_num_colors=4 if case == "lap" else 8, norm_type=norm_type,
              custom_norm=custom, save_residual_history=1, max_iterations=2000)
    x0 = P.uniform(21, s.A.n, -1.0, 1.0)
    if s.A.is_complex:
        x0 = x0 + 1j * P.uniform(22, s.A.n, -1.0, 1.0)
    cfg = _lib.default_config(**kw)
    x, res, hist = _lib.IcHandle(s.A, cfg).solve(method, s.b, x0, cfg)
    fn = oracle.solve_iccg if case == "lap" else oracle.solve_cocr
    xo, ro = fn(s.A, s.b, oracle.config(**kw), x0=x0)
    assert bool(res.converged) == ro["converged"]
    assert abs(res.iterations - ro["iterations"]) <= it_tol(ro["iterations"])
    assert rel(x, xo) < 1e-9
    m = min(len(hist), len(ro["residual_history"]), 6)
    np.testing.assert_allclose(hist[:m], ro["residual_history"][:m], rtol=1e-10)


@pytest.mark.parametrize("system", ["hermitian", "complex_symmetric"])
@pytest.mark.parametrize("method", ["iccg", "cocr"])
def test_conjugate_with_ic_matches_oracle(oracle, system, method):
    # conjugate=True (Hermitian dots) with the ABMC IC(0) preconditioner; the
    # oracle is pinned to the reference build on these systems
    # (test_oracle_golden.py).  Where the oracle converges: the full bar;
    # where it does not (the unconjugated IC(0) of a Hermitian matrix is no
    # Hermitian preconditioner): the early residual history and the best
    # residual's true value
    import scipy.sparse as sp
    L = P.laplace7(7).to_scipy()
    if system == "hermitian":
        S = sp.triu(L, 1) * 0.3j
        A = P.CSR.from_scipy((L.astype(complex) + S - S.T).tocsr())
    else:
        A = P.CSR.from_scipy((L.astype(complex) + 0.4j * sp.identity(343)).tocsr())
    z = P.gaussian(9, 2 * A.n)
    b = z[0::2] + 1j * z[1::2]
    kw = dict(use_abmc=1, conjugate=1, save_residual_history=1, max_iterations=400)
    cfg = _lib.default_config(**kw)
    code = _lib.METHOD_ICCG if method == "iccg" else _lib.METHOD_ICCOCR
    x, res, hist = _lib.IcHandle(A, cfg).solve(code, b, None, cfg)
    fn = oracle.solve_iccg if method == "iccg" else oracle.solve_cocr
    xo, ro = fn(A, b, oracle.config(**kw))
    assert bool(res.converged) == ro["converged"]
    np.testing.assert_allclose(hist[:10], ro["residual_history"][:10], rtol=1e-8)
    if ro["converged"]:
        assert abs(res.iterations - ro["iterations"]) <= it_tol(ro["iterations"])
        assert rel(x, xo) < 1e-9
    else:
        t = np.linalg.norm(b - A.to_scipy() @ x) / np.linalg.norm(b)
        assert abs(t - res.final_residual) <= 1e-8 * t


def test_maxiter_zero():
    s = P.poisson(6)
    cfg = _lib.default_config(use_abmc=1, max_iterations=0)
    h = _lib.IcHandle(s.A, cfg)
    x, res, _ = h.solve(_lib.METHOD_ICCG, s.b, None, cfg)
    assert not res.converged and res.iterations == 0


def test_ngs_cocr_on_eddy_current(oracle):
    # the reference's TestCOCRSolverStandalone checks iterations < 500 and a
    # non-zero solution on an eddy-current problem
    import sparsesolv_amd as ss
    e = P.eddy(7)
    pre = ss.ICPreconditioner(e.A, shift=1.05)
    pre.use_abmc = True
    pre.Update()
    inv = ss.COCRSolver(e.A, pre, maxiter=500, tol=1e-8)
    x = inv * e.b
    assert np.linalg.norm(x) > 0 and inv.iterations < 500
    # the recursive residual reached 1e-8; the true one is floored near it by
    # rounding (||x|| ~ 1e5 ||b|| on this near-singular system)
    assert np.linalg.norm(e.b - e.A.to_scipy() @ x) / np.linalg.norm(e.b) <= 1e-7


def test_python_dropin_iccg_and_cocr(oracle):
    import sparsesolv_amd as ss
    s = P.poisson(14)
    solver = ss.SparseSolvSolver(s.A, method="ICCG", use_abmc=True, save_residual_history=True)
    x = solver * s.b
    xo, ro = oracle.solve_iccg(s.A, s.b, oracle.config(use_abmc=1))
    assert solver.last_result.converged and rel(x, xo) < 1e-9
    assert abs(solver.last_result.iterations - ro["iterations"]) <= it_tol(ro["iterations"])
    A, b = complex_shifted_laplace(10)
    pre = ss.ICPreconditioner(A, shift=1.05)
    pre.use_abmc = True
    pre.abmc_num_colors = 8
    pre.Update()
    inv = ss.COCRSolver(A, pre, maxiter=500, tol=1e-8)
    xc = inv * b
    assert 0 < inv.iterations < 500
    xd = np.linalg.solve(A.to_scipy().toarray(), b)
    assert rel(xc, xd) < 1e-7
    # COCRSolverNGS stops at res/rhs_norm <= tol without the absolute test;
    # otherwise its loop is solve_cocr's
    xo, ro = oracle.solve_cocr(A, b, oracle.config(use_abmc=1, abmc_num_colors=8,
                                                   tolerance=1e-8, abs_tolerance=0.0))
    assert abs(inv.iterations - ro["iterations"]) <= it_tol(ro["iterations"])
    assert rel(xc, xo) < 1e-9


def test_python_dropin_freedofs(oracle):
    import sparsesolv_amd as ss
    from sparsesolv_amd.solvers import _with_freedofs
    s = P.poisson(10)
    free = np.ones(s.A.n, dtype=bool)
    free[::7] = False
    solver = ss.SparseSolvSolver(s.A, method="ICCG", freedofs=free, use_abmc=True)
    sol = np.full(s.A.n, 3.0)
    res = solver.Solve(s.b, sol)
    assert res.converged
    assert np.all(sol[~free] == 3.0)          # constrained entries untouched
    M = _with_freedofs(s.A, free)
    xo, _ = oracle.solve_iccg(M, np.where(free, s.b, 0), oracle.config(use_abmc=1),
                              x0=np.where(free, 3.0, 0.0))
    assert rel(sol[free], xo[free]) < 1e-9


def test_plain_methods_match_oracle(oracle):
    s = P.poisson(9)
    cfg = _lib.default_config(use_abmc=1, max_iterations=500)
    h = _lib.IcHandle(s.A, cfg)
    x, res, _ = h.solve(_lib.METHOD_CG, s.b, None, cfg)
    xo, ro = oracle.solve_plain(s.A, s.b, oracle.config(max_iterations=500), "CG")
    assert res.converged and abs(res.iterations - ro["iterations"]) <= it_tol(ro["iterations"])
    assert rel(x, xo) < 1e-9


@pytest.mark.parametrize("m,colour_kernels", [(32, True), (8, None)])
def test_profile_kinds(m, colour_kernels):
    # small grids may leave no timed colour kernel (tail kernels, colours
    # with oversized-block work): then both figures are zero
    s = P.poisson(m)
    h = handle(s.A, abmc_block_size=4, abmc_num_colors=4)
    for what in (_lib.PROFILE_SPMV, _lib.PROFILE_FORWARD, _lib.PROFILE_BACKWARD,
                 _lib.PROFILE_APPLY):
        us, by = h.profile(what, 2)
        assert us > 0 and by > 0, what
    us, by = h.profile(_lib.PROFILE_COLOUR_KERNELS, 2)
    assert (us > 0 and by > 0) or (colour_kernels is None and us == 0 and by == 0)
    # the colour kernels' bytes are a share of one forward + backward sweep
    _, by_fwd = h.profile(_lib.PROFILE_FORWARD, 1)
    _, by_launch = h.profile(_lib.PROFILE_COLOUR_KERNELS, 1)
    assert by_launch < 2 * by_fwd
    with pytest.raises(_lib.SsaError):
        h.profile(99, 1)
    # the sweeps after profiling still give the preconditioner's result
    x = np.linspace(-1.0, 1.0, s.A.n).astype(s.A.values.dtype)
    y1 = h.apply(x)
    h.profile(_lib.PROFILE_COLOUR_KERNELS, 1)
    assert np.array_equal(h.apply(x), y1)


# ---------------------------------------------------------------------------
# Level-schedule path (use_abmc = false, the factories' default)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name,kw", [("lap12", {}), ("cc8", dict(auto_shift=1)),
                                     ("eddy8", {})])
def test_level_apply_matches_oracle(oracle, name, kw):
    A = {"lap12": lambda: P.poisson(12).A, "cc8": lambda: P.curlcurl(8).A,
         "eddy8": lambda: P.eddy(8).A}[name]()
    h = _lib.IcHandle(A, _lib.default_config(use_abmc=0, **kw))
    x = P.uniform(3, A.n, -1.0, 1.0)
    if A.is_complex:
        x = x + 1j * P.uniform(4, A.n, -1.0, 1.0)
    y = h.apply(x)
    yo = oracle.ic(A, oracle.config(use_abmc=0, **kw)).apply(x)
    assert rel(y, yo) < 1e-12


@pytest.mark.parametrize("case", ["lap20", "varcoef16", "zlap12"])
def test_level_solve_matches_oracle(oracle, case):
    if case == "lap20":
        s = P.poisson(20)
        A, b, method = s.A, s.b, "iccg"
    elif case == "varcoef16":
        coef = 10.0 ** P.uniform(11, 16 ** 3, -1.0, 1.0)
        A, b, method = P.laplace7_varcoef(16, coef), P.uniform(12, 16 ** 3, -1.0, 1.0), "iccg"
    else:
        A, b = complex_shifted_laplace(12)
        method = "cocr"
    cfg = _lib.default_config(use_abmc=0, max_iterations=2000)
    h = _lib.IcHandle(A, cfg)
    code = _lib.METHOD_ICCG if method == "iccg" else _lib.METHOD_ICCOCR
    x, res, _ = h.solve(code, b, None, cfg)
    fn = oracle.solve_iccg if method == "iccg" else oracle.solve_cocr
    xo, ro = fn(A, b, oracle.config(use_abmc=0, max_iterations=2000))
    assert res.converged and ro["converged"]
    assert abs(res.iterations - ro["iterations"]) <= it_tol(ro["iterations"])
    assert rel(x, xo) < 1e-9


def test_level_path_is_the_dropin_default(oracle):
    import sparsesolv_amd as ss
    s = P.poisson(14)
    solver = ss.SparseSolvSolver(s.A, method="ICCG")          # use_abmc defaults to False
    assert solver.use_abmc is False
    sol = np.zeros(s.A.n)
    res = solver.Solve(s.b, sol)
    xo, ro = oracle.solve_iccg(s.A, s.b, oracle.config())
    assert res.converged and abs(res.iterations - ro["iterations"]) <= it_tol(ro["iterations"])
    assert rel(sol, xo) < 1e-9


def test_sparsesolv_solver_divergence_stop_and_printrates(capsys):
    # SparseSolvSolver with divergence_check (StagnationCount) stops where
    # the reference does (golden cc6_divergence: 31 iterations, residual
    # >= 2x the best for 5 iterations running, best x restored), and
    # printrates writes the reference's lines (sparsesolv_solvers.hpp:140-155)
    import sparsesolv_amd as ss
    g = load_golden("cc6_divergence")
    s = P.curlcurl(6)
    solver = ss.SparseSolvSolver(s.A, method="ICCG", use_abmc=True, printrates=True)
    solver.auto_shift = True
    solver.divergence_check = True
    solver.divergence_threshold = 2.0
    solver.divergence_count = 5
    sol = np.zeros(s.A.n)
    res = solver.Solve(s.b, sol)
    assert not res.converged and not bool(g["converged"])
    assert abs(res.iterations - int(g["iterations"])) <= it_tol(int(g["iterations"]))
    # near-singular: by iteration ~25 summation order moves the residuals by
    # a few per cent (the first ones agree to 1e-8)
    np.testing.assert_allclose(res.residual_history[:5], g["history"][:5], rtol=1e-8)
    np.testing.assert_allclose(res.final_residual, float(g["final_residual"]), rtol=5e-2)
    r = np.linalg.norm(s.b - s.A.to_scipy() @ sol) / np.linalg.norm(s.b)
    assert abs(r - res.final_residual) <= 1e-8 * r
    lines = capsys.readouterr().out.splitlines()
    assert len(lines) == len(res.residual_history) + 1
    assert lines[0] == f"ICCG iteration 0, residual = {g['history'][0]:g}"
    assert lines[3] == f"ICCG iteration 3, residual = {res.residual_history[3]:g}"
    assert lines[-1] == (f"ICCG NOT converged after {res.iterations} iterations, "
                         f"residual = {res.final_residual:g}")
