"""The C oracle against the reference's own outputs.

tests/golden/*.npz were written by tests/golden/make_golden.py from the
reference headers compiled unchanged (oracle/_ref).  The oracle restates the
serial reference, so integer outputs and real factors must agree bit for bit,
complex ones too (both use libgcc's complex helpers); solver histories are
compared exactly as well.
"""
import numpy as np
import pytest

from conftest import load_golden
from sparsesolv_amd import problems as P

import make_golden as G


@pytest.mark.parametrize("name", sorted(G.CASES))
def test_oracle_matches_reference_golden(oracle, name):
    g = load_golden(name)
    build, kw, solver = G.CASES[name]
    sysm = build()
    A, b = sysm.A, sysm.b
    assert A.n == int(g["n"]) and A.nnz == int(g["nnz"])
    np.testing.assert_allclose(np.sum(A.values), g["values_sum"][0], rtol=1e-14)
    cfg = oracle.config(save_residual_history=1, **kw)
    if solver in ("iccg", "cocr"):
        f = oracle.ic(A, cfg).factor()
        for k in ("L_row_ptr", "L_col_idx"):
            np.testing.assert_array_equal(f[k], g[k])
        np.testing.assert_array_equal(f["L_values"], g["L_values"])
        np.testing.assert_array_equal(f["inv_diag"], g["inv_diag"])
        assert f["actual_shift"] == float(g["actual_shift"])
        if "ordering" in g:
            for k in ("ordering", "color_offsets", "block_offsets"):
                np.testing.assert_array_equal(f[k], g[k])
        y = oracle.ic(A, cfg).apply(g["apply_x"])
        np.testing.assert_array_equal(y, g["apply_y"])
    if solver.startswith("sgs"):
        np.testing.assert_array_equal(oracle.sgs_apply(A, g["apply_x"]), g["apply_y"])
    if solver == "iccg":
        x, r = oracle.solve_iccg(A, b, cfg)
    elif solver == "cocr":
        x, r = oracle.solve_cocr(A, b, cfg)
    elif solver == "sgsmrtr":
        x, r = oracle.solve_sgsmrtr(A, b, cfg)
    elif solver in ("sgs_cg", "sgs_cocr"):
        x, r = oracle.solve_sgs(A, b, cfg, "COCR" if solver == "sgs_cocr" else "CG")
    elif solver == "plain_cg":
        x, r = oracle.solve_plain(A, b, cfg, "CG")
    else:
        x, r = oracle.solve_plain(A, b, cfg, "COCR")
    assert r["iterations"] == int(g["iterations"])
    assert r["converged"] == bool(g["converged"])
    np.testing.assert_array_equal(r["residual_history"], g["history"])
    np.testing.assert_array_equal(x, g["x"])
    assert r["final_residual"] == float(g["final_residual"])


@pytest.mark.parametrize("m,B,nc", [(10, 4, 4), (16, 4, 8), (12, 8, 4), (12, 16, 16), (9, 1, 2), (7, 3, 5)])
def test_oracle_abmc_matches_reference_build(oracle, ref, m, B, nc):
    A = P.laplace7(m)
    a = oracle.abmc(A, B, nc)
    r = ref.abmc(A, B, nc)
    for k in a:
        np.testing.assert_array_equal(a[k], r[k])


@pytest.mark.parametrize("m,B,nc", [(6, 4, 4), (8, 4, 8), (7, 8, 4)])
def test_oracle_abmc_curlcurl_matches_reference_build(oracle, ref, m, B, nc):
    A = P.curlcurl(m).A
    a = oracle.abmc(A, B, nc)
    r = ref.abmc(A, B, nc)
    for k in a:
        np.testing.assert_array_equal(a[k], r[k])


@pytest.mark.parametrize("case", ["lap", "cc", "eddy"])
def test_oracle_solvers_match_reference_build(oracle, ref, case):
    if case == "lap":
        s = P.poisson(11)
        kw = dict(use_abmc=1, abmc_block_size=4, abmc_num_colors=4)
        fn = "solve_iccg"
    elif case == "cc":
        s = P.curlcurl(7)
        kw = dict(use_abmc=1, auto_shift=1, max_iterations=3000)
        fn = "solve_iccg"
    else:
        s = P.eddy(7)
        kw = dict(use_abmc=1, abmc_num_colors=8, max_iterations=3000)
        fn = "solve_cocr"
    xo, ro = getattr(oracle, fn)(s.A, s.b, oracle.config(save_residual_history=1, **kw))
    xr, rr = getattr(ref, fn)(s.A, s.b, ref.config(save_residual_history=1, **kw))
    assert ro["iterations"] == rr["iterations"]
    np.testing.assert_array_equal(ro["residual_history"], rr["residual_history"])
    np.testing.assert_array_equal(xo, xr)


@pytest.mark.parametrize("case", ["lap9", "zlap7", "eddy6", "cc6"])
def test_oracle_sgs_matches_reference_build(oracle, ref, case):
    # SGS apply, CG / COCR with SGS and SGS-MRTR, bit for bit, on further
    # systems (complex conjugated and unconjugated dots included)
    import scipy.sparse as sp
    if case == "lap9":
        s = P.poisson(9)
        A, b = s.A, s.b
    elif case == "zlap7":
        A = P.CSR.from_scipy((P.laplace7(7).to_scipy().astype(complex) + 0.4j * sp.identity(343)).tocsr())
        z = P.gaussian(4, 2 * 343)
        b = z[0::2] + 1j * z[1::2]
    else:
        s = P.eddy(6) if case == "eddy6" else P.curlcurl(6)
        A, b = s.A, s.b
    x = P.uniform(8, A.n, -1.0, 1.0) * (1.0 + (0.3j if A.is_complex else 0.0))
    np.testing.assert_array_equal(oracle.sgs_apply(A, x), ref.sgs_apply(A, x))
    variants = [dict(), dict(max_iterations=9), dict(divergence_check=1, divergence_threshold=1.5,
                                                     divergence_count=3)]
    if A.is_complex:
        variants.append(dict(conjugate=1, max_iterations=300))
    for kw in variants:
        kw = dict(dict(save_residual_history=1, max_iterations=3000), **kw)
        xo, ro = oracle.solve_sgsmrtr(A, b, oracle.config(**kw))
        xr, rr = ref.solve_sgsmrtr(A, b, ref.config(**kw))
        assert ro["iterations"] == rr["iterations"] and ro["converged"] == rr["converged"], kw
        np.testing.assert_array_equal(ro["residual_history"], rr["residual_history"])
        np.testing.assert_array_equal(xo, xr)
    for method in ("CG", "COCR") if A.is_complex else ("CG",):
        cfg = dict(save_residual_history=1, max_iterations=3000)
        xo, ro = oracle.solve_sgs(A, b, oracle.config(**cfg), method)
        xr, rr = ref.solve_sgs(A, b, ref.config(**cfg), method)
        assert ro["iterations"] == rr["iterations"]
        np.testing.assert_array_equal(xo, xr)


@pytest.mark.parametrize("norm_type,custom", [(1, 1.0), (2, 3.5), (2, 1e-3)])
@pytest.mark.parametrize("case", ["lap", "zlap", "eddy"])
def test_oracle_norm_types_match_reference_build(oracle, ref, case, norm_type, custom):
    # NormType::InitialResidual / Custom (core/types.hpp:42-46,
    # iterative_solver.hpp:155-160) with a nonzero initial guess, so that
    # ||r_0|| differs from ||b||
    if case == "lap":
        s = P.poisson(9)
    elif case == "eddy":
        s = P.eddy(5)
    else:
        import scipy.sparse as sp
        A = P.laplace7(9).to_scipy().astype(complex) + sp.identity(9 ** 3) * 0.3j
        z = P.gaussian(6, 2 * 9 ** 3)
        s = P.System(P.CSR.from_scipy(A), z[0::2] + 1j * z[1::2], "zlap9", {})
    fn = "solve_iccg" if case == "lap" else "solve_cocr"
    kw = dict(use_abmc=1, abmc_num_colors=4 if case == "lap" else 8, norm_type=norm_type,
              custom_norm=custom, save_residual_history=1, max_iterations=2000)
    x0 = P.uniform(21, s.A.n, -1.0, 1.0)
    if s.A.is_complex:
        x0 = x0 + 1j * P.uniform(22, s.A.n, -1.0, 1.0)
    xo, ro = getattr(oracle, fn)(s.A, s.b, oracle.config(**kw), x0=x0)
    xr, rr = getattr(ref, fn)(s.A, s.b, ref.config(**kw), x0=x0)
    assert ro["iterations"] == rr["iterations"] and ro["converged"] == rr["converged"]
    np.testing.assert_array_equal(ro["residual_history"], rr["residual_history"])
    np.testing.assert_array_equal(xo, xr)


@pytest.mark.parametrize("case", ["cc6_params", "pairs_params", "pairs_repl"])
def test_oracle_shift_settings_match_reference_build(oracle, ref, case):
    # auto-shift increment, cap, pivot threshold, trial limit and the clamp
    # value away from their defaults (core/solver_config.hpp:81-100,
    # ic_preconditioner.hpp:1377-1412): factor and solve bit for bit
    import scipy.sparse as sp
    if case == "cc6_params":
        s = P.curlcurl(6)
        A, b = s.A, s.b
        kw = dict(auto_shift=1, shift_parameter=0.5, shift_increment=0.2, max_shift_value=3.0,
                  min_diagonal_threshold=1e-3, max_shift_trials=3)
    else:
        A = P.CSR.from_scipy(sp.kron(sp.identity(300), np.array([[1.0, 1.0], [1.0, 1.0]])).tocsr())
        b = P.uniform(3, A.n, -1.0, 1.0)
        kw = (dict(shift_parameter=1.0, auto_shift=1, shift_increment=0.01, max_shift_value=1.02,
                   max_shift_trials=10) if case == "pairs_params"
              else dict(shift_parameter=1.0, zero_diagonal_replacement=1e-4))
    kw = dict(use_abmc=1, save_residual_history=1, max_iterations=200, **kw)
    fo = oracle.ic(A, oracle.config(**kw)).factor()
    fr = ref.ic(A, ref.config(**kw)).factor()
    for k in ("ordering", "L_values", "inv_diag"):
        np.testing.assert_array_equal(fo[k], fr[k])
    assert fo["actual_shift"] == fr["actual_shift"]
    xo, ro = oracle.solve_iccg(A, b, oracle.config(**kw))
    xr, rr = ref.solve_iccg(A, b, ref.config(**kw))
    assert ro["iterations"] == rr["iterations"] and ro["actual_shift"] == rr["actual_shift"]
    np.testing.assert_array_equal(ro["residual_history"], rr["residual_history"])
    np.testing.assert_array_equal(xo, xr)


@pytest.mark.parametrize("fn", ["solve_iccg", "solve_cocr"])
def test_oracle_conjugate_matches_reference_build(oracle, ref, fn):
    # conjugate=True: the Hermitian inner product in CG / COCR's dots
    # (iterative_solver.hpp:284-322) with the IC(0) preconditioner, on a
    # Hermitian positive definite system (Laplacian + a skew-symmetric
    # imaginary part) and on a complex-symmetric one
    import scipy.sparse as sp
    L = P.laplace7(7).to_scipy()
    S = sp.triu(L, 1) * 0.3j
    for A in (P.CSR.from_scipy((L.astype(complex) + S - S.T).tocsr()),
              P.CSR.from_scipy((L.astype(complex) + 0.4j * sp.identity(343)).tocsr())):
        z = P.gaussian(9, 2 * A.n)
        b = z[0::2] + 1j * z[1::2]
        kw = dict(use_abmc=1, conjugate=1, save_residual_history=1, max_iterations=400)
        xo, ro = getattr(oracle, fn)(A, b, oracle.config(**kw))
        xr, rr = getattr(ref, fn)(A, b, ref.config(**kw))
        assert ro["iterations"] == rr["iterations"] and ro["converged"] == rr["converged"]
        np.testing.assert_array_equal(ro["residual_history"], rr["residual_history"])
        np.testing.assert_array_equal(xo, xr)
