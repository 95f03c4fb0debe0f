"""The oracle on tiny and ragged systems, bit for bit against the
reference's headers compiled here (oracle/_ref): one and two rows, and row
counts on either side of the 64-row slice and the 256-row workgroup, for
ICCG on the ABMC and level paths, COCR, SGS and SGS-MRTR; and the
pivot clamp on a system whose pivots vanish.  (An empty
system is outside what the reference accepts: its SparseMatrixView asserts
non-null column indices.)"""
import numpy as np
import pytest

from tiny_systems import GRIDS, grid_laplacian, rhs

CASES = sorted(GRIDS, key=lambda k: int(k[1:]))


def same(oracle, ref, fn, A, b, **kw):
    kw = dict(dict(save_residual_history=1, max_iterations=500), **kw)
    xo, ro = getattr(oracle, fn)(A, b, oracle.config(**kw))
    xr, rr = getattr(ref, fn)(A, b, ref.config(**kw))
    assert ro["iterations"] == rr["iterations"] and ro["converged"] == rr["converged"], (fn, kw)
    np.testing.assert_array_equal(ro["residual_history"], rr["residual_history"])
    np.testing.assert_array_equal(xo, xr)
    return ro


@pytest.mark.parametrize("name", CASES)
def test_tiny_systems_match_reference_build(oracle, ref, name):
    a, b_ = GRIDS[name]
    A = grid_laplacian(a, b_)
    b = rhs(A.n, 5)
    for abmc in (1, 0):
        r = same(oracle, ref, "solve_iccg", A, b, use_abmc=abmc)
        assert r["converged"]
    same(oracle, ref, "solve_iccg", A, b, use_abmc=1, abmc_block_size=8, abmc_num_colors=6)
    same(oracle, ref, "solve_sgsmrtr", A, b)
    Z = grid_laplacian(a, b_, complex_=True)
    z = rhs(Z.n, 6, complex_=True)
    r = same(oracle, ref, "solve_cocr", Z, z, use_abmc=1, abmc_num_colors=8)
    assert r["converged"]
    np.testing.assert_array_equal(oracle.sgs_apply(Z, z), ref.sgs_apply(Z, z))


def test_pivot_clamp_matches_reference_build(oracle, ref):
    # uncoupled 2x2 blocks [[1, 1], [1, 1]] at shift 1.0 without the
    # auto-shift: every second pivot is 1 - 1 = 0, clamped to
    # zero_diagonal_replacement (ic_preconditioner.hpp:1261-1267)
    import scipy.sparse as sp
    from sparsesolv_amd import problems as P
    A = P.CSR.from_scipy(sp.kron(sp.identity(300), np.array([[1.0, 1.0], [1.0, 1.0]])).tocsr())
    for abmc in (1, 0):
        fo = oracle.ic(A, oracle.config(use_abmc=abmc, shift_parameter=1.0))
        fr = ref.ic(A, ref.config(use_abmc=abmc, shift_parameter=1.0))
        a, b = fo.factor(), fr.factor()
        for k in a:
            np.testing.assert_array_equal(a[k], b[k])
        d = a["L_values"][a["L_row_ptr"][1:] - 1]
        assert np.sum(d == 1e-10) == 300
        x = rhs(A.n, 7)
        np.testing.assert_array_equal(fo.apply(x), fr.apply(x))
