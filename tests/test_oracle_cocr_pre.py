"""COCR with the IC preconditioner set up on another matrix (the oracle's
solve_cocr_pre) pinned against the reference build, and against the
oracle's own COCR when the two matrices are one."""
import numpy as np
import scipy.sparse as sp

from sparsesolv_amd import problems as P
from test_gpu_parity import complex_shifted_laplace


def shifted(A, s):
    return P.CSR.from_scipy((A.to_scipy() + sp.identity(A.n) * s).tocsr())


def test_same_matrix_is_plain_cocr(oracle):
    A, b = complex_shifted_laplace(8)
    cfg = oracle.config(use_abmc=1, abmc_num_colors=8)
    x1, r1 = oracle.solve_cocr_pre(A, A, b, cfg)
    x0, r0 = oracle.solve_cocr(A, b, cfg)
    assert r1["iterations"] == r0["iterations"] and np.array_equal(x1, x0)


def test_other_preconditioner_matrix_matches_reference(oracle, ref):
    A, b = complex_shifted_laplace(8)
    for M, kw in ((shifted(A, 0.5), dict(use_abmc=1, abmc_num_colors=8)),
                  (shifted(A, 2.0j), dict(use_abmc=0, save_residual_history=1))):
        x0, r0 = ref.solve_cocr_pre(A, M, b, ref.config(**kw))
        x1, r1 = oracle.solve_cocr_pre(A, M, b, oracle.config(**kw))
        assert r0["converged"] and r1["iterations"] == r0["iterations"]
        assert np.linalg.norm(x1 - x0) <= 1e-12 * np.linalg.norm(x0)
        assert np.allclose(r1["residual_history"], r0["residual_history"], rtol=1e-9, atol=0)
        # a solution of A x = b, not of M x = b
        assert np.linalg.norm(A.to_scipy() @ x1 - b) <= 1e-9 * np.linalg.norm(b)
