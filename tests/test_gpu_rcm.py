"""abmc_use_rcm on the GPU against the reference build (oracle/_ref; the C
restatement has no RCM).  The factor and schedule are bit-exact on the host
(tests/test_host_setup.py); here the preconditioner's apply (through the
composite original -> RCM -> ABMC permutation) and the solves: ICCG runs the
reference's Path 2 (CG in the RCM space, SpMV on the RCM-ordered matrix),
which the GPU's ABMC-ordered loop matches to rounding, as it matches
Path 3."""
import math

import numpy as np
import pytest

from sparsesolv_amd import _lib
from sparsesolv_amd import problems as P
from test_gpu_parity import complex_shifted_laplace

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def it_tol(n_ref):
    return max(1, math.ceil(0.02 * n_ref))


def varcoef(m, seed=21):
    return P.laplace7_varcoef(m, 10.0 ** P.uniform(seed, m ** 3, -2.0, 2.0))


@pytest.mark.parametrize("name", ["lap12", "eddy6"])
def test_apply_with_rcm(ref, name):
    if name == "lap12":
        A, kw = P.laplace7(12), dict()
        x = P.uniform(5, A.n, -1.0, 1.0)
    else:
        A, kw = P.eddy(6).A, dict(abmc_num_colors=8)
        x = P.uniform(5, A.n, -1.0, 1.0) + 1j * P.uniform(6, A.n, -1.0, 1.0)
    kw.update(use_abmc=1, abmc_use_rcm=1)
    y = _lib.IcHandle(A, _lib.default_config(**kw)).apply(x)
    yo = ref.ic(A, ref.config(**kw)).apply(x)
    assert rel(y, yo) <= 1e-12


@pytest.mark.parametrize("scaling", [0, 1])
def test_iccg_with_rcm(ref, scaling):
    A = varcoef(10)
    b = P.uniform(3, A.n, -1.0, 1.0)
    kw = dict(use_abmc=1, abmc_use_rcm=1, diagonal_scaling=scaling, max_iterations=2000)
    cfg = _lib.default_config(**kw)
    x, res, _ = _lib.IcHandle(A, cfg).solve(_lib.METHOD_ICCG, b, None, cfg)
    xo, ro = ref.solve_iccg(A, b, ref.config(**kw))
    assert res.converged and ro["converged"]
    assert abs(res.iterations - ro["iterations"]) <= it_tol(ro["iterations"])
    assert rel(x, xo) < 1e-9


def test_cocr_with_rcm(ref):
    A, b = complex_shifted_laplace(10)
    kw = dict(use_abmc=1, abmc_use_rcm=1, abmc_num_colors=8)
    cfg = _lib.default_config(**kw)
    x, res, _ = _lib.IcHandle(A, cfg).solve(_lib.METHOD_ICCOCR, b, None, cfg)
    xo, ro = ref.solve_cocr(A, b, ref.config(**kw))
    assert res.converged and ro["converged"]
    assert abs(res.iterations - ro["iterations"]) <= it_tol(ro["iterations"])
    assert rel(x, xo) < 1e-9


def test_batch_and_blocks_with_rcm():
    cfg = _lib.default_config(use_abmc=1, abmc_use_rcm=1, auto_shift=1)
    mats = [varcoef(8, seed=5), P.curlcurl(5).A, P.laplace7(9)]
    bs = [P.uniform(40 + k, A.n, -1.0, 1.0) for k, A in enumerate(mats)]
    hb = _lib.BatchHandle(mats, cfg)
    x, results = hb.solve_batch(_lib.METHOD_ICCG, np.concatenate(bs), None, cfg)
    for k, (A, b, xk) in enumerate(zip(mats, bs, hb.split(x))):
        xs, rs, _ = _lib.IcHandle(A, cfg).solve(_lib.METHOD_ICCG, b, None, cfg)
        assert np.array_equal(xk, xs) and results[k].iterations == rs.iterations
    A = mats[0]
    B = np.stack([P.uniform(60 + k, A.n, -1.0, 1.0) for k in range(4)])
    h = _lib.IcHandle(A, cfg)
    X, res = h.solve_multi(_lib.METHOD_ICCG, B, None, cfg)
    for k in range(4):
        xk, rk, _ = h.solve(_lib.METHOD_ICCG, B[k], None, cfg)
        assert np.array_equal(X[k], xk) and res[k].iterations == rk.iterations
