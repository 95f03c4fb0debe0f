"""In-solve kernel timing (ssa_ic_trace): the timed solves give the same
results as untimed ones, and every launch of the selected class inside the
iteration loop is counted with its algorithmic bytes."""
import numpy as np
import pytest

from sparsesolv_amd import _lib
from sparsesolv_amd import problems as P

pytestmark = pytest.mark.gpu

CLASSES = (_lib.TRACE_SPMV, _lib.TRACE_COLOUR, _lib.TRACE_UPDATE, _lib.TRACE_PQ, _lib.TRACE_DOT)


@pytest.mark.parametrize("case", ["iccg", "cocr-block"])
def test_trace_keeps_results_and_counts_launches(case):
    # sizes with colours large enough to run as colour kernels of their own
    # (smaller colours share the small-colour kernel)
    if case == "iccg":
        s = P.poisson(48)
        A, method, nrhs = s.A, _lib.METHOD_ICCG, 1
        B = s.b[None, :]
        cfg = _lib.default_config(use_abmc=1, max_iterations=300)
    else:
        A = P.eddy(16, cond_frac=1.0, sigma=0.1).A
        method, nrhs = _lib.METHOD_ICCOCR, 4
        B = P.complex_gaussian_block(A.n, nrhs, 2)
        cfg = _lib.default_config(use_abmc=1, abmc_num_colors=8, max_iterations=300,
                                  precision=_lib.PRECISION_MIXED)
    h = _lib.IcHandle(A, cfg)
    X0, r0 = h.solve_multi(method, B, None, cfg)
    its = max(r.iterations for r in r0)
    for cls in CLASSES:
        h.trace(cls)
        X, r = h.solve_multi(method, B, None, cfg)
        k, us, by = h.trace_read()
        h.trace(_lib.TRACE_OFF)
        assert np.array_equal(X, X0), cls
        assert [q.iterations for q in r] == [q.iterations for q in r0], cls
        # one launch per iteration (colours: several); iterations enqueued
        # past the stop launch too
        assert k >= its, (cls, k, its)
        assert us > 0 and by > 0, cls
        k2, us2, by2 = h.trace_read()
        assert (k2, us2, by2) == (k, us, by)
    # off: nothing more is counted
    h.solve_multi(method, B, None, cfg)
    assert h.trace_read()[0] == k


def test_trace_rejects_unknown_class():
    s = P.poisson(6)
    h = _lib.IcHandle(s.A, _lib.default_config(use_abmc=1))
    with pytest.raises(_lib.SsaError):
        h.trace(17)
