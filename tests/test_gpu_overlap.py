"""The straggler walks on a second stream (walk_overlap) against the serial
sweeps: the same kernels on the same data in another order, so every
result is bit for bit the serial one.  walk_overlap = 1 takes the overlap
wherever the schedule allows it (these small systems: one colour with a
straggler block, which no later forward colour reads); the default (-1)
takes it only when the walks are long, as on the 64^3 curl-curl."""
import numpy as np
import pytest

from sparsesolv_amd import _lib
from sparsesolv_amd import problems as P
from test_gpu_parity import complex_shifted_laplace

pytestmark = pytest.mark.gpu


def pair(A, **kw):
    on = _lib.IcHandle(A, _lib.default_config(walk_overlap=1, **kw))
    off = _lib.IcHandle(A, _lib.default_config(walk_overlap=0, **kw))
    return on, off


CASES = [("cc16", lambda: P.curlcurl(16).A, dict(use_abmc=1, auto_shift=1)),
         ("lap24", lambda: P.laplace7(24), dict(use_abmc=1)),
         ("lap24_b8", lambda: P.laplace7(24), dict(use_abmc=1, abmc_block_size=8, abmc_num_colors=6))]


@pytest.mark.parametrize("name,mk,kw", CASES, ids=[c[0] for c in CASES])
def test_apply_overlapped_equals_serial(name, mk, kw):
    A = mk()
    on, off = pair(A, **kw)
    x = P.uniform(3, A.n, -1.0, 1.0)
    y_on, y_off = on.apply(x), off.apply(x)
    assert on.info()["walk_overlap"] == 1 and off.info()["walk_overlap"] == 0
    assert np.array_equal(y_on, y_off)
    # repeated applies on the same handle: the side stream's work of one
    # apply is joined before the next one's
    for k in range(3):
        x = P.uniform(10 + k, A.n, -1.0, 1.0)
        assert np.array_equal(on.apply(x), off.apply(x))


@pytest.mark.parametrize("use_graph", [1, 0])
@pytest.mark.parametrize("precision", [_lib.PRECISION_FP64, _lib.PRECISION_MIXED])
def test_iccg_overlapped_equals_serial(use_graph, precision):
    A = P.curlcurl(16).A
    b = P.uniform(4, A.n, -1.0, 1.0)
    kw = dict(use_abmc=1, auto_shift=1, max_iterations=400, use_graph=use_graph, precision=precision)
    on, off = pair(A, **kw)
    x1, r1, _ = on.solve(_lib.METHOD_ICCG, b, None, _lib.default_config(walk_overlap=1, **kw))
    x0, r0, _ = off.solve(_lib.METHOD_ICCG, b, None, _lib.default_config(walk_overlap=0, **kw))
    assert on.info()["walk_overlap"] == 1
    assert r1.iterations == r0.iterations and r1.status == r0.status
    assert np.array_equal(x1, x0)


def test_cocr_overlapped_equals_serial():
    A, b = complex_shifted_laplace(24)
    kw = dict(use_abmc=1, abmc_num_colors=8)
    on, off = pair(A, **kw)
    x1, r1, _ = on.solve(_lib.METHOD_ICCOCR, b, None, _lib.default_config(walk_overlap=1, **kw))
    x0, r0, _ = off.solve(_lib.METHOD_ICCOCR, b, None, _lib.default_config(walk_overlap=0, **kw))
    assert on.info()["walk_overlap"] == 1
    assert r1.converged and r1.iterations == r0.iterations
    assert np.array_equal(x1, x0)


def test_blocks_and_scaling_overlapped_equal_serial():
    A = P.laplace7(24)
    B = np.stack([P.uniform(30 + k, A.n, -1.0, 1.0) for k in range(5)])
    kw = dict(use_abmc=1, abmc_reorder_spmv=1, diagonal_scaling=1)
    on, off = pair(A, **kw)
    X1, res1 = on.solve_multi(_lib.METHOD_ICCG, B, None, _lib.default_config(walk_overlap=1, **kw))
    X0, res0 = off.solve_multi(_lib.METHOD_ICCG, B, None, _lib.default_config(walk_overlap=0, **kw))
    assert on.info()["walk_overlap"] == 1
    assert np.array_equal(X1, X0)
    assert [r.iterations for r in res1] == [r.iterations for r in res0]


def test_default_takes_the_overlap_for_long_walks_only():
    # 64^3 curl-curl: a 4,609-row straggler block, ~130 steps per walk
    A = P.curlcurl(64).A
    h = _lib.IcHandle(A, _lib.default_config(use_abmc=1, auto_shift=1))
    assert h.info()["walk_overlap"] == -1
    x = P.uniform(5, A.n, -1.0, 1.0)
    y = h.apply(x)
    assert h.info()["walk_overlap"] == 1
    y0 = _lib.IcHandle(A, _lib.default_config(use_abmc=1, auto_shift=1, walk_overlap=0)).apply(x)
    assert np.array_equal(y, y0)
    # 24^3 Laplacian: an 89-row block, short walks
    s = _lib.IcHandle(P.laplace7(24), _lib.default_config(use_abmc=1))
    s.apply(np.ones(24 ** 3))
    assert s.info()["walk_overlap"] == 0


def test_batch_overlapped_equals_serial():
    # two systems with the same pattern put their straggler blocks in the
    # same colour: one walk launch per direction holds both, on the side
    # stream; a 16^3 curl-curl beside them puts another straggler in another
    # colour, and the batch stays serial
    mats = [P.laplace7_varcoef(24, 10.0 ** P.uniform(s, 24 ** 3, -1.0, 1.0)) for s in (1, 2)]
    bs = [P.uniform(40 + k, A.n, -1.0, 1.0) for k, A in enumerate(mats)]
    for extra, expect in (([], 1), ([P.curlcurl(16).A], 0)):
        ms = mats + extra
        b = np.concatenate(bs + [P.uniform(50, A.n, -1.0, 1.0) for A in extra])
        res = {}
        for ov in (1, 0):
            cfg = _lib.default_config(use_abmc=1, auto_shift=1, walk_overlap=ov)
            h = _lib.BatchHandle(ms, cfg)
            x, results = h.solve_batch(_lib.METHOD_ICCG, b, None, cfg)
            res[ov] = (x, [r.iterations for r in results], h.info()["walk_overlap"])
        assert res[1][2] == expect and res[0][2] == 0
        assert np.array_equal(res[1][0], res[0][0]) and res[1][1] == res[0][1]
