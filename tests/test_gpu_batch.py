"""Batched solves: several independent systems in one handle.

Each system keeps the launch shape of a one-system solve (the same
workgroups over its rows, the same reduction trees), so a batched solve must
reproduce every system's own solve bit for bit, while the systems stop at
different iterations (per-system stop status)."""
import numpy as np
import pytest

from sparsesolv_amd import _lib
from sparsesolv_amd import problems as P

pytestmark = pytest.mark.gpu


def varcoef(m, seed):
    coef = 10.0 ** P.uniform(seed, m ** 3, -2.0, 2.0)
    return P.laplace7_varcoef(m, coef), P.uniform(seed + 100, m ** 3, -1.0, 1.0)


def real_batch():
    out = []
    for k, m in enumerate((10, 7, 13, 5)):
        A, b = varcoef(m, 30 + k)
        out.append((A, b))
    s = P.poisson(9)
    out.append((s.A, s.b))
    return out


def check_against_singles(systems, cfg, method):
    hb = _lib.BatchHandle([A for A, _ in systems], cfg)
    assert hb.systems() == len(systems)
    b = np.concatenate([b for _, b in systems])
    x, results = hb.solve_batch(method, b, None, cfg)
    iters = []
    for k, (A, bk) in enumerate(systems):
        h = _lib.IcHandle(A, cfg)
        xk, rk, _ = h.solve(method, bk, None, cfg)
        seg = hb.split(x)[k]
        assert np.array_equal(seg, xk), k
        assert results[k].iterations == rk.iterations, k
        assert results[k].converged == rk.converged, k
        assert results[k].status == rk.status, k
        assert results[k].final_residual == rk.final_residual, k
        iters.append(rk.iterations)
    return iters


def test_batch_iccg_matches_single_solves():
    cfg = _lib.default_config(use_abmc=1, max_iterations=800)
    iters = check_against_singles(real_batch(), cfg, _lib.METHOD_ICCG)
    assert len(set(iters)) > 1          # the systems stop at different iterations


def test_batch_graph_and_eager_agree():
    systems = real_batch()
    b = np.concatenate([b for _, b in systems])
    out = []
    for graph in (1, 0):
        cfg = _lib.default_config(use_abmc=1, use_graph=graph)
        x, res = _lib.BatchHandle([A for A, _ in systems], cfg).solve_batch(_lib.METHOD_ICCG, b, None, cfg)
        out.append((x, [r.iterations for r in res]))
    assert np.array_equal(out[0][0], out[1][0]) and out[0][1] == out[1][1]


def test_batch_auto_shift_curlcurl():
    systems = [(P.curlcurl(m).A, P.curlcurl(m).b) for m in (5, 6, 4)]
    cfg = _lib.default_config(use_abmc=1, auto_shift=1, max_iterations=300)
    check_against_singles(systems, cfg, _lib.METHOD_ICCG)


def test_batch_cocr_complex():
    systems = []
    for m in (6, 5, 7):
        s = P.eddy(m)
        systems.append((s.A, s.b))
    cfg = _lib.default_config(use_abmc=1, abmc_num_colors=8, max_iterations=400)
    check_against_singles(systems, cfg, _lib.METHOD_ICCOCR)


def test_batch_level_path():
    cfg = _lib.default_config(use_abmc=0, max_iterations=800)
    check_against_singles(real_batch()[:3], cfg, _lib.METHOD_ICCG)


def test_batch_block8_and_plain_cg():
    systems = real_batch()[:3]
    check_against_singles(systems, _lib.default_config(use_abmc=1, abmc_block_size=8,
                                                       abmc_num_colors=6), _lib.METHOD_ICCG)
    check_against_singles(systems, _lib.default_config(use_abmc=1), _lib.METHOD_CG)


def test_batch_apply_and_spmv(oracle):
    systems = real_batch()
    hb = _lib.BatchHandle([A for A, _ in systems], _lib.default_config(use_abmc=1))
    x = P.uniform(9, hb.n, -1.0, 1.0)
    y = hb.apply(x)
    Ax = hb.spmv(x)
    for k, (A, _) in enumerate(systems):
        xk = hb.split(x)[k]
        yo = oracle.ic(A, oracle.config(use_abmc=1)).apply(xk)
        assert np.linalg.norm(hb.split(y)[k] - yo) <= 1e-12 * np.linalg.norm(yo)
        np.testing.assert_allclose(hb.split(Ax)[k], A.to_scipy() @ xk, rtol=1e-12, atol=1e-12)


def test_batch_handle_rejects_single_system_calls():
    systems = real_batch()[:2]
    cfg = _lib.default_config(use_abmc=1)
    hb = _lib.BatchHandle([A for A, _ in systems], cfg)
    b = np.zeros(hb.n)
    with pytest.raises(_lib.SsaError):
        hb.solve(_lib.METHOD_ICCG, b, None, cfg)
    with pytest.raises(_lib.SsaError):
        hb.ordering()


def test_batch_mixed_precision_matches_single_solves():
    # per-system iterative refinement: each system restarts, stops improving
    # and converges on its own, as in its own mixed-precision solve
    systems = real_batch()[:4]
    cfg = _lib.default_config(use_abmc=1, precision=_lib.PRECISION_MIXED, max_iterations=1500)
    check_against_singles(systems, cfg, _lib.METHOD_ICCG)
    cplx = [(P.eddy(m).A, P.eddy(m).b) for m in (5, 6)]
    check_against_singles(cplx, _lib.default_config(use_abmc=1, precision=_lib.PRECISION_MIXED,
                                                    abmc_num_colors=8, max_iterations=1500),
                          _lib.METHOD_ICCOCR)


def test_autotune_on_gpu():
    from sparsesolv_amd import autotune as at
    A, b = varcoef(12, 40)
    # block sizes x colours 4, 8, then the colour counts within 3 of the best
    r = at.autotune(A, b, block_sizes=(2, 4), colours=tuple(range(4, 10)), reps=1, max_iterations=1000)
    stage1 = [(t.block_size, t.num_colors) for t in r.trials[:4]]
    assert stage1 == [(2, 4), (2, 8), (4, 4), (4, 8)]
    assert len(r.trials) > 4 and all(t.converged for t in r.trials)
    assert r.best.seconds == min(t.seconds for t in r.trials)
    cfg = _lib.default_config(**r.config(max_iterations=1000))
    x, res, _ = _lib.IcHandle(A, cfg).solve(_lib.METHOD_ICCG, b, None, cfg)
    assert res.converged and np.linalg.norm(b - A.to_scipy() @ x) < 1e-9 * np.linalg.norm(b)


def test_width_tuning_keeps_the_results():
    # the right-hand sides per launch change only the launch shape: every
    # width gives the same bits
    from sparsesolv_amd import autotune as at
    A, _ = varcoef(10, 41)
    B = np.stack([P.uniform(50 + k, A.n, -1.0, 1.0) for k in range(10)])
    kw = dict(use_abmc=1, max_iterations=1000)
    h = _lib.IcHandle(A, _lib.default_config(**kw))
    w = at.autotune_width(h, B, kw, method=_lib.METHOD_ICCG)
    assert set(w.seconds) == {4, 8, 16} and w.width in w.seconds
    X = {}
    for width in (1, 4, 8, 16):
        X[width], res = h.solve_multi(_lib.METHOD_ICCG, B, None, _lib.default_config(rhs_block=width, **kw))
        assert all(r.converged for r in res)
    for width in (4, 8, 16):
        np.testing.assert_array_equal(X[width], X[1])
    with pytest.raises(_lib.SsaError):
        h.solve_multi(_lib.METHOD_ICCG, B, None, _lib.default_config(rhs_block=3, **kw))


def test_batch_tuned_per_size_class():
    # per-system block sizes and colour counts (ssa_batch_create_each): each
    # system's batched solve is its own solve with its parameters, bit for bit
    from sparsesolv_amd import autotune as at
    from types import SimpleNamespace
    systems = [SimpleNamespace(A=A, b=b) for A, b in real_batch()]
    kw = dict(use_abmc=1, max_iterations=1000)
    t = at.autotune_batch(systems, classes=2, block_sizes=(2, 4), colours=(4, 6, 8), refine=0, **kw)
    assert len(t.classes) == 2
    pairs = list(zip(t.block_sizes, t.num_colors))
    cfg = _lib.default_config(**kw)
    hb = _lib.BatchHandle([s.A for s in systems], cfg, block_sizes=t.block_sizes, num_colors=t.num_colors)
    b = np.concatenate([s.b for s in systems])
    x, results = hb.solve_batch(_lib.METHOD_ICCG, b, None, cfg)
    for k, s in enumerate(systems):
        bs, nc = pairs[k]
        c1 = _lib.default_config(abmc_block_size=bs, abmc_num_colors=nc, **kw)
        xk, rk, _ = _lib.IcHandle(s.A, c1).solve(_lib.METHOD_ICCG, s.b, None, c1)
        assert results[k].converged and results[k].iterations == rk.iterations
        np.testing.assert_array_equal(hb.split(x)[k], xk)
