"""Host half of ssa_ic_create (ABMC schedule + IC(0) factor, host C++)
against the oracle, through the C ABI with a host-only handle (no GPU)."""
import numpy as np
import pytest
import scipy.sparse as sp

from sparsesolv_amd import _lib
from sparsesolv_amd import problems as P


CASES = [
    ("lap9", lambda: P.laplace7(9), dict()),
    ("lap12_b8", lambda: P.laplace7(12), dict(abmc_block_size=8, abmc_num_colors=6)),
    ("lap10_b16", lambda: P.laplace7(10), dict(abmc_block_size=16, abmc_num_colors=16)),
    ("lap8_b3", lambda: P.laplace7(8), dict(abmc_block_size=3, abmc_num_colors=5)),
    ("cc6_auto", lambda: P.curlcurl(6).A, dict(auto_shift=1)),
    ("lap8_auto03", lambda: P.laplace7(8), dict(shift_parameter=0.3, auto_shift=1)),
    ("eddy6", lambda: P.eddy(6).A, dict(abmc_num_colors=8)),
    ("eddy5_b2", lambda: P.eddy(5).A, dict(abmc_block_size=2)),
    # colours of hundreds of blocks: the factor runs on host threads
    ("lap24", lambda: P.laplace7(24), dict()),
    ("lap24_auto03", lambda: P.laplace7(24), dict(shift_parameter=0.3, auto_shift=1)),
    ("cc16_auto", lambda: P.curlcurl(16).A, dict(auto_shift=1)),
    ("eddy16", lambda: P.eddy(16).A, dict(abmc_num_colors=8)),
    # 10,000 uncoupled 2x2 blocks [[1, 1], [1, 1]] at shift 1.0: every
    # second pivot is 1 - 1/1 = 0, so the auto-shift restarts (at 1.05)
    # after a colour of 5,000 blocks factored on host threads
    ("pairs_restart", lambda: P.CSR.from_scipy(
        sp.kron(sp.identity(10000), np.array([[1.0, 1.0], [1.0, 1.0]])).tocsr()),
     dict(shift_parameter=1.0, auto_shift=1)),
    # the same pairs without the auto-shift: every second pivot is 0 and is
    # clamped to zero_diagonal_replacement (ic_preconditioner.hpp:1261-1267)
    ("pairs_clamp", lambda: P.CSR.from_scipy(
        sp.kron(sp.identity(3000), np.array([[1.0, 1.0], [1.0, 1.0]])).tocsr()),
     dict(shift_parameter=1.0)),
    # the auto-shift's own settings away from their defaults
    # (core/solver_config.hpp:81-100): a larger first increment, a lower
    # shift cap, a looser pivot test and a trial limit that ends the search
    ("cc6_auto_params", lambda: P.curlcurl(6).A,
     dict(auto_shift=1, shift_parameter=0.5, shift_increment=0.2, max_shift_value=3.0,
          min_diagonal_threshold=1e-3, max_shift_trials=3)),
    ("pairs_auto_params", lambda: P.CSR.from_scipy(
        sp.kron(sp.identity(300), np.array([[1.0, 1.0], [1.0, 1.0]])).tocsr()),
     dict(shift_parameter=1.0, auto_shift=1, shift_increment=0.01, max_shift_value=1.02,
          max_shift_trials=10)),
    ("pairs_clamp_repl", lambda: P.CSR.from_scipy(
        sp.kron(sp.identity(300), np.array([[1.0, 1.0], [1.0, 1.0]])).tocsr()),
     dict(shift_parameter=1.0, zero_diagonal_replacement=1e-4)),
]


@pytest.mark.parametrize("name,builder,kw", CASES, ids=[c[0] for c in CASES])
def test_host_setup_bit_exact(oracle, name, builder, kw):
    A = builder()
    h = _lib.IcHandle(A, _lib.default_config(use_abmc=1, **kw), host_only=True)
    f = oracle.ic(A, oracle.config(use_abmc=1, **kw)).factor()
    np.testing.assert_array_equal(h.ordering(), f["ordering"])
    co, bo = h.schedule()
    np.testing.assert_array_equal(co, f["color_offsets"])
    np.testing.assert_array_equal(bo, f["block_offsets"])
    g = h.factor()
    np.testing.assert_array_equal(g["L_row_ptr"], f["L_row_ptr"])
    np.testing.assert_array_equal(g["L_col_idx"], f["L_col_idx"])
    np.testing.assert_array_equal(g["L_values"], f["L_values"])
    np.testing.assert_array_equal(g["inv_diag"], f["inv_diag"])
    assert g["actual_shift"] == f["actual_shift"]


def test_host_setup_missing_diagonal():
    import scipy.sparse as sp
    from sparsesolv_amd.problems import CSR
    A = sp.csr_matrix(np.array([[2.0, -1.0, 0.0], [-1.0, 0.0, -1.0], [0.0, -1.0, 2.0]]))
    A.eliminate_zeros()
    with pytest.raises(_lib.SsaError, match="missing diagonal") as e:
        _lib.IcHandle(CSR.from_scipy(A), _lib.default_config(use_abmc=1), host_only=True)
    assert e.value.code == _lib.SSA_ERR_RUNTIME


def test_host_setup_missing_diagonal_on_threads(oracle):
    # three missing diagonals in a 24^3 Laplacian: the colours run on host
    # threads, and the error names the first failing row of the serial
    # order, as the reference's message does
    A = P.laplace7(24).to_scipy().tolil()
    for r in (5000, 9000, 12000):
        A[r, r] = 0.0
    A = A.tocsr()
    A.eliminate_zeros()
    C = P.CSR.from_scipy(A)
    with pytest.raises(Exception) as eo:
        oracle.ic(C, oracle.config(use_abmc=1)).factor()
    with pytest.raises(_lib.SsaError) as e:
        _lib.IcHandle(C, _lib.default_config(use_abmc=1), host_only=True)
    assert "missing diagonal at row" in str(eo.value)
    assert str(e.value).endswith(str(eo.value).split("missing diagonal")[1])


def test_host_setup_rejects_bad_input():
    from sparsesolv_amd.problems import CSR
    A = P.laplace7(3)
    bad = CSR(A.n, A.row_ptr, A.col_idx.copy(), A.values)
    bad.col_idx[5] = A.n + 3
    with pytest.raises(_lib.SsaError) as e:
        _lib.IcHandle(bad, _lib.default_config(use_abmc=1), host_only=True)
    assert e.value.code == _lib.SSA_ERR_ARGUMENT


def test_host_only_handle_refuses_device_work():
    A = P.laplace7(4)
    h = _lib.IcHandle(A, _lib.default_config(use_abmc=1), host_only=True)
    with pytest.raises(_lib.SsaError) as e:
        h.apply(np.ones(A.n))
    assert e.value.code == _lib.SSA_ERR_ARGUMENT


def test_layout_info_reports_big_block():
    # 48^3 / B=4 leaves one 366-row straggler block with 27 dependency levels
    A = P.laplace7(48)
    h = _lib.IcHandle(A, _lib.default_config(use_abmc=1), host_only=True)
    co, bo = h.schedule()
    sizes = np.diff(bo)
    assert len(co) - 1 == 8 and sizes.max() == 366


LEVEL_CASES = [
    ("lap9", lambda: P.laplace7(9), dict()),
    ("cc6_auto", lambda: P.curlcurl(6).A, dict(auto_shift=1)),
    ("lap8_auto03", lambda: P.laplace7(8), dict(shift_parameter=0.3, auto_shift=1)),
    ("eddy6", lambda: P.eddy(6).A, dict()),
]


@pytest.mark.parametrize("name,builder,kw", LEVEL_CASES, ids=[c[0] for c in LEVEL_CASES])
def test_level_path_host_setup(oracle, name, builder, kw):
    # use_abmc = false: the factor is the natural-order one, bit for bit
    A = builder()
    h = _lib.IcHandle(A, _lib.default_config(use_abmc=0, **kw), host_only=True)
    f = oracle.ic(A, oracle.config(use_abmc=0, **kw)).factor()
    g = h.factor()
    for k in ("L_row_ptr", "L_col_idx", "L_values", "inv_diag"):
        np.testing.assert_array_equal(g[k], f[k], err_msg=k)
    assert g["actual_shift"] == f["actual_shift"]
    # the device order is the level order of core/level_schedule.hpp:58-89:
    # level(i) = 1 + max level(j) over the columns j < i of row i of L,
    # rows grouped by level, natural order inside a level
    rp, ci = f["L_row_ptr"], f["L_col_idx"]
    level = np.zeros(A.n, dtype=np.int64)
    for i in range(A.n):
        js = ci[rp[i]:rp[i + 1]]
        js = js[js < i]
        level[i] = level[js].max() + 1 if len(js) else 0
    expect = np.lexsort((np.arange(A.n), level))       # new -> old
    ordering = h.ordering()
    np.testing.assert_array_equal(np.argsort(ordering), expect)
    co, bo = h.schedule()
    np.testing.assert_array_equal(np.diff(co), np.bincount(level))
    np.testing.assert_array_equal(bo, np.arange(A.n + 1))


def test_batch_host_handle():
    # host-only batched handle: the setup of every system runs, and the
    # one-system queries are refused
    mats = [P.laplace7(6), P.laplace7(4), P.curlcurl(4).A]
    h = _lib.BatchHandle(mats, _lib.default_config(use_abmc=1, auto_shift=1), host_only=True)
    assert h.systems() == 3
    info = h.info()
    assert info["n"] == sum(A.n for A in mats)
    with pytest.raises(_lib.SsaError):
        h.factor()
    with pytest.raises(TypeError):
        _lib.BatchHandle([mats[0], P.eddy(4).A], _lib.default_config(use_abmc=1), host_only=True)


@pytest.mark.parametrize("use_abmc", [1, 0])
def test_host_setup_diagonal_scaling(oracle, use_abmc):
    # the factor of S A S, S = diag(1/sqrt|a_ii|), bit for bit
    m = 10
    A = P.laplace7_varcoef(m, 10.0 ** P.uniform(21, m ** 3, -2.0, 2.0))
    kw = dict(use_abmc=use_abmc, diagonal_scaling=1)
    h = _lib.IcHandle(A, _lib.default_config(**kw), host_only=True)
    f = oracle.ic(A, oracle.config(**kw)).factor()
    g = h.factor()
    for k in ("L_row_ptr", "L_col_idx", "L_values", "inv_diag"):
        np.testing.assert_array_equal(g[k], f[k], err_msg=k)


def test_batch_with_diagonal_scaling_builds():
    # each system planned with its own S (the batched solves are checked
    # against one-system solves on the GPU)
    cfg = _lib.default_config(use_abmc=1, diagonal_scaling=1)
    h = _lib.BatchHandle([P.laplace7(4), P.laplace7(5)], cfg, host_only=True)
    assert h.systems() == 2


RCM_CASES = [
    ("lap9", lambda: P.laplace7(9), dict()),
    ("lap24", lambda: P.laplace7(24), dict()),
    ("cc8_auto", lambda: P.curlcurl(8).A, dict(auto_shift=1)),
    ("eddy6", lambda: P.eddy(6).A, dict(abmc_num_colors=8)),
    ("varcoef10_b8", lambda: P.laplace7_varcoef(10, 10.0 ** P.uniform(21, 1000, -2.0, 2.0)),
     dict(abmc_block_size=8, abmc_num_colors=6)),
]


@pytest.mark.parametrize("name,builder,kw", RCM_CASES, ids=[c[0] for c in RCM_CASES])
def test_host_setup_rcm_matches_reference(ref, name, builder, kw):
    # abmc_use_rcm: ABMC of the RCM-reordered matrix.  The reference build
    # pins it (the C restatement has no RCM): its schedule and factor, bit
    # for bit; our ordering is the composite original -> RCM -> ABMC one
    A = builder()
    cfgk = dict(use_abmc=1, abmc_use_rcm=1, **kw)
    h = _lib.IcHandle(A, _lib.default_config(**cfgk), host_only=True)
    f = ref.ic(A, ref.config(**cfgk)).factor()
    co, bo = h.schedule()
    np.testing.assert_array_equal(co, f["color_offsets"])
    np.testing.assert_array_equal(bo, f["block_offsets"])
    g = h.factor()
    for k in ("L_row_ptr", "L_col_idx", "L_values", "inv_diag"):
        np.testing.assert_array_equal(g[k], f[k], err_msg=k)
    assert g["actual_shift"] == f["actual_shift"]
    # the composite ordering is a permutation and differs from plain ABMC's
    o = h.ordering()
    assert np.array_equal(np.sort(o), np.arange(A.n))
    plain = _lib.IcHandle(A, _lib.default_config(use_abmc=1, **kw), host_only=True).ordering()
    assert not np.array_equal(o, plain)
