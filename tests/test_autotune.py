"""Autotuner selection logic (CPU: a stand-in evaluator replaces the GPU)."""
import pytest

from sparsesolv_amd import autotune as at
from sparsesolv_amd import problems as P


def test_candidates_cover_the_grid():
    c = at.candidates()
    assert len(c) == len(at.BLOCK_SIZES) * len(at.COLOURS)
    assert (4, 4) in c and (16, 16) in c


def test_choose_prefers_fast_converged_trials():
    trials = [at.Trial(2, 4, 0.5, 300, True, 0.1), at.Trial(4, 4, 0.3, 310, True, 0.1),
              at.Trial(8, 8, 0.1, 2000, False, 0.1), at.Trial(16, 4, 0.3, 290, True, 0.1)]
    best = at.choose(trials)
    assert (best.block_size, best.num_colors) == (16, 4)     # tie on time, fewer iterations
    with pytest.raises(RuntimeError):
        at.choose([at.Trial(2, 4, 0.1, 10, False, 0.0)])


def test_autotune_with_stand_in_evaluator():
    s = P.poisson(6)
    seen = []

    def fake(A, b, method, cfg, reps):
        seen.append((cfg.abmc_block_size, cfg.abmc_num_colors, cfg.tolerance))
        # pretend bigger blocks are faster but 16 does not converge
        bs = cfg.abmc_block_size
        return 1.0 / bs, 100, bs != 16, 0.01

    r = at.autotune(s.A, s.b, evaluate=fake, tolerance=1e-9)
    assert (r.best.block_size, r.best.num_colors) == (8, 4)
    # stage 1: 4 block sizes x colours 4, 8, 12, 16; stage 2: colours 5-7 at block 8
    assert len(seen) == 16 + 3 and all(t == 1e-9 for *_, t in seen)
    assert [(b, c) for b, c, _ in seen[16:]] == [(8, 5), (8, 6), (8, 7)]
    kw = r.config(tolerance=1e-9)
    assert kw["abmc_block_size"] == 8 and kw["use_abmc"] == 1 and kw["tolerance"] == 1e-9


def test_refinement_finds_a_colour_count_between_the_coarse_ones():
    s = P.poisson(6)

    def fake(A, b, method, cfg, reps):
        # fastest at 4-row blocks and 10 colours
        t = abs(cfg.abmc_num_colors - 10) + (0 if cfg.abmc_block_size == 4 else 5)
        return 1.0 + t, 50, True, 0.0

    r = at.autotune(s.A, s.b, evaluate=fake)
    assert (r.best.block_size, r.best.num_colors) == (4, 10)
    full = at.autotune(s.A, s.b, evaluate=fake, refine=-1)
    assert len(full.trials) == len(at.candidates()) and full.best.num_colors == 10


def test_size_classes_and_batch_tuning():
    assert at.size_classes([50, 10, 30, 20, 40, 60], 3) == [[1, 3], [2, 4], [0, 5]]
    assert at.size_classes([5, 5], 3) == [[0], [1]]
    systems = [P.poisson(m) for m in (4, 9, 5, 8, 6, 7)]
    tuned_on = []

    def fake(A, b, method, cfg, reps):
        tuned_on.append(A.n)
        # small systems like many colours, large ones few
        want = 16 if A.n < 300 else 4
        return 1.0 + abs(cfg.abmc_num_colors - want) + cfg.abmc_block_size / 100, 10, True, 0.0

    r = at.autotune_batch(systems, classes=2, evaluate=fake)
    # classes {4, 5, 6}^3 and {7, 8, 9}^3, tuned on 6^3 = 216 and 9^3 = 729 rows
    assert sorted(set(tuned_on)) == [216, 729]
    assert [c[1] for c in r.classes] == [4, 1]
    assert r.num_colors == [16, 4, 16, 4, 16, 4] and r.block_sizes == [2] * 6


def test_width_choice():
    calls = []

    def fake(h, B, method, cfg, reps):
        calls.append(cfg.rhs_block)
        return {4: 3.0, 8: 2.5, 16: 2.5}[cfg.rhs_block]

    r = at.autotune_width(None, [[0.0]] * 16, {}, evaluate=fake)
    assert calls == [4, 8, 16] and r.width == 16      # a tie goes to the wider block
    calls.clear()
    r = at.autotune_width(None, [[0.0]] * 3, {}, evaluate=fake)
    assert calls == [4] and r.width == 4


def test_batch_handle_with_per_system_parameters_host_only():
    from sparsesolv_amd import _lib
    systems = [P.poisson(m).A for m in (5, 7, 6)]
    cfg = _lib.default_config(use_abmc=1)
    bs, nc = [2, 4, 8], [4, 8, 16]
    hb = _lib.BatchHandle(systems, cfg, host_only=True, block_sizes=bs, num_colors=nc)
    singles = [_lib.IcHandle(A, _lib.default_config(use_abmc=1, abmc_block_size=b, abmc_num_colors=c),
                             host_only=True) for A, b, c in zip(systems, bs, nc)]
    info = hb.info()
    assert info["num_blocks"] == sum(h.info()["num_blocks"] for h in singles)
    assert info["num_colors"] == max(h.info()["num_colors"] for h in singles)
    assert info["nnz_l"] == sum(h.info()["nnz_l"] for h in singles)
    with pytest.raises(ValueError):
        _lib.BatchHandle(systems, cfg, host_only=True, block_sizes=[4, 4])
    with pytest.raises(_lib.SsaError):
        _lib.BatchHandle(systems, cfg, host_only=True, num_colors=[4, 0, 4])
