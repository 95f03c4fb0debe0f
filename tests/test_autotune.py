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
    assert len(seen) == len(at.candidates()) and all(t == 1e-9 for *_, t in seen)
    kw = r.config(tolerance=1e-9)
    assert kw["abmc_block_size"] == 8 and kw["use_abmc"] == 1 and kw["tolerance"] == 1e-9
