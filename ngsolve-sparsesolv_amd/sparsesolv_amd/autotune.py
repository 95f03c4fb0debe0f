"""Autotuner for the ABMC parameters of the GPU path.

The ABMC block size and colour count change both the preconditioner (so the
iteration count) and the launch shape (so the time per iteration).  The
tuner solves a representative system once per candidate on the GPU and keeps
the candidate with the shortest time to tolerance; setup time is reported
but not part of the choice (a factor is built once and used for many
solves).  No reference counterpart: the reference takes abmc_block_size and
abmc_num_colors from the caller.
"""
from __future__ import annotations

import itertools
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib

BLOCK_SIZES = (2, 4, 8, 16)
COLOURS = (4, 8, 16)


@dataclass
class Trial:
    block_size: int
    num_colors: int
    seconds: float          # time to tolerance (solve only)
    iterations: int
    converged: bool
    setup_seconds: float


@dataclass
class TuneResult:
    best: Trial
    trials: list = field(default_factory=list)

    def config(self, **base):
        """Keyword arguments for default_config with the chosen parameters."""
        return dict(base, use_abmc=1, abmc_block_size=self.best.block_size,
                    abmc_num_colors=self.best.num_colors)


def choose(trials):
    """The converged trial with the shortest solve; ties go to fewer
    iterations, then to the earlier candidate."""
    ok = [t for t in trials if t.converged]
    if not ok:
        raise RuntimeError("autotune: no candidate reached the tolerance")
    return min(ok, key=lambda t: (t.seconds, t.iterations))


def candidates(block_sizes=BLOCK_SIZES, colours=COLOURS):
    return list(itertools.product(block_sizes, colours))


def autotune(A, b, method: int = _lib.METHOD_ICCG, block_sizes=BLOCK_SIZES, colours=COLOURS,
             reps: int = 2, evaluate=None, **cfg_kw) -> TuneResult:
    """Time every (block size, colours) candidate on (A, b).

    `evaluate(A, b, method, cfg, reps) -> (seconds, iterations, converged,
    setup)` runs one candidate; the default builds a handle and times `reps`
    solves, keeping the fastest."""
    evaluate = evaluate or _evaluate_gpu
    trials = []
    for bs, nc in candidates(block_sizes, colours):
        cfg = _lib.default_config(**dict(cfg_kw, use_abmc=1, abmc_block_size=bs,
                                         abmc_num_colors=nc))
        sec, it, conv, setup = evaluate(A, b, method, cfg, reps)
        trials.append(Trial(bs, nc, sec, it, conv, setup))
    return TuneResult(choose(trials), trials)


def _evaluate_gpu(A, b, method, cfg, reps):
    # host vectors: two transfers of n values per solve, small against the
    # solve, and no second GPU runtime in the process
    t = time.perf_counter()
    h = _lib.IcHandle(A, cfg)
    setup = time.perf_counter() - t
    best, res = None, None
    for _ in range(max(reps, 1)):
        t = time.perf_counter()
        _, res, _ = h.solve(method, b, None, cfg)
        s = time.perf_counter() - t
        best = s if best is None else min(best, s)
    h.close()
    return best, int(res.iterations), bool(res.converged), setup
