"""Autotuner for the schedule parameters of the GPU path.

The ABMC block size and colour count change both the preconditioner (so the
iteration count) and the launch shape (so the time per iteration); the
right-hand sides per launch (1, 4, 8 or 16) change only the launch shape.
The tuner solves a representative system once per candidate on the GPU and
keeps the candidate with the shortest time to tolerance; setup time is
reported but not part of the choice (a factor is built once and used for
many solves).

  * autotune        block size x colour count on one system: the block sizes
                    against the colour counts of a coarse grid (4, 8, 12,
                    16), then every colour count within 3 of the best at its
                    block size;
  * autotune_width  right-hand sides per launch for a block solve;
  * autotune_batch  a batch of uneven systems: the systems split into size
                    classes, each class tuned on its largest system, and the
                    per-system choice for ssa_batch_create_each.

Rows per wavefront are not a separate dimension: every kernel gives a
wavefront 64 rows (SELL-64 slices of the SpMV, one row per lane in the
sweeps, the rows of an ABMC block in consecutive lanes), and what a
wavefront carries per row is the block width -- the right-hand sides per
lane that autotune_width chooses.

No reference counterpart: the reference takes abmc_block_size and
abmc_num_colors from the caller and solves one vector per call.
"""
from __future__ import annotations

import itertools
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib

BLOCK_SIZES = (2, 4, 8, 16)
COLOURS = tuple(range(4, 17))
WIDTHS = (4, 8, 16)


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
    """The full grid of (block size, colour count) pairs."""
    return list(itertools.product(block_sizes, colours))


def coarse_colours(colours=COLOURS):
    """The colour counts of the first stage: the multiples of 4 among
    `colours` (all of them when there are none)."""
    c = [k for k in colours if k % 4 == 0]
    return tuple(c) if c else tuple(colours)


def autotune(A, b, method: int = _lib.METHOD_ICCG, block_sizes=BLOCK_SIZES, colours=COLOURS,
             reps: int = 2, evaluate=None, refine: int = 3, **cfg_kw) -> TuneResult:
    """Choose (block size, colour count) on (A, b).

    Stage 1 times every block size against coarse_colours(colours); stage 2
    the remaining colour counts of `colours` within `refine` of the best
    stage-1 count, at the best stage-1 block size (refine=0: stage 1 only;
    refine < 0: the full grid).  `evaluate(A, b, method, cfg, reps) ->
    (seconds, iterations, converged, setup)` runs one candidate; the default
    builds a handle and times `reps` solves, keeping the fastest."""
    evaluate = evaluate or _evaluate_gpu
    trials = []

    def run(bs, nc):
        cfg = _lib.default_config(**dict(cfg_kw, use_abmc=1, abmc_block_size=bs,
                                         abmc_num_colors=nc))
        sec, it, conv, setup = evaluate(A, b, method, cfg, reps)
        trials.append(Trial(bs, nc, sec, it, conv, setup))

    first = candidates(block_sizes, colours if refine < 0 else coarse_colours(colours))
    for bs, nc in first:
        run(bs, nc)
    if refine > 0:
        best = choose(trials)
        done = {(t.block_size, t.num_colors) for t in trials}
        for nc in colours:
            if abs(nc - best.num_colors) <= refine and (best.block_size, nc) not in done:
                run(best.block_size, nc)
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


@dataclass
class WidthResult:
    width: int
    seconds: dict           # width -> seconds of the whole block solve


def autotune_width(h, B, cfg_kw: dict, method: int = _lib.METHOD_ICCOCR, widths=WIDTHS,
                   reps: int = 1, evaluate=None) -> WidthResult:
    """Right-hand sides per launch for solving the rows of B on handle h:
    the widths of `widths` timed on the whole block, up to the first that
    takes all rows in one launch (a wider one would only pad more); the
    results are the same bits for every width, so only time decides."""
    B = np.asarray(B)
    evaluate = evaluate or _evaluate_width_gpu
    secs = {}
    for w in sorted(widths):
        if secs and max(secs) >= B.shape[0]:
            break
        cfg = _lib.default_config(**dict(cfg_kw, rhs_block=w))
        secs[w] = evaluate(h, B, method, cfg, reps)
    return WidthResult(min(secs, key=lambda w: (secs[w], -w)), secs)


def _evaluate_width_gpu(h, B, method, cfg, reps):
    best = None
    for _ in range(max(reps, 1)):
        t = time.perf_counter()
        h.solve_multi(method, B, None, cfg)
        s = time.perf_counter() - t
        best = s if best is None else min(best, s)
    return best


@dataclass
class BatchTuneResult:
    block_sizes: list       # per system
    num_colors: list        # per system
    classes: list           # per class: (system indices, the tuned system, TuneResult)


def size_classes(sizes, classes: int = 3):
    """Indices of the systems split into `classes` groups of consecutive
    sizes (as equal in count as possible), smallest first."""
    order = sorted(range(len(sizes)), key=lambda i: (sizes[i], i))
    k = max(1, min(classes, len(order)))
    return [list(part) for part in np.array_split(np.array(order, dtype=np.int64), k) if len(part)]


def autotune_batch(systems, classes: int = 3, method: int = _lib.METHOD_ICCG, reps: int = 1,
                   evaluate=None, **kw) -> BatchTuneResult:
    """Per size class, autotune on the class's largest system (its solve
    takes the longest in every batched launch); every system gets its
    class's choice.  `systems` hold .A and .b."""
    sizes = [int(s.A.n) for s in systems]
    bs = [0] * len(systems)
    nc = [0] * len(systems)
    out = []
    for idx in size_classes(sizes, classes):
        rep = max(idx, key=lambda i: (sizes[i], i))
        r = autotune(systems[rep].A, systems[rep].b, method=method, reps=reps, evaluate=evaluate, **kw)
        for i in idx:
            bs[i], nc[i] = r.best.block_size, r.best.num_colors
        out.append((list(map(int, idx)), int(rep), r))
    return BatchTuneResult(bs, nc, out)
