#!/usr/bin/env python3
"""Benchmark of the ABMC-IC(0) Krylov path on one MI355X.

Headline workload (`value`): mixed-precision multi-RHS ABMC-COCR on the
complex-symmetric eddy-current system of a 78^3 hex grid, lowest-order edge
unknowns, boundary edges eliminated (n = 1,387,386, nnz = 17,821,650):
A = K + j*s*M_e with K = C^T diag(nu_f) C + 1e-6 diag(nu_e) (nu = 1e-3 in a
central iron cube of 20% volume), M_e the lumped edge mass and s = 10 on the
edges of a central conductor box of 15% volume, 0 elsewhere.  A block of 16
seeded complex-Gaussian right-hand sides (seed 2) is solved in one block
solve: IC(0) factor (block 4, 8 colours, shift 1.05, unconjugated) stored and
applied in complex<float>, the COCR iteration, its dots and the true
residual in complex<double> with iterative refinement, tol 1e-10, at most
1000 iterations per right-hand side.  The complex-Gaussian right-hand sides
excite the gradient fields outside the conductor, which only the 1e-6 term
touches: neither this path nor the reference's COCR reaches 1e-10 on this
system in 1000 iterations, so a step is the solve exactly as configured and
the line reports the residuals it reached, and each right-hand side's error
against its FP64 block solution.

One step = one block solve from X = 0 with B, X, the operator and the factor
resident in HBM (setup excluded; reported apart).  `value` = seconds per
block solve.

Beside it (`config`): the FP64 block solve of the same 16 right-hand sides
and the plain FP64 one-vector solve of the first (the configured complex
system run as the plain FP64 GPU path) with its time per right-hand side.

`roofline`: the kernel with the largest share of an iteration, from an extra
solve of the same block per kernel class in which HIP events on the
library's stream bracket every launch of that class (ssa_ic_trace; these
solves run eagerly, over the timed solve's own iterations unless
`--trace-iterations` caps them, so their launches are the timed solve's
mix -- the early iterations, whose residual improves, also store the best
iterate in the update kernel, the later ones rarely do): algorithmic
bytes per launch / mean launch time, against the 8 TB/s HBM peak.
`traffic` is the HBM bytes per launch that rocprofv3's FETCH_SIZE +
WRITE_SIZE counters read for the same kernel, from the counter summary under
profiles/ (null when there is none for the kernel and workload).

`cpu_baseline`: the reference's own COCR (its headers compiled with OpenMP,
oracle/_ref/libssref_omp.so) on the first right-hand side of the same
system, on this host's cores: set-up + a bounded number of iterations,
extrapolated to 1000 iterations per right-hand side and the 16 right-hand
sides of the block (the reference solves one vector per call, in FP64).

Extra workloads (`extras`, one line of figures each; `--quick` skips them):
  laplace48   7-point Laplacian, 48^3 interior grid, ABMC-ICCG fp64 (to 1e-10);
  curlcurl64  real curl-curl, 64^3 hex grid, auto-shift ABMC-ICCG fp64, 1000
              iterations, and the same in mixed precision;
  eddy78      the headline system, one right-hand side, mixed precision;
  batch64     64 independent variable-coefficient Laplacians (16.2M DOFs)
              in one batched solve, against the same systems one by one;
  eddy78-lossy-multi16  the headline block on eddy78's grid with the
              conductor filling the box (A = K + j*0.1*M_e on every edge),
              where COCR reaches 1e-10: a true time to tolerance;
  laplace48-sgsmrtr  SGS-MRTR (solve_sgsmrtr) on laplace48's system, against
              the reference's own solve_sgsmrtr.
Every figure of every workload also goes to stderr as one JSON document, and
to --detail PATH when given; stdout carries only the final line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ngsolve-sparsesolv_amd"))

PEAK_GBS = 8000.0
METRIC = "time-to-1e-10 (s) & HBM GB/s: FP64 vs FP32-precond batched ABMC-ICCG/COCR"
HEADLINE = "eddy78-multi16"
MAX_LINE = 1600   # bytes of the final stdout line (the driver keeps the last 2000 of the output)

WORKLOADS = {
    "curlcurl64": dict(
        desc=("real curl-curl, 64^3 hex grid, interior edges (n=762,048, nnz=9,762,984), "
              "auto-shift ABMC-ICCG fp64, shift 1.05, block 4, 4 colours, tol 1e-10, "
              "max 1000 iterations"),
        method="iccg",
        cfg=dict(use_abmc=1, auto_shift=1, shift_parameter=1.05, abmc_block_size=4,
                 abmc_num_colors=4, tolerance=1e-10, max_iterations=1000)),
    "laplace48": dict(
        desc=("7-point Laplacian, 48^3 interior grid (n=110,592, nnz=760,320), ABMC-ICCG fp64, "
              "shift 1.05, block 4, 4 colours, tol 1e-10"),
        method="iccg",
        cfg=dict(use_abmc=1, shift_parameter=1.05, abmc_block_size=4, abmc_num_colors=4,
                 tolerance=1e-10, max_iterations=1000)),
    "eddy78": dict(
        desc=("complex-symmetric eddy current, 78^3 hex grid (n=1,387,386, nnz=17,821,650), "
              "A = K + j*10*M_e in a 15% conductor box, ABMC-IC(0)-COCR complex<double>, "
              "shift 1.05, block 4, 8 colours, tol 1e-10, max 1000 iterations"),
        short=("78^3 eddy current, A=K+j*10*M_e (15% conductor box), ABMC-IC(0) block 4, "
               "8 colours, shift 1.05, COCR tol 1e-10, max 1000 it"),
        method="cocr",
        cfg=dict(use_abmc=1, shift_parameter=1.05, abmc_block_size=4, abmc_num_colors=8,
                 tolerance=1e-10, max_iterations=1000)),
    "eddy78-lossy": dict(
        desc=("complex-symmetric eddy current in a lossy medium, 78^3 hex grid (n=1,387,386, "
              "nnz=17,821,650): the conductor fills the box, A = K + j*0.1*M_e on every edge; "
              "ABMC-IC(0)-COCR complex<double>, shift 1.05, block 4, 8 colours, tol 1e-10, "
              "max 1000 iterations"),
        short=("78^3 lossy eddy current, A=K+j*0.1*M_e on every edge, ABMC-IC(0) block 4, "
               "8 colours, shift 1.05, COCR tol 1e-10, max 1000 it"),
        method="cocr",
        cfg=dict(use_abmc=1, shift_parameter=1.05, abmc_block_size=4, abmc_num_colors=8,
                 tolerance=1e-10, max_iterations=1000)),
    "laplace48-sgsmrtr": dict(
        desc=("7-point Laplacian, 48^3 interior grid (n=110,592, nnz=760,320), SGS-MRTR fp64 "
              "(MRTR on D A D with the split SGS sweeps of L = lower(D A D), level schedule in "
              "natural order), tol 1e-10, max 1000 iterations"),
        method="sgsmrtr",
        cfg=dict(preconditioner=2, tolerance=1e-10, max_iterations=1000)),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


for _base in list(WORKLOADS):
    _w = WORKLOADS[_base]
    WORKLOADS[_base + "-mixed"] = dict(
        desc=_w["desc"].replace("fp64", "fp64 outer").replace(
            "complex<double>", "complex<double> outer") +
        "; IC factor and sweeps in float / complex<float> (mixed precision)",
        method=_w["method"], cfg=dict(_w["cfg"], precision=1))

for _base in ("curlcurl64", "laplace48"):
    _w = WORKLOADS[_base]
    WORKLOADS[_base + "-level"] = dict(
        desc=_w["desc"].replace("ABMC-ICCG", "level-scheduled ICCG (use_abmc=False, the "
                                             "factories' default)"),
        method=_w["method"], cfg=dict(_w["cfg"], use_abmc=0))

_SYSTEMS = {}


def make_system(name):
    from sparsesolv_amd import problems as P
    base = name.replace("-mixed", "").replace("-level", "").replace("-sgsmrtr", "")
    if base not in _SYSTEMS:
        if base == "curlcurl64":
            _SYSTEMS[base] = P.curlcurl(64)
        elif base == "laplace48":
            _SYSTEMS[base] = P.poisson(48)
        elif base == "eddy78":
            _SYSTEMS[base] = P.eddy(78)
        elif base == "eddy78-lossy":
            _SYSTEMS[base] = P.eddy(78, cond_frac=1.0, sigma=0.1)
        else:
            raise SystemExit(f"unknown workload {name}")
    return _SYSTEMS[base]


def bytes_per_iteration(n, nnz, nnz_l, s, cocr, s_low=None, nrhs=1):
    """Compulsory bytes of one Krylov iteration in ABMC order: every operand
    touched once, gathers served from cache.  s = scalar size of the matrix
    and the Krylov vectors; s_low = scalar size of the factor and of the
    preconditioner's internal vectors (s unless mixed); nnz_l counts the
    diagonal.  A block of nrhs right-hand sides reads the matrix and the
    factor once and every vector operand once per right-hand side."""
    s_f = s_p = s if s_low is None else s_low
    S_A = nnz * (s + 4) + 4 * (n + 1)
    S_L = nnz_l * (s_f + 4) + 4 * (n + 1)
    vec = n * ((20 if cocr else 14) * s + 2 * s_p) + n * s_f
    vec_rhs = n * ((20 if cocr else 14) * s + 2 * s_p)
    return S_A + 2 * S_L + (vec if nrhs == 1 else n * s_f + nrhs * vec_rhs)


def mrtr_bytes_per_iteration(n, nnz_l, s):
    """Compulsory bytes of one SGS-MRTR iteration: the factor L = lower(D A D)
    read by both sweeps (value + index per entry, a row pointer per row), and
    the vector operands once each: the sweeps' inputs and outputs (4), rd - u
    (3), Ard and its three dots (5), the p, x2, y, rd, rd / L_ii update (11)
    and the best-x copy (2)."""
    S_L = nnz_l * (s + 4) + 4 * (n + 1)
    return 2 * S_L + 25 * n * s


TRACE_NAMES = ("spmv", "colour", "update", "pq", "dot")
TRACE_KERNELS = {
    "spmv": "k_spmv (SpMV with the fused dot / convergence check)",
    "colour": "k_sweep (ABMC colour kernel, forward and backward)",
    "update": "x, r, rt update with the |r|^2 partials",
    "pq": "p, q update with the best-x copy",
    "dot": "dot partials after the preconditioner",
}


def trace_kernels(h, solve, iterations):
    """Per kernel class: one solve (`solve()` returns the iterations it ran)
    with every launch of that class bracketed by HIP events on the stream it
    runs on.  Returns launches, mean us and algorithmic bytes per launch,
    GB/s, fraction of the HBM peak and us per iteration."""
    from sparsesolv_amd import _lib
    out = {}
    for name in TRACE_NAMES:
        h.trace(getattr(_lib, "TRACE_" + name.upper()))
        try:
            its = max(1, solve())
            k, us, by = h.trace_read()
        finally:
            h.trace(_lib.TRACE_OFF)
        if k == 0 or us <= 0:
            continue
        out[name] = {"launches": int(k), "us": round(us / k, 2), "bytes": int(by / k),
                     "GB_s": round(by / us / 1e3, 1), "frac": round(by / us / 1e3 / PEAK_GBS, 4),
                     "us_per_iteration": round(us / its, 1), "iterations": int(its)}
    log(f"  traced ({iterations} iterations): " +
        ", ".join(f"{k}: {v['us']}us x{v['launches']} {v['GB_s']}GB/s" for k, v in out.items()))
    return out


def dominant(trace):
    if not trace:
        return None, None
    name = max(trace, key=lambda k: trace[k]["us_per_iteration"])
    return name, trace[name]


def recorded_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the rocprofv3
    counter summary kept beside this script (calibrated FETCH_SIZE +
    WRITE_SIZE), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.md")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1]), os.path.relpath(path, ROOT)
    return None, None


def run_gpu(name, steps, warmup, budget=None, trace_iterations=0):
    """One-vector workload: `steps` timed solves (device-resident vectors)."""
    import torch
    from sparsesolv_amd import _lib
    w = WORKLOADS[name]
    kw = dict(w["cfg"])
    if budget:
        kw["max_iterations"] = int(budget)
    t0 = time.perf_counter()
    sysm = make_system(name)
    A, b = sysm.A, sysm.b
    build_s = time.perf_counter() - t0
    cfg = _lib.default_config(**kw)
    t0 = time.perf_counter()
    h = _lib.IcHandle(A, cfg)
    setup_wall = time.perf_counter() - t0
    info = h.info()
    log(f"[{name}] n={A.n} nnz={A.nnz} colours={info['num_colors']} blocks={info['num_blocks']} "
        f"oversized={info['big_blocks']}x{info['max_big_rows']} levels={info['max_levels']} "
        f"shift={info['actual_shift']} setup={setup_wall:.2f}s (matrix build {build_s:.1f}s)")
    dt = torch.complex128 if A.is_complex else torch.float64
    dev = torch.device("cuda:0")
    b_d = torch.from_numpy(np.ascontiguousarray(b)).to(dev)
    x_d = torch.zeros(A.n, dtype=dt, device=dev)
    code = {"iccg": _lib.METHOD_ICCG, "cocr": _lib.METHOD_ICCOCR,
            "sgsmrtr": _lib.METHOD_SGSMRTR}[w["method"]]
    res = None
    for _ in range(warmup):
        x_d.zero_()
        torch.cuda.synchronize()
        res = h.solve_device(code, b_d.data_ptr(), x_d.data_ptr(), cfg)
    torch.cuda.synchronize()
    times, iters = [], []
    for k in range(steps):
        x_d.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        res = h.solve_device(code, b_d.data_ptr(), x_d.data_ptr(), cfg)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
        iters.append(res.iterations)
        log(f"[{name}] step {k}: {times[-1]:.4f}s iterations={res.iterations} "
            f"converged={res.converged} residual={res.final_residual:.3e}")
    x = x_d.cpu().numpy()
    true_rel = float(np.linalg.norm(b - A.to_scipy() @ x) / np.linalg.norm(b))
    # the same solve through host buffers (b in, x out over PCIe)
    t = time.perf_counter()
    h.solve(code, b, None, cfg)
    host_s = time.perf_counter() - t
    sec = float(np.mean(times))
    it = int(np.median(iters))
    s = 16 if A.is_complex else 8
    mixed = kw.get("precision", 0) == 1
    if w["method"] == "sgsmrtr":
        model = mrtr_bytes_per_iteration(A.n, info["nnz_l"], s)
    else:
        model = bytes_per_iteration(A.n, A.nnz, info["nnz_l"], s, w["method"] == "cocr",
                                    s // 2 if mixed else None)
    trace = {}
    if trace_iterations != 0:
        # the traced solve runs the timed solve's iterations (its mix of
        # improving iterations, whose update also stores the best iterate)
        # unless a cap is given
        tits = kw.get("max_iterations", 1000) if trace_iterations < 0 else trace_iterations
        tcfg = _lib.default_config(**dict(kw, max_iterations=tits))

        def one():
            x_d.zero_()
            return h.solve_device(code, b_d.data_ptr(), x_d.data_ptr(), tcfg).iterations
        trace = trace_kernels(h, one, tits)
    out = {
        "workload": name, "desc": w["desc"], "method": w["method"], "n": A.n, "nnz": A.nnz,
        "nnz_l": info["nnz_l"], "colours": info["num_colors"],
        "blocks": info["num_blocks"], "oversized_block_rows": info["max_big_rows"],
        "oversized_block_levels": info["max_levels"], "actual_shift": info["actual_shift"],
        "seconds": sec, "iterations": it, "converged": bool(res.converged),
        "final_residual": res.final_residual, "true_rel_residual": true_rel,
        "setup_s": round(setup_wall, 3), "host_setup_s": round(info["setup_seconds"], 3),
        "host_buffers_solve_s": round(host_s, 5),
        "setup_plus_solve_s": round(setup_wall + sec, 4),
        "iterations_per_s": round(it / sec, 1) if sec > 0 else None,
        "us_per_iteration": round(1e6 * sec / max(it, 1), 2),
        "bytes_per_iteration_model": int(model),
        "effective_GB_s": round(model * it / sec / 1e9, 1) if sec > 0 else None,
        "kernels_in_solve": trace, "times": [round(t, 5) for t in times],
    }
    return out, (A, b, kw, w["method"], x)


BATCH_DESC = ("64 independent 7-point variable-coefficient Laplacians, m ~ U{16..96} (seed 3), "
              "16,184,264 DOFs in total, cell coefficients log-uniform in [1e-2, 1e2], one uniform "
              "RHS each; ABMC-ICCG fp64, shift 1.05, block 4, 4 colours, tol 1e-10; all systems "
              "in one batched solve with per-system stop status")


def run_batch(steps, warmup, cpu=True, budget_s=12.0, tune=True):
    """The 64-system batch: one batched solve per step (device-resident
    vectors), then the same systems solved one after another on the GPU, and
    the reference on the host on a bounded sample of the systems."""
    import torch
    from sparsesolv_amd import _lib
    from sparsesolv_amd import problems as P
    t0 = time.perf_counter()
    systems = P.varcoef_batch(64)
    build_s = time.perf_counter() - t0
    kw = dict(use_abmc=1, shift_parameter=1.05, abmc_block_size=4, abmc_num_colors=4,
              tolerance=1e-10, max_iterations=2000)
    cfg = _lib.default_config(**kw)
    t0 = time.perf_counter()
    hb = _lib.BatchHandle([s.A for s in systems], cfg)
    setup_s = time.perf_counter() - t0
    dev = torch.device("cuda:0")
    b_d = torch.from_numpy(np.concatenate([s.b for s in systems])).to(dev)
    x_d = torch.zeros(hb.n, dtype=torch.float64, device=dev)
    log(f"[batch64] {len(systems)} systems, {hb.n} DOFs, setup {setup_s:.2f}s "
        f"(matrices {build_s:.1f}s)")
    res = None
    for _ in range(warmup):
        x_d.zero_()
        torch.cuda.synchronize()
        res = hb.solve_batch_device(_lib.METHOD_ICCG, b_d.data_ptr(), x_d.data_ptr(), cfg)
    times = []
    for k in range(steps):
        x_d.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        res = hb.solve_batch_device(_lib.METHOD_ICCG, b_d.data_ptr(), x_d.data_ptr(), cfg)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
        log(f"[batch64] step {k}: {times[-1]:.4f}s max iterations={max(r.iterations for r in res)} "
            f"all converged={all(r.converged for r in res)}")
    x = x_d.cpu().numpy()
    worst = 0.0
    for k, s in enumerate(systems):
        xk = hb.split(x)[k]
        worst = max(worst, float(np.linalg.norm(s.b - s.A.to_scipy() @ xk) / np.linalg.norm(s.b)))
    iters = [int(r.iterations) for r in res]
    # the same systems, one solve each (handles built beforehand, untimed)
    singles = [_lib.IcHandle(s.A, cfg) for s in systems]
    bufs = [(torch.from_numpy(s.b).to(dev), torch.zeros(s.A.n, dtype=torch.float64, device=dev))
            for s in systems]
    torch.cuda.synchronize()
    t = time.perf_counter()
    for h, (bb, xx) in zip(singles, bufs):
        h.solve_device(_lib.METHOD_ICCG, bb.data_ptr(), xx.data_ptr(), cfg)
    torch.cuda.synchronize()
    seq_s = time.perf_counter() - t
    del singles, bufs
    out = {
        "workload": "batch64", "desc": BATCH_DESC, "systems": len(systems), "dofs": hb.n,
        "seconds": float(np.mean(times)), "times": [round(t, 5) for t in times],
        "iterations_max": max(iters), "iterations_mean": round(float(np.mean(iters)), 1),
        "all_converged": bool(all(r.converged for r in res)),
        "worst_true_rel_residual": worst, "setup_s": round(setup_s, 3),
        "sequential_single_solves_s": round(seq_s, 4),
        "batching_speedup": round(seq_s / float(np.mean(times)), 2),
    }
    if cpu:
        out["cpu_baseline"] = batch_cpu_baseline(systems, kw, iters, budget_s)
    if tune:
        # autotuned ABMC parameters per size class (three classes of 21-22
        # systems, each tuned on its largest system), one batch with each
        # system's choice
        from sparsesolv_amd import autotune as at
        t0 = time.perf_counter()
        r = at.autotune_batch(systems, classes=3, tolerance=1e-10, max_iterations=2000,
                              shift_parameter=1.05)
        tune_s = time.perf_counter() - t0
        tcfg = cfg
        del hb
        ht = _lib.BatchHandle([q.A for q in systems], tcfg, block_sizes=r.block_sizes,
                              num_colors=r.num_colors)
        ht.solve_batch_device(_lib.METHOD_ICCG, b_d.data_ptr(), x_d.data_ptr(), tcfg)
        tt = []
        for _ in range(max(1, steps)):
            x_d.zero_()
            torch.cuda.synchronize()
            t = time.perf_counter()
            tres = ht.solve_batch_device(_lib.METHOD_ICCG, b_d.data_ptr(), x_d.data_ptr(), tcfg)
            torch.cuda.synchronize()
            tt.append(time.perf_counter() - t)
        classes = []
        for idx, rep, tr in r.classes:
            classes.append({
                "systems": len(idx), "n_range": [min(systems[i].A.n for i in idx),
                                                 max(systems[i].A.n for i in idx)],
                "tuned_on_n": systems[rep].A.n, "abmc_block_size": tr.best.block_size,
                "abmc_num_colors": tr.best.num_colors,
                "trials": [[t.block_size, t.num_colors, round(t.seconds, 4), t.iterations,
                            t.converged] for t in tr.trials]})
        out["tuned"] = {
            "per_class": classes, "tune_s": round(tune_s, 2), "seconds": float(np.mean(tt)),
            "iterations_max": max(int(q.iterations) for q in tres),
            "all_converged": bool(all(q.converged for q in tres))}
        log(f"[batch64] tuned per class: " +
            ", ".join(f"n<={c['n_range'][1]}: block {c['abmc_block_size']} colours "
                      f"{c['abmc_num_colors']}" for c in classes) +
            f"; {out['tuned']['seconds']:.4f}s (tuning {tune_s:.1f}s)")
    return out


MULTI_DESC = ("{sys}; 16 seeded complex-Gaussian right-hand sides (seed 2; the first is "
              "the one-vector workload's); IC factor and sweeps in complex<float> inside the "
              "complex<double> COCR with iterative refinement, all 16 in one block solve")


def run_multi(system, steps, warmup, nrhs=16, trace_iterations=0, budget=None):
    """`system` with a block of nrhs right-hand sides: the mixed-precision
    block solve per step (device-resident vectors), the FP64 block solve
    (each right-hand side's FP64 solution, for its error), and the one-vector
    FP64 and mixed solves of the first right-hand side for the time per
    vector."""
    import torch
    from sparsesolv_amd import _lib
    from sparsesolv_amd import problems as P
    w = WORKLOADS[system]
    tag = f"[{system}-multi{nrhs}]"
    t0 = time.perf_counter()
    sysm = make_system(system)
    A = sysm.A
    B = P.complex_gaussian_block(A.n, nrhs, 2)
    build_s = time.perf_counter() - t0
    kw = dict(w["cfg"])
    if budget:
        kw["max_iterations"] = int(budget)
    cfg64 = _lib.default_config(**kw)
    cfgm = _lib.default_config(**dict(kw, precision=1))
    t0 = time.perf_counter()
    hm = _lib.IcHandle(A, cfgm)
    setup_s = time.perf_counter() - t0
    h64 = _lib.IcHandle(A, cfg64)
    info = hm.info()
    log(f"{tag} n={A.n} nnz={A.nnz} colours={info['num_colors']} setup={setup_s:.2f}s "
        f"(matrix build {build_s:.1f}s)")
    dev = torch.device("cuda:0")
    B_d = torch.from_numpy(B).to(dev)
    X_d = torch.zeros_like(B_d)
    code = _lib.METHOD_ICCOCR

    def block(h, cfg):
        X_d.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        res = h.solve_multi_device(code, nrhs, B_d.data_ptr(), X_d.data_ptr(), cfg)
        torch.cuda.synchronize()
        return time.perf_counter() - t, res

    def single(h, cfg):
        b0 = B_d[0].contiguous()
        x0 = torch.zeros_like(b0)
        h.solve_device(code, b0.data_ptr(), x0.data_ptr(), cfg)    # warm-up
        x0.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = h.solve_device(code, b0.data_ptr(), x0.data_ptr(), cfg)
        torch.cuda.synchronize()
        return time.perf_counter() - t, r

    for _ in range(warmup):
        block(hm, cfgm)
    times, res = [], None
    for k in range(steps):
        t, res = block(hm, cfgm)
        times.append(t)
        log(f"{tag} step {k}: {t:.4f}s iterations max={max(r.iterations for r in res)} "
            f"all converged={all(r.converged for r in res)}")
    Xm = X_d.cpu().numpy()
    block(h64, cfg64)
    t64, res64 = block(h64, cfg64)
    X64 = X_d.cpu().numpy()
    log(f"{tag} fp64 block: {t64:.4f}s iterations max={max(r.iterations for r in res64)}")
    t1_64, r1_64 = single(h64, cfg64)
    t1_m, r1_m = single(hm, cfgm)
    log(f"{tag} one vector: fp64 {t1_64:.4f}s ({r1_64.iterations} it, residual "
        f"{r1_64.final_residual:.3e}), mixed {t1_m:.4f}s ({r1_m.iterations} it)")
    As = A.to_scipy()
    err, trel = [], []
    for k in range(nrhs):
        err.append(float(np.linalg.norm(Xm[k] - X64[k]) / np.linalg.norm(X64[k])))
        trel.append(float(np.linalg.norm(B[k] - As @ Xm[k]) / np.linalg.norm(B[k])))
    sec = float(np.mean(times))
    its = [int(r.iterations) for r in res]
    trace = {}
    if trace_iterations != 0:
        tits = kw.get("max_iterations", 1000) if trace_iterations < 0 else trace_iterations
        tcfg = _lib.default_config(**dict(kw, precision=1, max_iterations=tits))

        def one():
            X_d.zero_()
            rr = hm.solve_multi_device(code, nrhs, B_d.data_ptr(), X_d.data_ptr(), tcfg)
            return max(r.iterations for r in rr)
        trace = trace_kernels(hm, one, tits)
    model = bytes_per_iteration(A.n, A.nnz, info["nnz_l"], 16, True, 8, nrhs)
    it_mean = float(np.mean(its))
    out = {
        "workload": f"{system}-multi{nrhs}", "desc": MULTI_DESC.format(sys=w["desc"]),
        "nrhs": nrhs, "n": A.n, "nnz": A.nnz,
        "seconds": sec, "times": [round(t, 5) for t in times],
        "seconds_per_rhs": round(sec / nrhs, 5),
        "iterations": its, "converged": [bool(r.converged) for r in res],
        "all_converged": bool(all(r.converged for r in res)),
        "final_residuals": [float(r.final_residual) for r in res],
        "max_rel_error_vs_fp64": max(err), "rel_error_vs_fp64": err,
        "max_true_rel_residual": max(trel),
        "fp64_block": {"seconds": round(t64, 5),
                       "iterations": [int(r.iterations) for r in res64],
                       "all_converged": bool(all(r.converged for r in res64)),
                       "final_residuals": [float(r.final_residual) for r in res64]},
        "fp64_one_vector": {"seconds": round(t1_64, 5), "iterations": int(r1_64.iterations),
                            "converged": bool(r1_64.converged),
                            "final_residual": float(r1_64.final_residual)},
        "mixed_one_vector": {"seconds": round(t1_m, 5), "iterations": int(r1_m.iterations),
                             "final_residual": float(r1_m.final_residual)},
        "speedup_vs_fp64_one_vector_each": round(nrhs * t1_64 / sec, 2),
        "speedup_vs_fp64_block": round(t64 / sec, 2),
        "setup_s": round(setup_s, 3),
        "bytes_per_iteration_model": int(model),
        "effective_GB_s": round(model * it_mean / sec / 1e9, 1),
        "us_per_iteration": round(1e6 * sec / max(max(its), 1), 1),
        "kernels_in_solve": trace,
    }
    return out, (A, B, kw, Xm, X64)


def _reference(kind_pref="ref_omp"):
    # thread placement for the OpenMP runtime, set before the library loads
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    kind = "reference" if pyoracle.available(kind_pref) else "port"
    orc = pyoracle.Oracle(kind_pref if kind == "reference" else "oracle")
    return pyoracle, orc, kind


def batch_cpu_baseline(systems, kw, iters, budget_s):
    """The reference solving the systems one after another on this host:
    systems taken in order until the time budget is spent, each timed for the
    GPU's iteration count of that system; the total is extrapolated by
    DOFs x iterations."""
    _, orc, kind = _reference()
    fields = dict(orc.config()._fields_)
    cfg = orc.config(**{k: v for k, v in kw.items() if k in fields})
    spent, work_done, k = 0.0, 0.0, 0
    while k < len(systems) and spent < budget_s:
        s = systems[k]
        t_it, _ = orc.time_iterations(s.A, s.b, cfg, max(iters[k], 1), "CG")
        spent += t_it
        work_done += s.A.n * iters[k]
        k += 1
    work_all = float(sum(s.A.n * it for s, it in zip(systems, iters)))
    return {"value": round(spent * work_all / max(work_done, 1.0), 3), "unit": "s",
            "cores": orc.num_threads(), "kind": kind, "cpu": cpu_model(),
            "sample": (f"the first {k} of {len(systems)} systems solved by the reference "
                       f"({spent:.2f}s, each for the GPU's iteration count, setup apart), "
                       f"extrapolated by DOFs x iterations")}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(A, b, kw, method, gpu_iters, x_gpu=None, budget_s=25.0, single=True):
    """The reference's own ICCG / COCR (its headers compiled with OpenMP) on
    this host.  When it fits `budget_s`: one full solve (solution compared
    with the GPU's), then the same number of iterations timed three times
    (setup and iterations apart, the fastest run kept).  Otherwise setup + a
    bounded number of iterations, extrapolated to the GPU's count."""
    _, orc, kind = _reference()
    fields = dict(orc.config()._fields_)
    cfg = orc.config(**{k: v for k, v in kw.items() if k in fields})
    mname = "COCR" if method == "cocr" else "CG"
    probe = 5
    t_probe, setup = orc.time_iterations(A, b, cfg, probe, mname)
    per_it = max(t_probe / probe, 1e-6)
    out = {"unit": "s", "cores": orc.num_threads(), "kind": kind, "cpu": cpu_model()}
    if 2 * setup + 4 * per_it * gpu_iters <= budget_s:
        x, r = (orc.solve_cocr if method == "cocr" else orc.solve_iccg)(A, b, cfg)
        its = max(int(r["iterations"]), 1)
        runs = [orc.time_iterations(A, b, cfg, its, mname) for _ in range(3)]
        t_it = min(t for t, _ in runs)
        setup = min(su for _, su in runs)
        out.update(value=round(t_it, 4), setup_s=round(setup, 3), iterations=its,
                   sample=(f"the reference's full solve of the same system ({its} iterations, "
                           f"converged={r['converged']}, residual {r['final_residual']:.3e}); "
                           f"value = its {its} iterations timed alone, fastest of 3 runs, "
                           f"setup apart"),
                   ms_per_iteration=round(1e3 * t_it / its, 3))
        if x_gpu is not None:
            out["gpu_vs_cpu_rel_diff"] = float(np.linalg.norm(x_gpu - x) /
                                               max(np.linalg.norm(x), 1e-300))
    else:
        iters = int(max(probe, min(gpu_iters, (budget_s - setup) / per_it)))
        t_it, setup = orc.time_iterations(A, b, cfg, iters, mname)
        per_it = t_it / iters
        out.update(value=round(per_it * gpu_iters, 4), setup_s=round(setup, 3),
                   sample=(f"setup + {iters} {mname} iterations of the same system timed "
                           f"({setup:.2f}s + {t_it:.2f}s), extrapolated to the GPU's {gpu_iters} "
                           f"iterations; value excludes setup"),
                   ms_per_iteration=round(per_it * 1e3, 3))
    if single and kind == "reference" and out["cores"] > 1:
        out["single_thread"] = cpu_single_thread(orc, A, b, cfg, mname,
                                                 out.get("iterations", gpu_iters), out["cores"],
                                                 budget_s=0.4 * budget_s)
    return out


def mrtr_cpu_baseline(A, b, kw, x_gpu=None):
    """The reference's own solve_sgsmrtr (its headers compiled with OpenMP)
    on this host: the whole solve (its setup of D and L included), fastest
    of three runs, its solution compared with the GPU's."""
    _, orc, kind = _reference()
    fields = dict(orc.config()._fields_)
    cfg = orc.config(**{k: v for k, v in kw.items() if k in fields})
    runs = []
    for _ in range(3):
        t = time.perf_counter()
        x, r = orc.solve_sgsmrtr(A, b, cfg)
        runs.append(time.perf_counter() - t)
    its = int(r["iterations"])
    out = {"value": round(min(runs), 4), "unit": "s", "cores": orc.num_threads(), "kind": kind,
           "cpu": cpu_model(), "iterations": its,
           "ms_per_iteration": round(1e3 * min(runs) / max(its, 1), 3),
           "sample": (f"the reference's full solve_sgsmrtr of the same system ({its} iterations, "
                      f"converged={r['converged']}, residual {r['final_residual']:.3e}), setup "
                      f"of D and L included, fastest of 3 runs")}
    if x_gpu is not None:
        out["gpu_vs_cpu_rel_diff"] = float(np.linalg.norm(x_gpu - x) / max(np.linalg.norm(x), 1e-300))
    return out


def cpu_single_thread(orc, A, b, cfg, mname, gpu_iters, threads, budget_s):
    """The same reference build on one thread: setup + a bounded number of
    iterations, extrapolated to the GPU's iteration count."""
    orc.set_threads(1)
    try:
        t_probe, setup = orc.time_iterations(A, b, cfg, 2, mname)
        per_it = max(t_probe / 2, 1e-6)
        iters = int(max(2, min(gpu_iters, (budget_s - setup) / per_it)))
        t_it, setup = orc.time_iterations(A, b, cfg, iters, mname)
    finally:
        orc.set_threads(threads)
    per_it = t_it / iters
    return {"cores": 1, "value": round(per_it * gpu_iters, 4), "setup_s": round(setup, 3),
            "ms_per_iteration": round(per_it * 1e3, 3),
            "sample": (f"setup + {iters} {mname} iterations on one thread "
                       f"({setup:.2f}s + {t_it:.2f}s), extrapolated to {gpu_iters} iterations; "
                       f"value excludes setup")}


def multi_cpu_baseline(A, b, kw, nrhs, iterations, budget_s=20.0):
    """The reference's FP64 COCR on the first right-hand side: setup + a
    bounded number of iterations on this host's cores, extrapolated to
    `iterations` per right-hand side and nrhs right-hand sides (the reference
    solves one vector per call)."""
    _, orc, kind = _reference()
    fields = dict(orc.config()._fields_)
    cfg = orc.config(**{k: v for k, v in kw.items() if k in fields})
    t_probe, setup = orc.time_iterations(A, b, cfg, 3, "COCR")
    per_it = max(t_probe / 3, 1e-6)
    iters = int(max(3, min(iterations, (budget_s - setup) / per_it)))
    t_it, setup = orc.time_iterations(A, b, cfg, iters, "COCR")
    per_it = t_it / iters
    return {"value": round(per_it * iterations * nrhs, 3), "unit": "s", "cores": orc.num_threads(),
            "kind": kind, "cpu": cpu_model(),
            "sample": (f"reference FP64 COCR on RHS 0: setup + {iters} iterations timed "
                       f"({setup:.2f} + {t_it:.2f} s), scaled to {nrhs} x {iterations} iterations"),
            "per_rhs_s": round(per_it * iterations, 3), "setup_s": round(setup, 3),
            "ms_per_iteration": round(per_it * 1e3, 3)}


def _g(v, d=4):
    """Compact float for the summary line."""
    return None if v is None else float(f"{v:.{d}g}")


def summary(key, o):
    """One compact object per extra workload for the final line."""
    if key == "batch64":
        s = {"s": _g(o["seconds"]), "it_max": o["iterations_max"], "conv": o["all_converged"],
             "seq_s": _g(o["sequential_single_solves_s"]), "setup_s": _g(o["setup_s"], 3)}
        if "cpu_baseline" in o:
            s["cpu_s"] = _g(o["cpu_baseline"]["value"])
        if "tuned" in o:
            s["tuned_s"] = _g(o["tuned"]["seconds"])
            s["tuned_bc"] = [[c["abmc_block_size"], c["abmc_num_colors"]] for c in o["tuned"]["per_class"]]
        return s
    if "nrhs" in o:
        return {"s": _g(o["seconds"]), "it_max": max(o["iterations"]), "conv": o["all_converged"],
                "err_vs_fp64": _g(o["max_rel_error_vs_fp64"], 2),
                "fp64_block_s": _g(o["fp64_block"]["seconds"])}
    s = {"s": _g(o["seconds"]), "it": o["iterations"], "conv": o["converged"],
         "res": _g(o["final_residual"], 3)}
    if "cpu_baseline" in o:
        s["cpu_s"] = _g(o["cpu_baseline"]["value"])
    return s


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default=HEADLINE,
                    choices=sorted(WORKLOADS) + [HEADLINE, "eddy78-lossy-multi16"])
    ap.add_argument("--quick", action="store_true", help="headline workload only")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--budget", type=int, default=None,
                    help="iteration budget per solve instead of the workloads' own")
    ap.add_argument("--trace-iterations", type=int, default=-1,
                    help="iterations of each traced solve (-1: the timed solve's own budget; "
                         "0: no in-solve kernel timing)")
    ap.add_argument("--extras", default=None,
                    help="comma-separated extra workloads to run (default: all of them)")
    ap.add_argument("--detail", default=None, help="also write every figure as JSON to this path")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 1 GiB between two device buffers (a copy kernel of "
                         "known size for scaling hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path "
              f"(got --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)

    import torch
    # torch's own HIP runtime takes the GPU before the library's does
    torch.zeros(1, device="cuda:0")
    torch.cuda.synchronize()
    if args.copy_calibration:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")   # 1 GiB
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: 1073741824 bytes read, 1073741824 bytes written")
        del src, dst
    detail = {}
    cpu = None
    if args.workload.endswith("-multi16"):
        system = args.workload[:-len("-multi16")]
        main_out, (A, B, kw, Xm, X64) = run_multi(system, args.steps, args.warmup,
                                                  trace_iterations=args.trace_iterations,
                                                  budget=args.budget)
        if not args.no_cpu:
            log("cpu baseline ...")
            cpu = multi_cpu_baseline(A, B[0], kw, 16, max(main_out["iterations"]))
            log(f"cpu baseline: {cpu}")
        del Xm, X64
        cfg_line = {
            "workload": f"{args.workload}: {WORKLOADS[system]['short']}; 16 complex-Gaussian "
                        f"RHS (seed 2) in one block, c64 IC factor and sweeps in c128 COCR",
            "n": A.n, "nnz": A.nnz, "nrhs": 16, "precision": "mixed",
            "iterations_max": max(main_out["iterations"]),
            "converged": f"{sum(main_out['converged'])}/16",
            "final_residual_max": _g(max(main_out["final_residuals"]), 4),
            "err_vs_fp64_block_max": _g(main_out["max_rel_error_vs_fp64"], 3),
            "fp64_block_s": main_out["fp64_block"]["seconds"],
            "fp64_one_vector_s": main_out["fp64_one_vector"]["seconds"],
            "speedup_vs_fp64_block": main_out["speedup_vs_fp64_block"],
            "us_per_iteration": main_out["us_per_iteration"]}
        dtype = "c128 (c64 preconditioner)"
    else:
        main_out, (A, b, kw, method, x_gpu) = run_gpu(args.workload, args.steps, args.warmup,
                                                     args.budget, args.trace_iterations)
        if not args.no_cpu:
            log("cpu baseline ...")
            if method == "sgsmrtr":
                cpu = mrtr_cpu_baseline(A, b, kw, x_gpu)
            else:
                cpu = cpu_baseline(A, b, kw, method, main_out["iterations"], x_gpu)
            log(f"cpu baseline: {cpu}")
        cfg_line = {"workload": f"{args.workload}: {WORKLOADS[args.workload]['desc']}",
                    "n": A.n, "nnz": A.nnz, "iterations": main_out["iterations"],
                    "converged": main_out["converged"],
                    "final_residual": _g(main_out["final_residual"], 4),
                    "parallelism": "single-gpu"}
        mixed = WORKLOADS[args.workload]["cfg"].get("precision", 0) == 1
        base = "c128" if A.is_complex else "f64"
        dtype = base + ((" (c64 preconditioner)" if A.is_complex else " (f32 preconditioner)")
                        if mixed else "")
    detail[args.workload] = main_out
    kname, k = dominant(main_out["kernels_in_solve"])
    roofline = None
    if k is not None:
        traffic, traffic_src = recorded_traffic(kname, args.workload)
        roofline = {"bound": "hbm", "kernel": TRACE_KERNELS[kname], "achieved": k["GB_s"],
                    "peak": PEAK_GBS, "unit": "GB/s", "frac": k["frac"], "traffic": traffic,
                    "bytes_per_launch": k["bytes"], "us_per_launch": k["us"],
                    "share_us_per_iteration": k["us_per_iteration"]}
        if traffic_src:
            roofline["traffic_source"] = traffic_src

    names = ("laplace48", "curlcurl64", "curlcurl64-mixed", "eddy78-mixed", "batch64",
             "eddy78-lossy-multi16", "laplace48-sgsmrtr")
    wanted = set(names if args.extras is None else [e for e in args.extras.split(",") if e])
    extras = {}
    if not args.quick:
        esteps = max(1, min(args.steps, 3))
        for other in names:
            if other == args.workload or other not in wanted:
                continue
            if other == "batch64":
                o = run_batch(esteps, 1, cpu=not args.no_cpu)
            elif other.endswith("-multi16"):
                o, _ = run_multi(other[:-len("-multi16")], esteps, 1, budget=args.budget)
            else:
                o, (A2, b2, kw2, m2, x2) = run_gpu(other, esteps, 1, args.budget)
                if not args.no_cpu and other in ("laplace48", "curlcurl64"):
                    log(f"cpu baseline ({other}) ...")
                    o["cpu_baseline"] = cpu_baseline(A2, b2, kw2, m2, o["iterations"], x2,
                                                     budget_s=12.0, single=False)
                elif not args.no_cpu and m2 == "sgsmrtr":
                    log(f"cpu baseline ({other}) ...")
                    o["cpu_baseline"] = mrtr_cpu_baseline(A2, b2, kw2, x2)
            detail[other] = o
            extras[other] = summary(other, o)
            log(f"[{other}] {json.dumps(extras[other])}")
    line = {
        "metric": METRIC,
        "value": round(main_out["seconds"], 6),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * main_out["seconds"], 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": cfg_line,
        "roofline": roofline,
        "cpu_baseline": None if cpu is None else {
            k2: cpu[k2] for k2 in ("value", "unit", "cores", "kind", "sample", "setup_s",
                                   "ms_per_iteration") if k2 in cpu},
        "extras": extras,
    }
    full = dict(line, detail=detail)
    text = json.dumps(full)
    log("detail: " + text)
    if args.detail:
        with open(args.detail, "w") as f:
            f.write(text + "\n")
    print(final_line(line), flush=True)


def final_line(line):
    """The stdout line: compact JSON of at most MAX_LINE bytes (the extras'
    figures are cut to their times in seconds first, then dropped)."""
    out = json.dumps(line, separators=(",", ":"))
    if len(out) > MAX_LINE:
        line = dict(line, extras={k: v.get("s") for k, v in line["extras"].items()})
        out = json.dumps(line, separators=(",", ":"))
    if len(out) > MAX_LINE:
        line = dict(line, extras={})
        out = json.dumps(line, separators=(",", ":"))
    return out


if __name__ == "__main__":
    main()
