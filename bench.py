#!/usr/bin/env python3
"""Benchmark of the ABMC-IC(0) Krylov path on one MI355X.

Primary workload (the `value` line): the real curl-curl system on a 64^3 hex
grid, lowest-order edge unknowns, boundary edges eliminated (n = 762,048,
nnz = 9,762,984): A = C^T diag(nu_f) C + 1e-6 diag(nu_e), nu = 1e-3 in a
central iron cube, RHS = C^T g with Gaussian face currents (seed 1).
Auto-shift ABMC-ICCG in FP64: shift 1.05, block 4, 4 colours requested,
tol 1e-10, max_iterations 1000 (SparseSolvSolver's default).  IC(0)-CG does
not reach 1e-10 on this near-singular operator within 1000 iterations (the
reference does not either): a step is the solve exactly as the reference
runs it, and the line reports the residual it reached.

One step = one solve from x = 0 with b, x, the operator and the factor
resident in HBM (setup excluded; it is reported separately).  `value` =
seconds per solve.

Extra workloads (`extra`, skipped with --quick):
  laplace48  7-point Laplacian on a 48^3 interior grid (n = 110,592), ABMC-ICCG
             shift 1.05, block 4, 4 colours, tol 1e-10: converges, so this is
             a true time-to-1e-10, with the error against the reference's
             solution;
  eddy78     complex-symmetric eddy-current system on 78^3 (n = 1,387,386),
             A = K + j*10*M_e in a conductor box, ABMC-IC(0)-COCR in
             complex<double>, block 4, 8 colours, tol 1e-10, 1000 iterations;
  *-mixed    each of the three with the IC factor and both sweeps in float /
             complex<float> inside the FP64 Krylov loop (SpMV, vectors, dots
             and the final true-residual check in FP64);
  *-level    curlcurl64 and laplace48 on the level-scheduled path
             (use_abmc=False, the factories' default), with the reference's
             own level path as the CPU baseline;
  batch64    64 independent variable-coefficient Laplacians (16.2M DOFs)
             in one batched solve, against the same systems solved one by
             one on the GPU and the reference on a sample of them;
  eddy78-multi16  eddy78 with a block of 16 seeded complex-Gaussian
             right-hand sides in one mixed-precision block solve (the SpMV
             and the colour sweeps load each matrix and factor entry once for
             all 16), with each right-hand side's error against its FP64
             solution, beside the FP64 block solve and the one-vector FP64
             and mixed solves;
  eddy78-lossy-multi16  the same on eddy78's grid with the conductor
             filling the box (A = K + j*0.1*M_e on every edge): COCR reaches
             1e-10 there, so the block solve is a time to tolerance and the
             errors compare converged solutions (and the reference's, for the
             first right-hand side).

The JSON line also carries
  roofline      the dominant kernel by time, the ABMC colour kernel k_sweep:
                algorithmic bytes per launch / mean launch time.  Every colour
                launched on its own runs `--profile-reps` times back to back
                between two HIP events on the library's stream; the figures
                are means over those colours, against the 8 TB/s HBM peak.
                `traffic` is the HBM bytes per launch that rocprofv3's
                FETCH_SIZE + WRITE_SIZE counters read for the same launches,
                taken from the counter summary under profiles/ (null when
                there is none for the workload);
  cpu_baseline  the reference (its own headers, OpenMP build) on this host's
                cores, on the same system: the full solve when it fits the
                time budget, else a bounded sample extrapolated.
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
              "ABMC-IC(0)-COCR complex<double>, shift 1.05, block 4, 8 colours, tol 1e-10, "
              "max 1000 iterations"),
        method="cocr",
        cfg=dict(use_abmc=1, shift_parameter=1.05, abmc_block_size=4, abmc_num_colors=8,
                 tolerance=1e-10, max_iterations=1000)),
    "eddy78-lossy": dict(
        desc=("complex-symmetric eddy current in a lossy medium, 78^3 hex grid (n=1,387,386, "
              "nnz=17,821,650): eddy78's operator with the conductor filling the box, "
              "A = K + j*0.1*M_e on every edge; ABMC-IC(0)-COCR complex<double>, shift 1.05, "
              "block 4, 8 colours, tol 1e-10, max 1000 iterations"),
        method="cocr",
        cfg=dict(use_abmc=1, shift_parameter=1.05, abmc_block_size=4, abmc_num_colors=8,
                 tolerance=1e-10, max_iterations=1000)),
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
    base = name.replace("-mixed", "").replace("-level", "")
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


def run_gpu(name, steps, warmup, budget=None, profile_reps=20):
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
    code = _lib.METHOD_ICCG if w["method"] == "iccg" else _lib.METHOD_ICCOCR
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
    model = bytes_per_iteration(A.n, A.nnz, info["nnz_l"], s, w["method"] == "cocr",
                                s // 2 if mixed else None)
    kern = {}
    for what, label in ((_lib.PROFILE_SPMV, "spmv"), (_lib.PROFILE_FORWARD, "forward_sweep"),
                        (_lib.PROFILE_BACKWARD, "backward_sweep"),
                        (_lib.PROFILE_APPLY, "precond_apply"),
                        (_lib.PROFILE_COLOUR_KERNELS, "colour_kernel")):
        if profile_reps <= 0:
            break
        us, by = h.profile(what, profile_reps)
        if us <= 0:
            continue   # no colour launched on its own (all in small-colour kernels)
        kern[label] = {"us": round(us, 2), "bytes": int(by), "GB_s": round(by / us / 1e3, 1),
                       "frac": round(by / us / 1e3 / PEAK_GBS, 4)}
        log(f"[{name}] {label}: {us:.2f} us  {by / 1e6:.2f} MB  {by / us / 1e3:.0f} GB/s")
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
        "kernels": kern, "times": [round(t, 5) for t in times],
    }
    return out, (A, b, kw, w["method"], x)


BATCH_DESC = ("64 independent 7-point variable-coefficient Laplacians, m ~ U{16..96} (seed 3), "
              "16,184,264 DOFs in total, cell coefficients log-uniform in [1e-2, 1e2], one uniform "
              "RHS each; ABMC-ICCG fp64, shift 1.05, block 4, 4 colours, tol 1e-10; all systems "
              "in one batched solve with per-system stop status")


def run_batch(steps, warmup, cpu=True, budget_s=25.0, profile_reps=10, tune=True):
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
    us, by = hb.profile(_lib.PROFILE_COLOUR_KERNELS, profile_reps) if profile_reps > 0 else (0.0, 0.0)
    out = {
        "workload": "batch64", "desc": BATCH_DESC, "systems": len(systems), "dofs": hb.n,
        "seconds": float(np.mean(times)), "times": [round(t, 5) for t in times],
        "iterations_max": max(iters), "iterations_mean": round(float(np.mean(iters)), 1),
        "all_converged": bool(all(r.converged for r in res)),
        "worst_true_rel_residual": worst, "setup_s": round(setup_s, 3),
        "sequential_single_solves_s": round(seq_s, 4),
        "batching_speedup": round(seq_s / float(np.mean(times)), 2),
        "colour_kernel": None if us <= 0 else {"us": round(us, 2), "bytes": int(by),
                                                "GB_s": round(by / us / 1e3, 1)},
    }
    if cpu:
        out["cpu_baseline"] = batch_cpu_baseline(systems, kw, iters, budget_s)
    if tune:
        # autotuned ABMC parameters, chosen on the largest system
        from sparsesolv_amd import autotune as at
        big = max(systems, key=lambda q: q.A.n)
        r = at.autotune(big.A, big.b, tolerance=1e-10, max_iterations=2000, shift_parameter=1.05)
        tkw = r.config(**kw)
        tcfg = _lib.default_config(**tkw)
        del hb
        ht = _lib.BatchHandle([q.A for q in systems], tcfg)
        ht.solve_batch_device(_lib.METHOD_ICCG, b_d.data_ptr(), x_d.data_ptr(), tcfg)
        tt = []
        for _ in range(max(1, steps)):
            x_d.zero_()
            torch.cuda.synchronize()
            t = time.perf_counter()
            tres = ht.solve_batch_device(_lib.METHOD_ICCG, b_d.data_ptr(), x_d.data_ptr(), tcfg)
            torch.cuda.synchronize()
            tt.append(time.perf_counter() - t)
        out["tuned"] = {
            "abmc_block_size": r.best.block_size, "abmc_num_colors": r.best.num_colors,
            "tuned_on": f"the largest system (n={big.A.n})", "seconds": float(np.mean(tt)),
            "iterations_max": max(int(q.iterations) for q in tres),
            "all_converged": bool(all(q.converged for q in tres)),
            "trials": [[t.block_size, t.num_colors, round(t.seconds, 4), t.iterations, t.converged]
                       for t in r.trials]}
        log(f"[batch64] tuned: {out['tuned']}")
    return out


MULTI_DESC = ("{sys}; 16 seeded complex-Gaussian right-hand sides (seed 2; the first is "
              "the one-vector workload's); IC factor and sweeps in complex<float> inside the "
              "complex<double> COCR with iterative refinement, all 16 in one block solve")


def run_multi(system, steps, warmup, cpu=True, profile_reps=10, nrhs=16, budget_s=40.0):
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
    sysm = make_system(system)
    A = sysm.A
    B = P.complex_gaussian_block(A.n, nrhs, 2)
    kw = dict(w["cfg"])
    cfg64 = _lib.default_config(**kw)
    cfgm = _lib.default_config(**dict(kw, precision=1))
    t0 = time.perf_counter()
    hm = _lib.IcHandle(A, cfgm)
    setup_s = time.perf_counter() - t0
    h64 = _lib.IcHandle(A, cfg64)
    info = hm.info()
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
    log(f"{tag} one vector: fp64 {t1_64:.4f}s ({r1_64.iterations} it), "
        f"mixed {t1_m:.4f}s ({r1_m.iterations} it)")
    As = A.to_scipy()
    err, trel = [], []
    for k in range(nrhs):
        err.append(float(np.linalg.norm(Xm[k] - X64[k]) / np.linalg.norm(X64[k])))
        trel.append(float(np.linalg.norm(B[k] - As @ Xm[k]) / np.linalg.norm(B[k])))
    sec = float(np.mean(times))
    its = [int(r.iterations) for r in res]
    kern = {}
    for what, label in ((_lib.PROFILE_SPMV, "spmv"), (_lib.PROFILE_FORWARD, "forward_sweep"),
                        (_lib.PROFILE_BACKWARD, "backward_sweep"),
                        (_lib.PROFILE_APPLY, "precond_apply"),
                        (_lib.PROFILE_COLOUR_KERNELS, "colour_kernel")):
        if profile_reps <= 0:
            break
        us, by = hm.profile_multi(what, nrhs, profile_reps)
        if us <= 0:
            continue
        kern[label] = {"us": round(us, 2), "bytes": int(by), "GB_s": round(by / us / 1e3, 1),
                       "frac": round(by / us / 1e3 / PEAK_GBS, 4)}
        log(f"{tag} {label} (x{nrhs}): {us:.2f} us  {by / 1e6:.2f} MB  {by / us / 1e3:.0f} GB/s")
    model = bytes_per_iteration(A.n, A.nnz, info["nnz_l"], 16, True, 8, nrhs)
    it_mean = float(np.mean(its))
    out = {
        "workload": f"{system}-multi{nrhs}", "desc": MULTI_DESC.format(sys=w["desc"]),
        "nrhs": nrhs, "n": A.n,
        "seconds": sec, "times": [round(t, 5) for t in times],
        "seconds_per_rhs": round(sec / nrhs, 5),
        "iterations": its, "all_converged": bool(all(r.converged for r in res)),
        "max_rel_error_vs_fp64": max(err), "rel_error_vs_fp64": err,
        "max_true_rel_residual": max(trel),
        "fp64_block": {"seconds": round(t64, 5),
                       "iterations": [int(r.iterations) for r in res64],
                       "all_converged": bool(all(r.converged for r in res64))},
        "fp64_one_vector": {"seconds": round(t1_64, 5), "iterations": int(r1_64.iterations),
                            "converged": bool(r1_64.converged)},
        "mixed_one_vector": {"seconds": round(t1_m, 5), "iterations": int(r1_m.iterations)},
        "speedup_vs_fp64_one_vector_each": round(nrhs * t1_64 / sec, 2),
        "speedup_vs_fp64_block": round(t64 / sec, 2),
        "setup_s": round(setup_s, 3),
        "bytes_per_iteration_model": int(model),
        "effective_GB_s": round(model * it_mean / sec / 1e9, 1),
        "kernels": kern,
    }
    if cpu:
        out["cpu_reference_rhs0"] = multi_cpu_check(A, B[0], kw, Xm[0], X64[0], int(r1_64.iterations),
                                                    budget_s)
    return out


def multi_cpu_check(A, b, kw, xm, x64, gpu_iters, budget_s):
    """The reference's FP64 COCR on the first right-hand side, when a full
    solve fits the budget (estimated from 5 timed iterations and the GPU's
    FP64 iteration count): the errors of the GPU's mixed and FP64 solutions
    against it."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    kind = "reference" if pyoracle.available("ref_omp") else "port"
    orc = pyoracle.Oracle("ref_omp" if kind == "reference" else "oracle")
    fields = dict(orc.config()._fields_)
    cfg = orc.config(**{k: v for k, v in kw.items() if k in fields})
    t_probe, setup = orc.time_iterations(A, b, cfg, 5, "COCR")
    est = setup + t_probe / 5 * 1.2 * min(gpu_iters, kw.get("max_iterations", 1000))
    if est > budget_s:
        return {"skipped": f"estimated {est:.0f}s for a full reference solve > {budget_s:.0f}s"}
    t = time.perf_counter()
    x, r = orc.solve_cocr(A, b, cfg)
    return {"kind": kind, "cores": orc.num_threads(), "seconds": round(time.perf_counter() - t, 3),
            "iterations": int(r["iterations"]), "converged": bool(r["converged"]),
            "mixed_rel_error": float(np.linalg.norm(xm - x) / np.linalg.norm(x)),
            "fp64_rel_error": float(np.linalg.norm(x64 - x) / np.linalg.norm(x))}


def batch_cpu_baseline(systems, kw, iters, budget_s):
    """The reference solving the systems one after another on this host:
    systems taken in order until the time budget is spent, each timed for the
    GPU's iteration count of that system; the total is extrapolated by
    DOFs x iterations."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    kind = "reference" if pyoracle.available("ref_omp") else "port"
    orc = pyoracle.Oracle("ref_omp" if kind == "reference" else "oracle")
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


def cpu_baseline(A, b, kw, method, gpu_iters, x_gpu=None, budget_s=25.0):
    """The reference's own ICCG / COCR (its headers compiled with OpenMP) on
    this host.  When it fits `budget_s`: one full solve (solution compared
    with the GPU's), then the same number of iterations timed three times
    (setup and iterations apart, the fastest run kept).  Otherwise setup + a
    bounded number of iterations, extrapolated to the GPU's count."""
    # thread placement for the OpenMP runtime, set before the library loads
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    kind = "reference" if pyoracle.available("ref_omp") else "port"
    orc = pyoracle.Oracle("ref_omp" if kind == "reference" else "oracle")
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
    if kind == "reference" and out["cores"] > 1:
        out["single_thread"] = cpu_single_thread(orc, A, b, cfg, mname,
                                                 out.get("iterations", gpu_iters), out["cores"],
                                                 budget_s=0.4 * budget_s)
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


def recorded_traffic(workload):
    """HBM bytes per colour-kernel launch from the rocprofv3 counter summary
    kept beside this script (calibrated FETCH_SIZE + WRITE_SIZE), or None."""
    path = os.path.join(ROOT, "profiles", f"colour_kernel_{workload}_traffic.md")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1]), os.path.relpath(path, ROOT)
    return None, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="curlcurl64", choices=sorted(WORKLOADS))
    ap.add_argument("--quick", action="store_true", help="primary workload only")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--budget", type=int, default=None,
                    help="fixed iteration budget per solve (per-iteration timing runs)")
    ap.add_argument("--profile-reps", type=int, default=20,
                    help="repetitions in the HIP-event kernel timing (0: skip)")
    ap.add_argument("--extras", default=None,
                    help="comma-separated extra workloads to run (default: all of them)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 1 GiB between two device buffers (a copy kernel of "
                         "known size for scaling hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: this build has no multi-GPU path "
              f"(got --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        import torch
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")   # 1 GiB
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copy calibration: 1073741824 bytes read, 1073741824 bytes written")
        del src, dst
    main_out, (A, b, kw, method, x_gpu) = run_gpu(args.workload, args.steps, args.warmup,
                                                 args.budget, args.profile_reps)
    extra = {"primary": main_out}
    names = ("curlcurl64-mixed", "curlcurl64-level", "laplace48", "laplace48-mixed",
             "laplace48-level", "eddy78", "eddy78-mixed", "batch64", "eddy78-multi16",
             "eddy78-lossy-multi16")
    wanted = set(names if args.extras is None else args.extras.split(","))
    if not args.quick:
        for other in names[:7]:
            if other == args.workload or other not in wanted:
                continue
            o, (A2, b2, kw2, m2, x2) = run_gpu(other, max(1, args.steps), 1, args.budget,
                                               args.profile_reps)
            if not args.no_cpu and other in ("laplace48", "laplace48-level", "curlcurl64-level"):
                log(f"cpu baseline ({other}) ...")
                o["cpu_baseline"] = cpu_baseline(A2, b2, kw2, m2, o["iterations"], x2)
                log(f"cpu baseline ({other}): {o['cpu_baseline']}")
            extra[other] = o
    if not args.quick and "batch64" in wanted:
        extra["batch64"] = run_batch(max(1, args.steps), 1, cpu=not args.no_cpu,
                                     profile_reps=args.profile_reps)
        log(f"[batch64] {json.dumps({k: v for k, v in extra['batch64'].items() if k != 'desc'})}")
    for system in ("eddy78", "eddy78-lossy"):
        key = f"{system}-multi16"
        if not args.quick and key in wanted:
            extra[key] = run_multi(system, max(1, args.steps), 1, cpu=not args.no_cpu,
                                   profile_reps=args.profile_reps)
            log(f"[{key}] {json.dumps({k: v for k, v in extra[key].items() if k != 'desc'})}")
    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(A, b, kw, method, main_out["iterations"], x_gpu)
        log(f"cpu baseline: {cpu}")
    ck = main_out["kernels"].get("colour_kernel")
    traffic, traffic_src = recorded_traffic(args.workload)
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
        "dtype": "c128" if A.is_complex else "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {WORKLOADS[args.workload]['desc']}",
                   "n": A.n, "nnz": A.nnz, "iterations": main_out["iterations"],
                   "converged": main_out["converged"],
                   "final_residual": main_out["final_residual"],
                   "parallelism": "single-gpu"},
        "roofline": None if ck is None else {
            "bound": "hbm", "kernel": "k_sweep (ABMC colour kernel, forward and backward)",
            "achieved": ck["GB_s"], "peak": PEAK_GBS, "unit": "GB/s", "frac": ck["frac"],
            "traffic": traffic, "traffic_source": traffic_src,
            "bytes_per_launch": ck["bytes"], "us_per_launch": ck["us"]},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
