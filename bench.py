#!/usr/bin/env python3
"""Time-to-tolerance benchmark of the ABMC-IC(0) Krylov path on one MI355X.

Workload (default): the real curl-curl system on a 64^3 hex grid with
lowest-order edge unknowns, boundary edges eliminated (n = 762,048,
nnz = 9,762,984), A = C^T diag(nu_f) C + 1e-6 diag(nu_e), nu = 1e-3 in a
central iron cube; RHS = C^T g with Gaussian face currents (seed 1).
Auto-shift ICCG in FP64, shift 1.05, ABMC block 4 / 4 colours, tol 1e-10.

One step = one full solve from x = 0 to ||r||/||b|| < 1e-10 with b, x and
the factor resident in HBM.  `value` = seconds per solve (time to 1e-10).

The JSON line also carries:
  roofline      the SpMV kernel (largest single kernel): algorithmic bytes per
                launch / mean launch time, timed with HIP events on the
                library's stream, against the 8 TB/s HBM peak;
  cpu_baseline  the reference (its own headers, OpenMP build) timed on this
                host on a bounded sample of the same workload;
  extra         iterations, iterations/s, setup time, the sweep kernels' rates,
                the effective per-iteration bandwidth and (with --all) the
                complex eddy-current COCR workload.
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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def bytes_model(n, nnz, nnz_l, s_a, s_f, s_v, s_p, cocr):
    """Compulsory bytes of one Krylov iteration in ABMC order: every operand
    touched once, gathers served from cache."""
    S_A = nnz * (s_a + 4) + 4 * (n + 1)
    S_L = nnz_l * (s_f + 4) + 4 * (n + 1)
    vec = n * ((20 if cocr else 14) * s_v + 2 * s_p + s_f)
    return S_A + 2 * S_L + vec


def make_workload(name, budget=None):
    s, cfg, method = _make_workload(name)
    if budget:
        cfg = dict(cfg, max_iterations=int(budget))
    return s, cfg, method


def _make_workload(name):
    from sparsesolv_amd import problems as P
    if name == "curlcurl64":
        s = P.curlcurl(64)
        cfg = dict(use_abmc=1, auto_shift=1, shift_parameter=1.05, abmc_block_size=4,
                   abmc_num_colors=4, tolerance=1e-10, max_iterations=20000)
        return s, cfg, "iccg"
    if name == "eddy78":
        s = P.eddy(78)
        cfg = dict(use_abmc=1, shift_parameter=1.05, abmc_block_size=4, abmc_num_colors=8,
                   tolerance=1e-10, max_iterations=20000)
        return s, cfg, "cocr"
    if name == "laplace48":
        s = P.poisson(48)
        cfg = dict(use_abmc=1, shift_parameter=1.05, abmc_block_size=4, abmc_num_colors=4,
                   tolerance=1e-10, max_iterations=20000)
        return s, cfg, "iccg"
    raise SystemExit(f"unknown workload {name}")


def run_gpu(name, steps, warmup, budget=None, profile_reps=30):
    import torch
    from sparsesolv_amd import _lib
    s, kw, method = make_workload(name, budget)
    A, b = s.A, s.b
    cfg = _lib.default_config(**kw)
    t0 = time.perf_counter()
    h = _lib.IcHandle(A, cfg)
    setup_wall = time.perf_counter() - t0
    info = h.info()
    log(f"[{name}] n={A.n} nnz={A.nnz} colours={info['num_colors']} blocks={info['num_blocks']} "
        f"big={info['big_blocks']}x{info['max_big_rows']} levels={info['max_levels']} "
        f"shift={info['actual_shift']} setup={setup_wall:.2f}s")
    dt = torch.complex128 if A.is_complex else torch.float64
    dev = torch.device("cuda:0")
    b_d = torch.from_numpy(np.ascontiguousarray(b)).to(dev)
    x_d = torch.zeros(A.n, dtype=dt, device=dev)
    code = _lib.METHOD_ICCG if method == "iccg" else _lib.METHOD_ICCOCR
    res = None
    for _ in range(warmup):
        x_d.zero_()
        torch.cuda.synchronize()
        res = h.solve_device(code, b_d.data_ptr(), x_d.data_ptr(), cfg)
    torch.cuda.synchronize()
    times = []
    iters = []
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
    sec = float(np.mean(times))
    s_v = 16 if A.is_complex else 8
    model = bytes_model(A.n, A.nnz, info["nnz_l"], s_v, s_v, s_v, s_v, method == "cocr")
    kern = {}
    for what, label in ((_lib.PROFILE_SPMV, "spmv"), (_lib.PROFILE_FORWARD, "forward_sweep"),
                        (_lib.PROFILE_BACKWARD, "backward_sweep"), (_lib.PROFILE_APPLY, "precond_apply")):
        if profile_reps <= 0:
            break
        us, by = h.profile(what, profile_reps)
        kern[label] = {"us": round(us, 2), "bytes": int(by), "GB_s": round(by / us / 1e3, 1),
                       "frac": round(by / us / 1e3 / PEAK_GBS, 4)}
        log(f"[{name}] {label}: {us:.1f} us  {by / 1e6:.1f} MB  {by / us / 1e3:.0f} GB/s")
    out = {
        "workload": name, "method": method, "n": A.n, "nnz": A.nnz, "nnz_l": info["nnz_l"],
        "colours": info["num_colors"], "blocks": info["num_blocks"],
        "big_block_rows": info["max_big_rows"], "big_block_levels": info["max_levels"],
        "actual_shift": info["actual_shift"], "seconds": sec, "iterations": int(np.median(iters)),
        "converged": bool(res.converged), "final_residual": res.final_residual,
        "true_rel_residual": true_rel, "setup_s": round(setup_wall, 3),
        "host_setup_s": round(info["setup_seconds"], 3),
        "iterations_per_s": round(np.median(iters) / sec, 1),
        "us_per_iteration": round(1e6 * sec / max(np.median(iters), 1), 2),
        "bytes_per_iteration_model": int(model),
        "effective_GB_s": round(model * np.median(iters) / sec / 1e9, 1),
        "kernels": kern, "times": [round(t, 5) for t in times],
    }
    return out, (A, b, kw, method)


def cpu_baseline(A, b, kw, method, gpu_iters, budget_s=20.0):
    """The reference's own ICCG / COCR (headers compiled with OpenMP) on this
    host, on a bounded number of iterations of the same system."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    kind = "reference" if pyoracle.available("ref_omp") else "port"
    orc = pyoracle.Oracle("ref_omp" if kind == "reference" else "oracle")
    cfg = orc.config(**{k: v for k, v in kw.items() if k in dict(orc.config()._fields_)})
    probe = 5
    t_probe, setup = orc.time_iterations(A, b, cfg, probe, "COCR" if method == "cocr" else "CG")
    per_it = max(t_probe / probe, 1e-6)
    iters = int(max(probe, min(gpu_iters, (budget_s - setup) / per_it)))
    t_it, setup = orc.time_iterations(A, b, cfg, iters, "COCR" if method == "cocr" else "CG")
    per_it = t_it / iters
    est = setup + per_it * gpu_iters
    return {"value": round(est, 3), "unit": "s", "cores": orc.num_threads(), "kind": kind,
            "sample": (f"setup + {iters} {method.upper()} iterations of the same system timed "
                       f"({setup:.2f}s + {t_it:.2f}s); time-to-1e-10 extrapolated at the GPU's "
                       f"{gpu_iters} iterations ({per_it * 1e3:.2f} ms/iteration)"),
            "ms_per_iteration": round(per_it * 1e3, 3), "setup_s": round(setup, 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="curlcurl64",
                    choices=["curlcurl64", "eddy78", "laplace48"])
    ap.add_argument("--all", action="store_true", help="also run the complex eddy78 COCR workload")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--budget", type=int, default=None,
                    help="fixed iteration budget per solve (per-iteration timing runs)")
    ap.add_argument("--profile-reps", type=int, default=30,
                    help="launches per kernel in the HIP-event kernel timing (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    main_out, (A, b, kw, method) = run_gpu(args.workload, args.steps, args.warmup, args.budget,
                                           args.profile_reps)
    extra = {"primary": main_out}
    if args.all and args.workload != "eddy78":
        extra["eddy78"], _ = run_gpu("eddy78", max(1, args.steps // 2), 1)
    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(A, b, kw, method, main_out["iterations"])
        log(f"cpu baseline: {cpu}")
    sp = main_out["kernels"].get("spmv", {"GB_s": None, "frac": None})
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
        "config": {"workload": f"{args.workload}-{method}-fp64", "n": A.n, "nnz": A.nnz,
                   "abmc_block_size": kw["abmc_block_size"], "abmc_num_colors": kw["abmc_num_colors"],
                   "auto_shift": bool(kw.get("auto_shift", 0)), "tol": kw["tolerance"],
                   "parallelism": "single-gpu"},
        "roofline": {"bound": "hbm", "kernel": "k_spmv (SELL-64 SpMV + fused dot)",
                     "achieved": sp["GB_s"], "peak": PEAK_GBS, "unit": "GB/s",
                     "frac": sp["frac"], "traffic": None},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
