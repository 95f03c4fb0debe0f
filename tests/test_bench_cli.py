"""bench.py contract checks that need no GPU."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_bytes_per_iteration_model():
    # 7-point Laplacian 48^3: n = 110,592, nnz = 760,320, nnz_L = (nnz + n) / 2
    n, nnz = 110592, 760320
    assert bench.bytes_per_iteration(n, nnz, (nnz + n) // 2, 8, False) == 35_942_412
    # curl-curl 64^3 interior edges
    n, nnz = 762048, 9762984
    assert bench.bytes_per_iteration(n, nnz, (nnz + n) // 2, 8, False) == 356_239_308
    # eddy 78^3, complex<double> COCR
    n, nnz = 1387386, 17821650
    assert bench.bytes_per_iteration(n, nnz, (nnz + n) // 2, 16, True) == 1_267_820_412
    # the same system with a complex<float> factor and sweeps
    assert bench.bytes_per_iteration(n, nnz, (nnz + n) // 2, 16, True, 8) == \
        1_267_820_412 - 2 * ((nnz + n) // 2) * 8 - n * (2 * 8 + 8)


def test_cpu_baseline_reports_one_thread_beside_all_threads():
    import pytest
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    if not pyoracle.available("ref_omp"):
        pytest.skip("the reference's OpenMP build (oracle/_ref/libssref_omp.so) is not present")
    from sparsesolv_amd import problems as P
    s = P.poisson(12)
    out = bench.cpu_baseline(s.A, s.b, dict(use_abmc=1), "iccg", 30, budget_s=5.0)
    assert out["kind"] == "reference" and out["value"] > 0
    if out["cores"] > 1:
        one = out["single_thread"]
        assert one["cores"] == 1 and one["value"] > 0 and one["ms_per_iteration"] > 0


def _fake_outputs():
    """Figures shaped like the bench's own (the longest strings it writes)."""
    one = {"seconds": 0.2374, "iterations": 1000, "converged": False, "final_residual": 0.2371,
           "cpu_baseline": {"value": 2.27}}
    multi = {"nrhs": 16, "seconds": 1.6, "iterations": [517] * 16, "all_converged": True,
             "max_rel_error_vs_fp64": 8.8e-11, "fp64_block": {"seconds": 1.67}}
    batch = {"seconds": 0.53, "iterations_max": 309, "all_converged": True,
             "sequential_single_solves_s": 1.71, "setup_s": 3.2, "cpu_baseline": {"value": 8.1},
             "tuned": {"seconds": 0.52, "per_class": [{"abmc_block_size": 2, "abmc_num_colors": 4}, {"abmc_block_size": 4, "abmc_num_colors": 8}]}}
    extras = {k: bench.summary(k, v) for k, v in (
        ("laplace48", one), ("curlcurl64", one), ("curlcurl64-mixed", one),
        ("eddy78-mixed", one), ("batch64", batch), ("eddy78-lossy-multi16", multi))}
    return extras


def test_final_line_is_compact_json():
    import json
    desc = bench.WORKLOADS["eddy78"]["short"]
    line = {
        "metric": bench.METRIC, "value": 2.881234, "unit": "s", "n_gpus": 1, "steps": 20,
        "warmup": 5, "ms_per_step": 2881.234, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "c128 (c64 preconditioner)", "data": "synthetic",
        "config": {"workload": "eddy78-multi16: " + desc + "; 16 complex-Gaussian RHS "
                   "(seed 2) in one block, c64 IC factor and sweeps in c128 COCR",
                   "n": 1387386, "nnz": 17821650, "nrhs": 16,
                   "precision": "mixed", "iterations_max": 1000, "converged": "0/16",
                   "final_residual_max": 0.9123, "err_vs_fp64_block_max": 0.512,
                   "fp64_block_s": 3.36123, "fp64_one_vector_s": 0.47512,
                   "speedup_vs_fp64_block": 1.17, "us_per_iteration": 2881.2},
        "roofline": {"bound": "hbm", "kernel": bench.TRACE_KERNELS["colour"], "achieved": 1234.5,
                     "peak": 8000.0, "unit": "GB/s", "frac": 0.1543, "traffic": 85123456,
                     "bytes_per_launch": 23456789, "us_per_launch": 28.12,
                     "share_us_per_iteration": 634.1,
                     "traffic_source": "profiles/colour_eddy78-multi16_traffic.md"},
        "cpu_baseline": {"value": 1634.123, "unit": "s", "cores": 16, "kind": "reference",
                         "sample": "reference FP64 COCR on RHS 0: setup + 180 iterations "
                                   "timed (1.23 + 18.77 s), scaled to 16 x 1000 iterations",
                         "setup_s": 1.23,
                         "ms_per_iteration": 102.1},
        "extras": _fake_outputs(),
    }
    out = bench.final_line(line)
    assert len(out) <= bench.MAX_LINE, len(out)
    back = json.loads(out)
    assert back["value"] == line["value"] and back["roofline"]["frac"] == 0.1543
    assert "\n" not in out
    # every extra workload keeps at least its time
    assert set(back["extras"]) == set(line["extras"])
    # room to spare under the 2000 characters of output the driver keeps
    assert bench.MAX_LINE <= 1600


def test_dominant_kernel_is_the_largest_share():
    tr = {"spmv": {"us_per_iteration": 590.0}, "colour": {"us_per_iteration": 640.0},
          "pq": {"us_per_iteration": 630.0}}
    assert bench.dominant(tr)[0] == "colour"
    assert bench.dominant({}) == (None, None)
