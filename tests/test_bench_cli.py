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
