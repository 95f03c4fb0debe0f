"""HBM layouts checked on the host (tests/cpp/layout_check, built by build()).

The checker builds a system's plan with the library's own host code
(host_setup.cpp), checks every index the kernels form against its array,
runs the SpMV from its SELL-64 layout and both sweeps lane by lane from the
sweep layouts -- regular slices with in-block shuffles, oversized blocks by
the packed walk records or the level fallback -- and compares them with the
CSR product and a sequential substitution over the host factor.  This covers
the layouts of every shape case the GPU tests run, with no GPU."""
import os
import subprocess

import numpy as np
import pytest

from sparsesolv_amd import problems as P

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "layout_check")


def run_check(tmp_path, A, block=4, colours=4, abmc=True):
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/layout_check is not built (run __graft_entry__.build())")
    path = tmp_path / "A.bin"
    with open(path, "wb") as f:
        np.array([A.n], dtype=np.int32).tofile(f)
        A.row_ptr.astype(np.int32).tofile(f)
        A.col_idx.astype(np.int32).tofile(f)
        A.values.astype(np.complex128 if A.is_complex else np.float64).tofile(f)
    out = subprocess.run([BIN, str(path), str(int(A.is_complex)), str(block), str(colours),
                          str(int(abmc))], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok"), out.stdout
    info = dict(kv.split("=") for kv in out.stdout.split()[1:8])
    return {k: int(v.split("/")[0]) for k, v in info.items()}


def varcoef(m, seed=21):
    return P.laplace7_varcoef(m, 10.0 ** P.uniform(seed, m ** 3, -2.0, 2.0))


CASES = [
    ("lap9", lambda: P.laplace7(9), 4, 4),
    ("lap12_b8", lambda: P.laplace7(12), 8, 6),
    ("lap10_b16", lambda: P.laplace7(10), 16, 16),
    ("lap8_b3", lambda: P.laplace7(8), 3, 5),
    ("varcoef14", lambda: varcoef(14), 4, 4),
    ("cc6", lambda: P.curlcurl(6).A, 4, 4),
    ("cc6_b16", lambda: P.curlcurl(6).A, 16, 4),
    ("eddy6", lambda: P.eddy(6).A, 4, 8),
    ("eddy5_b2", lambda: P.eddy(5).A, 2, 4),
    # tens of thousands of rows: the host setup (permutation, factor,
    # transpose) runs on threads
    ("cc24", lambda: P.curlcurl(24).A, 4, 4),
]


@pytest.mark.parametrize("name,builder,block,colours", CASES, ids=[c[0] for c in CASES])
def test_layouts_abmc(tmp_path, name, builder, block, colours):
    run_check(tmp_path, builder(), block, colours)


def test_layouts_walk_of_oversized_block(tmp_path):
    # 48^3 Laplacian, B=4: the aggregation leaves one 366-row block, walked
    # step by step from the packed records (and its prefetch tail)
    info = run_check(tmp_path, P.laplace7(48))
    assert info["big"] == 1 and info["big_rows"] == 366 and info["walk_k"] > 0


def test_layouts_wide_blocks_level_fallback(tmp_path):
    # blocks wider than 16 rows: every block is level scheduled; the wide
    # curl-curl blocks have more than 8 in-block entries per row, so the
    # level kernel runs them instead of the walk
    info = run_check(tmp_path, P.curlcurl(5).A, block=64, colours=4)
    assert info["BR"] == 0 and info["big"] > 0 and info["walk_k"] == 0
    info = run_check(tmp_path, P.laplace7(10), block=32, colours=4)
    assert info["BR"] == 0 and info["big"] > 0


@pytest.mark.parametrize("name,builder", [
    ("varcoef9", lambda: varcoef(9, seed=7)),
    ("cc5", lambda: P.curlcurl(5).A),
    ("eddy5", lambda: P.eddy(5).A),
    ("cc24", lambda: P.curlcurl(24).A),
], ids=["varcoef9", "cc5", "eddy5", "cc24"])
def test_layouts_level_path(tmp_path, name, builder):
    # use_abmc = 0: one-row blocks, one colour per dependency level, the
    # device rows in level order and the factor in the original order
    info = run_check(tmp_path, builder(), abmc=False)
    assert info["BR"] == 1 and info["phases"] > 4


@pytest.mark.parametrize("abmc", [True, False], ids=["abmc", "level"])
@pytest.mark.parametrize("name", ["n1", "n1c", "diag100", "tri300", "n5"])
def test_layouts_degenerate_shapes(tmp_path, name, abmc):
    # one row; no off-diagonal entries (on the ABMC path the stragglers form
    # one oversized block of a single level); fewer rows than a wavefront
    import scipy.sparse as sp
    M = {"n1": lambda: sp.csr_matrix(np.array([[4.0]])),
         "n1c": lambda: sp.csr_matrix(np.array([[4.0 + 1.0j]])),
         "diag100": lambda: sp.diags(np.linspace(1.0, 5.0, 100)).tocsr(),
         "tri300": lambda: sp.diags([-np.ones(299), 2.0 * np.ones(300), -np.ones(299)],
                                    [-1, 0, 1]).tocsr(),
         "n5": lambda: sp.csr_matrix(np.eye(5) * 4.0 - np.eye(5, k=1) - np.eye(5, k=-1))}[name]()
    info = run_check(tmp_path, P.CSR.from_scipy(M), abmc=abmc)
    if name == "diag100" and abmc:
        assert info["big"] == 1 and info["big_rows"] == 76
