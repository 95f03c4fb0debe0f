"""Writes the golden fixtures of tests/golden/ from the reference itself.

The reference (ksugahar/ngsolve-sparsesolv) is a C++ header library; its
headers are compiled unchanged behind oracle/ref_harness.cpp into
oracle/_ref/libssref.so (``make -C oracle ref``).  This script runs that
build on small generated systems and stores inputs' identities and outputs
(ABMC arrays, IC factor, solver results) as .npz.  The inputs themselves are
regenerated by sparsesolv_amd.problems from their seeds.

Usage:  python tests/golden/make_golden.py [--all]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "ngsolve-sparsesolv_amd"))

import pyoracle  # noqa: E402
from sparsesolv_amd import problems as P  # noqa: E402


def varcoef6():
    """7-point variable-coefficient Laplacian, 6^3 cells, coefficients
    log-uniform in [1e-2, 1e2] (seed 13), uniform RHS (seed 14)."""
    A = P.laplace7_varcoef(6, 10.0 ** P.uniform(13, 216, -2.0, 2.0))
    return P.System(A, P.uniform(14, 216, -1.0, 1.0), "varcoef6", {})


def zvarcoef6():
    """varcoef6 + 0.3j I (complex symmetric), complex Gaussian RHS (seed 15)."""
    import scipy.sparse as sp
    A = varcoef6().A.to_scipy().astype(np.complex128) + sp.identity(216) * 0.3j
    z = P.gaussian(15, 2 * 216)
    return P.System(P.CSR.from_scipy(A.tocsr()), z[0::2] + 1j * z[1::2], "zvarcoef6", {})


# name -> (system builder, config overrides, solver)
CASES = {
    "lap6_abmc": (lambda: P.poisson(6), dict(use_abmc=1), "iccg"),
    "lap6_abmc_b8c6": (lambda: P.poisson(6), dict(use_abmc=1, abmc_block_size=8, abmc_num_colors=6), "iccg"),
    "lap6_path1": (lambda: P.poisson(6), dict(use_abmc=1, abmc_reorder_spmv=1), "iccg"),
    "lap6_level": (lambda: P.poisson(6), dict(use_abmc=0), "iccg"),
    "lap6_autoshift": (lambda: P.poisson(6), dict(use_abmc=1, shift_parameter=0.3, auto_shift=1), "iccg"),
    "lap6_maxiter": (lambda: P.poisson(6), dict(use_abmc=1, max_iterations=5), "iccg"),
    "cc4_autoshift": (lambda: P.curlcurl(4), dict(use_abmc=1, auto_shift=1), "iccg"),
    "eddy4_cocr": (lambda: P.eddy(4), dict(use_abmc=1, abmc_num_colors=8), "cocr"),
    "eddy4_cocg": (lambda: P.eddy(4), dict(use_abmc=1), "iccg"),
    "eddy4_plaincocr": (lambda: P.eddy(4), dict(max_iterations=400), "plain_cocr"),
    "lap6_plaincg": (lambda: P.poisson(6), dict(), "plain_cg"),
    # diagonal_scaling: the factor of S A S; the ABMC apply (Path 3, COCR)
    # runs it unscaled, Path 1 and the level path scale around it
    "vc6_scaled_abmc": (varcoef6, dict(use_abmc=1, diagonal_scaling=1), "iccg"),
    "vc6_scaled_path1": (varcoef6, dict(use_abmc=1, diagonal_scaling=1, abmc_reorder_spmv=1), "iccg"),
    "vc6_scaled_level": (varcoef6, dict(use_abmc=0, diagonal_scaling=1), "iccg"),
    "zvc6_scaled_cocr": (zvarcoef6, dict(use_abmc=1, abmc_num_colors=8, diagonal_scaling=1), "cocr"),
    # divergence_check = StagnationCount: a stop after progress (residual
    # >= 2x the best for 5 iterations running), and threshold 1.0, where
    # every improving iteration counts (the best is updated first)
    "cc6_divergence": (lambda: P.curlcurl(6), dict(use_abmc=1, auto_shift=1, divergence_check=1,
                                                  divergence_threshold=2.0, divergence_count=5), "iccg"),
    "eddy6_divergence": (lambda: P.eddy(6), dict(use_abmc=1, abmc_num_colors=8, divergence_check=1,
                                                divergence_threshold=2.0, divergence_count=5), "cocr"),
    "lap6_divergence_thr1": (lambda: P.poisson(6), dict(use_abmc=1, divergence_check=1,
                                                       divergence_threshold=1.0, divergence_count=4),
                             "iccg"),
    # SGS preconditioner (with CG / COCR) and SGS-MRTR (solve_sgsmrtr)
    "vc6_sgsmrtr": (varcoef6, dict(), "sgsmrtr"),
    "zvc6_sgsmrtr": (zvarcoef6, dict(), "sgsmrtr"),
    "cc5_sgsmrtr_maxiter": (lambda: P.curlcurl(5), dict(max_iterations=7), "sgsmrtr"),
    "eddy5_sgsmrtr": (lambda: P.eddy(5), dict(), "sgsmrtr"),
    "vc6_sgs_cg": (varcoef6, dict(), "sgs_cg"),
    "eddy5_sgs_cocr": (lambda: P.eddy(5), dict(), "sgs_cocr"),
}


def run(ref, name):
    build, kw, solver = CASES[name]
    sysm = build()
    A, b = sysm.A, sysm.b
    cfg = ref.config(save_residual_history=1, **kw)
    out = {"n": np.int64(A.n), "nnz": np.int64(A.nnz),
           "values_sum": np.array([np.sum(A.values)]).astype(np.complex128),
           "b_sum": np.array([np.sum(b)]).astype(np.complex128)}
    if solver in ("iccg", "cocr"):
        ic = ref.ic(A, cfg)
        f = ic.factor()
        for k in ("L_row_ptr", "L_col_idx", "L_values", "inv_diag"):
            out[k] = f[k]
        out["actual_shift"] = np.float64(f["actual_shift"])
        if "ordering" in f:
            for k in ("ordering", "color_offsets", "block_offsets"):
                out[k] = f[k]
        xr = np.linspace(-1.0, 1.0, A.n) * (1.0 + (0.5j if A.is_complex else 0.0))
        out["apply_x"] = xr
        out["apply_y"] = ic.apply(xr)
    if solver.startswith("sgs"):
        xr = np.linspace(-1.0, 1.0, A.n) * (1.0 + (0.5j if A.is_complex else 0.0))
        out["apply_x"] = xr
        out["apply_y"] = ref.sgs_apply(A, xr)
    if solver == "iccg":
        x, r = ref.solve_iccg(A, b, cfg)
    elif solver == "cocr":
        x, r = ref.solve_cocr(A, b, cfg)
    elif solver == "sgsmrtr":
        x, r = ref.solve_sgsmrtr(A, b, cfg)
    elif solver in ("sgs_cg", "sgs_cocr"):
        x, r = ref.solve_sgs(A, b, cfg, "COCR" if solver == "sgs_cocr" else "CG")
    elif solver == "plain_cg":
        x, r = ref.solve_plain(A, b, cfg, "CG")
    else:
        x, r = ref.solve_plain(A, b, cfg, "COCR")
    out["x"] = x
    out["converged"] = np.int64(r["converged"])
    out["iterations"] = np.int64(r["iterations"])
    out["final_residual"] = np.float64(r["final_residual"])
    out["history"] = r["residual_history"]
    return out


def main():
    # existing fixtures are kept unless --all is given
    ref = pyoracle.Oracle("ref")
    for name in CASES:
        if "--all" not in sys.argv and os.path.exists(os.path.join(HERE, name + ".npz")):
            continue
        data = run(ref, name)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **data)
        print(f"{name}: n={int(data['n'])} iterations={int(data['iterations'])} "
              f"converged={int(data['converged'])}")


if __name__ == "__main__":
    main()
