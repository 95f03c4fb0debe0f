"""The C ABI library loads, exports every function include/sparsesolv_amd.h
declares, and its defaults are the reference's SolverConfig defaults.
No GPU work here."""
import ctypes as C
import os
import re

import pytest

from sparsesolv_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sparsesolv_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(ssa_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "ssa_solve" in names and "ssa_ic_create" in names and len(names) >= 13


def test_library_exports_every_declared_function():
    lib = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared_functions()) <= set(_lib.EXPORTED)


def test_config_defaults_match_reference():
    # sparsesolv::SolverConfig defaults (core/solver_config.hpp:39-157)
    c = _lib.default_config()
    assert c.tolerance == 1e-10 and c.max_iterations == 1000 and c.abs_tolerance == 1e-12
    assert c.norm_type == 0 and c.custom_norm == 1.0 and c.shift_parameter == 1.05
    assert c.diagonal_scaling == 0 and c.auto_shift == 0 and c.shift_increment == 0.05
    assert c.max_shift_value == 5.0 and c.min_diagonal_threshold == 1e-6
    assert c.zero_diagonal_replacement == 1e-10 and c.max_shift_trials == 100
    assert c.divergence_check == 0 and c.divergence_threshold == 1000.0
    assert c.divergence_count == 100 and c.save_best_result == 1
    assert c.save_residual_history == 0 and c.conjugate == 0 and c.use_abmc == 0
    assert c.abmc_block_size == 4 and c.abmc_num_colors == 4
    assert c.abmc_reorder_spmv == 0 and c.abmc_use_rcm == 0


def test_config_struct_layout_matches_oracle_prefix():
    import pyoracle
    a = [f for f, _ in _lib.Config._fields_]
    b = [f for f, _ in pyoracle.OrcConfig._fields_]
    assert a[:len(b)] == b


def test_null_arguments_are_rejected():
    lib = _lib.load()
    assert lib.ssa_ic_apply(None, None, None, 0) == _lib.SSA_ERR_ARGUMENT
    assert "null" in lib.ssa_last_error().decode()
    out = C.c_void_p()
    assert lib.ssa_context_create(-1, C.byref(out)) != _lib.SSA_OK
