"""ctypes wrapper for the CPU checker libraries (test infrastructure only).

``Oracle("oracle")`` loads oracle/liboracle.so, the C restatement of the
reference path; ``Oracle("ref")`` loads oracle/_ref/libssref.so, the
reference headers compiled behind ref_harness.cpp (present only where the
reference sources were available at build time); ``Oracle("ref_omp")`` the
OpenMP build of the same.  Both export the functions declared in oracle.h.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBS = {
    "oracle": os.path.join(HERE, "liboracle.so"),
    "ref": os.path.join(HERE, "_ref", "libssref.so"),
    "ref_omp": os.path.join(HERE, "_ref", "libssref_omp.so"),
}

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class OrcConfig(C.Structure):
    """Mirror of orc_config in oracle.h (sparsesolv::SolverConfig)."""
    _fields_ = [
        ("tolerance", C.c_double), ("max_iterations", C.c_int),
        ("abs_tolerance", C.c_double), ("norm_type", C.c_int),
        ("custom_norm", C.c_double), ("shift_parameter", C.c_double),
        ("diagonal_scaling", C.c_int), ("auto_shift", C.c_int),
        ("shift_increment", C.c_double), ("max_shift_value", C.c_double),
        ("min_diagonal_threshold", C.c_double), ("zero_diagonal_replacement", C.c_double),
        ("max_shift_trials", C.c_int), ("divergence_check", C.c_int),
        ("divergence_threshold", C.c_double), ("divergence_count", C.c_int),
        ("save_best_result", C.c_int), ("save_residual_history", C.c_int),
        ("conjugate", C.c_int), ("use_abmc", C.c_int), ("abmc_block_size", C.c_int),
        ("abmc_num_colors", C.c_int), ("abmc_reorder_spmv", C.c_int),
        ("abmc_use_rcm", C.c_int),
    ]


class OrcResult(C.Structure):
    _fields_ = [("converged", C.c_int), ("iterations", C.c_int),
                ("final_residual", C.c_double), ("history_len", C.c_int),
                ("actual_shift", C.c_double)]


def available(kind: str) -> bool:
    return os.path.exists(LIBS[kind])


class Oracle:
    def __init__(self, kind: str = "oracle"):
        path = LIBS[kind]
        if not os.path.exists(path):
            raise FileNotFoundError(f"checker library not built: {path}")
        self.kind = kind
        L = C.CDLL(path)
        self.lib = L
        vp = C.c_void_p
        L.orc_config_default.argtypes = [C.POINTER(OrcConfig)]
        L.orc_abmc_build.restype = vp
        L.orc_abmc_build.argtypes = [C.c_int, _i32p, _i32p, C.c_int, C.c_int]
        L.orc_abmc_count.argtypes = [vp, C.c_int]
        L.orc_abmc_copy.argtypes = [vp, C.c_int, _i32p]
        L.orc_abmc_free.argtypes = [vp]
        L.orc_ic_setup.restype = vp
        L.orc_ic_setup.argtypes = [C.c_int, C.c_int, _i32p, _i32p, vp,
                                   C.POINTER(OrcConfig), C.c_char_p, C.c_int]
        L.orc_ic_count.argtypes = [vp, C.c_int]
        L.orc_ic_actual_shift.restype = C.c_double
        L.orc_ic_actual_shift.argtypes = [vp]
        L.orc_ic_copy_int.argtypes = [vp, C.c_int, _i32p]
        L.orc_ic_copy_val.argtypes = [vp, C.c_int, vp]
        L.orc_ic_apply.argtypes = [vp, vp, vp]
        L.orc_ic_apply_reordered.argtypes = [vp, vp, vp]
        L.orc_ic_free.argtypes = [vp]
        sargs = [C.c_int, C.c_int, _i32p, _i32p, vp, vp, vp, C.POINTER(OrcConfig),
                 C.POINTER(OrcResult), vp, C.c_char_p, C.c_int]
        L.orc_solve_iccg.argtypes = sargs
        L.orc_solve_cocr.argtypes = sargs
        L.orc_solve_cocr_pre.argtypes = [C.c_int, C.c_int, _i32p, _i32p, vp, _i32p, _i32p, vp,
                                         vp, vp, C.POINTER(OrcConfig), C.POINTER(OrcResult), vp,
                                         C.c_char_p, C.c_int]
        L.orc_solve_plain.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p, vp, vp, vp,
                                      C.POINTER(OrcConfig), C.POINTER(OrcResult), vp,
                                      C.c_char_p, C.c_int]
        L.orc_sgs_apply.argtypes = [C.c_int, C.c_int, _i32p, _i32p, vp, vp, vp]
        L.orc_solve_sgs.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p, vp, vp, vp,
                                    C.POINTER(OrcConfig), C.POINTER(OrcResult), vp,
                                    C.c_char_p, C.c_int]
        L.orc_solve_sgsmrtr.argtypes = sargs
        L.orc_time_iterations.restype = C.c_double
        L.orc_time_iterations.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p, vp, vp,
                                          C.POINTER(OrcConfig), C.c_int,
                                          C.POINTER(C.c_double)]
        self.has_threads = hasattr(L, "orc_num_threads")

    # -- config ------------------------------------------------------------
    def config(self, **kw) -> OrcConfig:
        c = OrcConfig()
        self.lib.orc_config_default(C.byref(c))
        for k, v in kw.items():
            if not hasattr(c, k):
                raise KeyError(k)
            setattr(c, k, v)
        return c

    # -- ABMC --------------------------------------------------------------
    def abmc(self, A, block_size: int, colors: int) -> dict:
        h = self.lib.orc_abmc_build(A.n, A.row_ptr, A.col_idx, block_size, colors)
        try:
            nc = self.lib.orc_abmc_count(h, 0)
            nb = self.lib.orc_abmc_count(h, 1)
            out = {}
            for which, name, size in ((0, "color_offsets", nc + 1), (1, "color_blocks", nb),
                                      (2, "block_offsets", nb + 1), (3, "ordering", A.n),
                                      (4, "reverse_ordering", A.n)):
                arr = np.empty(size, dtype=np.int32)
                if size:
                    self.lib.orc_abmc_copy(h, which, arr)
                out[name] = arr
            return out
        finally:
            self.lib.orc_abmc_free(h)

    # -- IC ----------------------------------------------------------------
    def ic(self, A, cfg: OrcConfig) -> "IC":
        return IC(self, A, cfg)

    # -- solvers -----------------------------------------------------------
    def _solve(self, fn, A, b, cfg, x0=None, extra=()):
        dt = np.complex128 if A.is_complex else np.float64
        vals = np.ascontiguousarray(A.values, dtype=dt)
        b = np.ascontiguousarray(b, dtype=dt)
        x = np.zeros(A.n, dtype=dt) if x0 is None else np.array(x0, dtype=dt)
        hist = np.zeros(cfg.max_iterations + 4, dtype=np.float64)
        res = OrcResult()
        err = C.create_string_buffer(512)
        rc = fn(*extra, 1 if A.is_complex else 0, A.n, A.row_ptr, A.col_idx,
                vals.ctypes.data, b.ctypes.data, x.ctypes.data, C.byref(cfg),
                C.byref(res), hist.ctypes.data, err, 512)
        if rc != 0:
            raise RuntimeError(err.value.decode())
        return x, {"converged": bool(res.converged), "iterations": res.iterations,
                   "final_residual": res.final_residual, "actual_shift": res.actual_shift,
                   "residual_history": hist[:res.history_len].copy()}

    def solve_iccg(self, A, b, cfg, x0=None):
        return self._solve(self.lib.orc_solve_iccg, A, b, cfg, x0)

    def solve_cocr(self, A, b, cfg, x0=None):
        return self._solve(self.lib.orc_solve_cocr, A, b, cfg, x0)

    def solve_cocr_pre(self, A, M, b, cfg, x0=None):
        """COCR on A with the IC preconditioner set up on M (same order)."""
        if M.n != A.n or M.is_complex != A.is_complex:
            raise ValueError("the preconditioner's matrix must match the operator's order and type")
        dt = np.complex128 if A.is_complex else np.float64
        mv = np.ascontiguousarray(M.values, dtype=dt)
        f = self.lib.orc_solve_cocr_pre
        def fn(dtype, n, rp, ci, vals, *rest):
            return f(dtype, n, rp, ci, vals, M.row_ptr, M.col_idx, mv.ctypes.data, *rest)
        return self._solve(fn, A, b, cfg, x0)

    def solve_plain(self, A, b, cfg, method: str, x0=None):
        f = self.lib.orc_solve_plain
        def fn(dtype, *rest):
            return f(dtype, 1 if method.upper() == "COCR" else 0, *rest)
        return self._solve(fn, A, b, cfg, x0)

    def sgs_apply(self, A, x):
        """y = M^{-1} x of the SGS preconditioner set up on A."""
        dt = np.complex128 if A.is_complex else np.float64
        vals = np.ascontiguousarray(A.values, dtype=dt)
        x = np.ascontiguousarray(x, dtype=dt)
        y = np.empty_like(x)
        self.lib.orc_sgs_apply(1 if A.is_complex else 0, A.n, A.row_ptr, A.col_idx,
                               vals.ctypes.data, x.ctypes.data, y.ctypes.data)
        return y

    def solve_sgs(self, A, b, cfg, method: str = "CG", x0=None):
        """CG / COCR on A preconditioned by the SGS preconditioner of A."""
        f = self.lib.orc_solve_sgs
        def fn(dtype, *rest):
            return f(dtype, 1 if method.upper() == "COCR" else 0, *rest)
        return self._solve(fn, A, b, cfg, x0)

    def solve_sgsmrtr(self, A, b, cfg, x0=None):
        return self._solve(self.lib.orc_solve_sgsmrtr, A, b, cfg, x0)

    def time_iterations(self, A, b, cfg, iters: int, method: str = "CG"):
        dt = np.complex128 if A.is_complex else np.float64
        vals = np.ascontiguousarray(A.values, dtype=dt)
        b = np.ascontiguousarray(b, dtype=dt)
        setup = C.c_double(0.0)
        t = self.lib.orc_time_iterations(1 if A.is_complex else 0,
                                         1 if method.upper() == "COCR" else 0, A.n,
                                         A.row_ptr, A.col_idx, vals.ctypes.data,
                                         b.ctypes.data, C.byref(cfg), iters, C.byref(setup))
        return t, setup.value

    def num_threads(self) -> int:
        return self.lib.orc_num_threads() if self.has_threads else 1

    def set_threads(self, n: int) -> None:
        """Threads of the library's later OpenMP regions (OpenMP builds only)."""
        if hasattr(self.lib, "orc_set_threads"):
            self.lib.orc_set_threads(int(n))


class IC:
    """An IC(0) preconditioner built by the checker library."""

    def __init__(self, orc: Oracle, A, cfg: OrcConfig):
        self.orc = orc
        self.n = A.n
        self.dtype = np.complex128 if A.is_complex else np.float64
        self._vals = np.ascontiguousarray(A.values, dtype=self.dtype)
        err = C.create_string_buffer(512)
        self.h = orc.lib.orc_ic_setup(1 if A.is_complex else 0, A.n, A.row_ptr, A.col_idx,
                                      self._vals.ctypes.data, C.byref(cfg), err, 512)
        if not self.h:
            raise RuntimeError(err.value.decode())

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.orc.lib.orc_ic_free(h)
            self.h = None

    def count(self, which: int) -> int:
        return self.orc.lib.orc_ic_count(self.h, which)

    @property
    def actual_shift(self) -> float:
        return self.orc.lib.orc_ic_actual_shift(self.h)

    def _int(self, which, size):
        a = np.empty(size, dtype=np.int32)
        if size:
            self.orc.lib.orc_ic_copy_int(self.h, which, a)
        return a

    def _val(self, which, size, dt=None):
        a = np.empty(size, dtype=dt or self.dtype)
        if size:
            self.orc.lib.orc_ic_copy_val(self.h, which, a.ctypes.data)
        return a

    def factor(self) -> dict:
        n, nnz = self.n, self.count(1)
        nc, nb = self.count(2), self.count(3)
        out = {"L_row_ptr": self._int(0, n + 1), "L_col_idx": self._int(1, nnz),
               "Lt_row_ptr": self._int(2, n + 1), "Lt_col_idx": self._int(3, nnz),
               "L_values": self._val(0, nnz), "Lt_values": self._val(1, nnz),
               "inv_diag": self._val(2, n), "actual_shift": self.actual_shift}
        if nc:
            out["ordering"] = self._int(4, n)
            out["color_offsets"] = self._int(5, nc + 1)
            out["block_offsets"] = self._int(6, nb + 1)
        return out

    def apply(self, x):
        x = np.ascontiguousarray(x, dtype=self.dtype)
        y = np.empty_like(x)
        self.orc.lib.orc_ic_apply(self.h, x.ctypes.data, y.ctypes.data)
        return y

    def apply_reordered(self, x):
        x = np.ascontiguousarray(x, dtype=self.dtype)
        y = np.empty_like(x)
        self.orc.lib.orc_ic_apply_reordered(self.h, x.ctypes.data, y.ctypes.data)
        return y
