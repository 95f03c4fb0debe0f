"""ctypes binding of libsparsesolv_amd.so (declarations: include/sparsesolv_amd.h).

The library must be built (``make`` in ngsolve-sparsesolv_amd/, or
``__graft_entry__.build()``); there is no fallback implementation: loading
fails loudly when it is missing.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsparsesolv_amd.so")

SSA_OK = 0
SSA_ERR_ARGUMENT = -1
SSA_ERR_RUNTIME = -2
SSA_ERR_DEVICE = -3
SSA_ERR_UNSUPPORTED = -4

SSA_F64 = 0
SSA_C128 = 1
PRECISION_FP64 = 0
PRECISION_MIXED = 1

METHOD_ICCG = 0
METHOD_ICCOCR = 1
METHOD_CG = 2
METHOD_COCR = 3
METHOD_NGS_ICCOCR = 4
METHOD_SGSMRTR = 5

PRECOND_IC = 0
PRECOND_SGS = 1
PRECOND_SGSMRTR = 2

EXPORTED = (
    "ssa_config_default", "ssa_last_error", "ssa_context_create", "ssa_context_destroy",
    "ssa_ic_create", "ssa_ic_destroy", "ssa_ic_info_get", "ssa_ic_ordering",
    "ssa_ic_schedule", "ssa_ic_factor", "ssa_ic_apply", "ssa_spmv", "ssa_solve",
    "ssa_ic_profile", "ssa_batch_create", "ssa_batch_solve", "ssa_ic_systems",
    "ssa_solve_multi", "ssa_ic_apply_multi", "ssa_spmv_multi", "ssa_ic_profile_multi",
    "ssa_ic_set_operator", "ssa_ic_trace", "ssa_ic_trace_read", "ssa_ic_apply_noalloc", "ssa_batch_create_each",
)

PROFILE_SPMV = 0
PROFILE_FORWARD = 1
PROFILE_BACKWARD = 2
PROFILE_APPLY = 3
PROFILE_COLOUR_KERNELS = 4

TRACE_OFF = -1
TRACE_SPMV = 0
TRACE_COLOUR = 1
TRACE_UPDATE = 2
TRACE_PQ = 3
TRACE_DOT = 4


class Config(C.Structure):
    """ssa_config (mirror of sparsesolv::SolverConfig plus GPU fields)."""
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
        ("abmc_use_rcm", C.c_int), ("precision", C.c_int), ("use_graph", C.c_int),
        ("poll_interval", C.c_int), ("walk_overlap", C.c_int),
        ("preconditioner", C.c_int), ("rhs_block", C.c_int),
    ]


class Result(C.Structure):
    _fields_ = [("converged", C.c_int), ("iterations", C.c_int),
                ("final_residual", C.c_double), ("history_len", C.c_int),
                ("status", C.c_int), ("actual_shift", C.c_double),
                ("solve_seconds", C.c_double)]


class IcInfo(C.Structure):
    _fields_ = [("n", C.c_int), ("nnz", C.c_int), ("nnz_l", C.c_int),
                ("num_colors", C.c_int), ("num_blocks", C.c_int),
                ("rows_per_lane", C.c_int), ("big_blocks", C.c_int),
                ("max_big_rows", C.c_int), ("max_levels", C.c_int),
                ("shift_trials", C.c_int), ("actual_shift", C.c_double),
                ("setup_seconds", C.c_double), ("device_bytes", C.c_size_t),
                ("factor_bytes", C.c_size_t), ("spmv_bytes", C.c_size_t),
                ("walk_overlap", C.c_int)]


class SsaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


_lib = None
_lock = threading.Lock()


def load() -> C.CDLL:
    """Load the library (once).  Raises ImportError when it is not built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `make -C ngsolve-sparsesolv_amd` "
                "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        L.ssa_config_default.argtypes = [C.POINTER(Config)]
        L.ssa_last_error.restype = C.c_char_p
        L.ssa_context_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.ssa_context_destroy.argtypes = [vp]
        L.ssa_ic_create.argtypes = [vp, C.c_int, C.c_int, i32p, i32p, vp, C.POINTER(Config),
                                    C.POINTER(vp)]
        L.ssa_ic_destroy.argtypes = [vp]
        L.ssa_ic_info_get.argtypes = [vp, C.POINTER(IcInfo)]
        L.ssa_ic_ordering.argtypes = [vp, i32p]
        L.ssa_ic_schedule.argtypes = [vp, vp, vp]
        L.ssa_ic_factor.argtypes = [vp, vp, vp, vp, vp]
        L.ssa_ic_apply.argtypes = [vp, vp, vp, C.c_int]
        L.ssa_spmv.argtypes = [vp, vp, vp, C.c_int]
        L.ssa_ic_set_operator.argtypes = [vp, C.c_int, i32p, i32p, vp]
        L.ssa_solve.argtypes = [vp, C.c_int, vp, vp, C.c_int, C.POINTER(Config),
                                C.POINTER(Result), vp]
        L.ssa_ic_profile.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]
        L.ssa_batch_create_each.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp, vp, C.POINTER(Config),
                                            vp, vp, C.POINTER(vp)]
        L.ssa_batch_create.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp, vp, C.POINTER(Config),
                                       C.POINTER(vp)]
        L.ssa_batch_solve.argtypes = [vp, C.c_int, vp, vp, C.c_int, C.POINTER(Config),
                                      C.POINTER(Result)]
        L.ssa_ic_systems.argtypes = [vp, C.POINTER(C.c_int)]
        L.ssa_solve_multi.argtypes = [vp, C.c_int, C.c_int, vp, vp, C.c_int, C.POINTER(Config),
                                      C.POINTER(Result)]
        L.ssa_ic_apply_multi.argtypes = [vp, C.c_int, vp, vp, C.c_int]
        L.ssa_spmv_multi.argtypes = [vp, C.c_int, vp, vp, C.c_int]
        L.ssa_ic_profile_multi.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double),
                                           C.POINTER(C.c_double)]
        L.ssa_ic_apply_noalloc.argtypes = [vp, vp, vp]
        L.ssa_ic_trace.argtypes = [vp, C.c_int]
        L.ssa_ic_trace_read.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_double),
                                        C.POINTER(C.c_double)]
        for name in EXPORTED:
            getattr(L, name).restype = C.c_char_p if name == "ssa_last_error" else (
                None if name == "ssa_config_default" else C.c_int)
        _lib = L
        return L


def check(rc: int) -> None:
    if rc != SSA_OK:
        msg = load().ssa_last_error().decode(errors="replace")
        raise SsaError(rc, msg)


def default_config(**kw) -> Config:
    c = Config()
    load().ssa_config_default(C.byref(c))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, int(v) if isinstance(v, bool) else v)
    return c


class Context:
    """A device context (HIP stream on one GPU)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.ssa_context_create(device, C.byref(h)))
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            self.lib.ssa_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = {}


def context(device: int = 0) -> Context:
    with _lock:
        ctx = _default_ctx.get(device)
    if ctx is None:
        ctx = Context(device)
        with _lock:
            _default_ctx[device] = ctx
    return ctx


class IcHandle:
    """ssa_ic: the operator and IC(0) factor resident on the GPU."""

    def __init__(self, A, cfg: Config, ctx: Context | None = None, host_only: bool = False):
        self.ctx = None if host_only else (ctx or context(0))
        self.lib = load()
        self.n = int(A.n)
        self.is_complex = bool(A.is_complex)
        self.dtype = np.complex128 if self.is_complex else np.float64
        vals = np.ascontiguousarray(A.values, dtype=self.dtype)
        rp = np.ascontiguousarray(A.row_ptr, dtype=np.int32)
        ci = np.ascontiguousarray(A.col_idx, dtype=np.int32)
        h = C.c_void_p()
        check(self.lib.ssa_ic_create(self.ctx.h if self.ctx else None, SSA_C128 if self.is_complex else SSA_F64,
                                     self.n, rp, ci, vals.ctypes.data, C.byref(cfg), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.ssa_ic_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        i = IcInfo()
        check(self.lib.ssa_ic_info_get(self.h, C.byref(i)))
        return {k: getattr(i, k) for k, _ in IcInfo._fields_}

    def ordering(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.int32)
        check(self.lib.ssa_ic_ordering(self.h, out))
        return out

    def schedule(self):
        inf = self.info()
        co = np.empty(inf["num_colors"] + 1, dtype=np.int32)
        bo = np.empty(inf["num_blocks"] + 1, dtype=np.int32)
        check(self.lib.ssa_ic_schedule(self.h, co.ctypes.data, bo.ctypes.data))
        return co, bo

    def factor(self) -> dict:
        inf = self.info()
        rp = np.empty(self.n + 1, dtype=np.int32)
        ci = np.empty(inf["nnz_l"], dtype=np.int32)
        v = np.empty(inf["nnz_l"], dtype=self.dtype)
        d = np.empty(self.n, dtype=self.dtype)
        check(self.lib.ssa_ic_factor(self.h, rp.ctypes.data, ci.ctypes.data, v.ctypes.data,
                                     d.ctypes.data))
        return {"L_row_ptr": rp, "L_col_idx": ci, "L_values": v, "inv_diag": d,
                "actual_shift": inf["actual_shift"]}

    def _vec(self, x):
        x = np.ascontiguousarray(x, dtype=self.dtype)
        if x.shape != (self.n,):
            raise ValueError(f"vector of length {self.n} expected, got {x.shape}")
        return x

    def apply(self, x) -> np.ndarray:
        x = self._vec(x)
        y = np.empty_like(x)
        check(self.lib.ssa_ic_apply(self.h, x.ctypes.data, y.ctypes.data, 0))
        return y

    def spmv(self, x) -> np.ndarray:
        x = self._vec(x)
        y = np.empty_like(x)
        check(self.lib.ssa_spmv(self.h, x.ctypes.data, y.ctypes.data, 0))
        return y

    def set_operator(self, A) -> None:
        """The operator of this handle's solves and SpMV becomes A (same order
        and dtype); the IC factor stays."""
        if A.is_complex != self.is_complex:
            raise TypeError("set_operator: the operator's dtype must match the handle's")
        vals = np.ascontiguousarray(A.values, dtype=self.dtype)
        rp = np.ascontiguousarray(A.row_ptr, dtype=np.int32)
        ci = np.ascontiguousarray(A.col_idx, dtype=np.int32)
        check(self.lib.ssa_ic_set_operator(self.h, A.n, rp, ci, vals.ctypes.data))

    def apply_device(self, x_ptr: int, y_ptr: int) -> None:
        check(self.lib.ssa_ic_apply(self.h, C.c_void_p(x_ptr), C.c_void_p(y_ptr), 1))

    def spmv_device(self, x_ptr: int, y_ptr: int) -> None:
        check(self.lib.ssa_spmv(self.h, C.c_void_p(x_ptr), C.c_void_p(y_ptr), 1))

    def solve(self, method: int, b, x0, cfg: Config):
        b = self._vec(b)
        x = np.array(self._vec(np.zeros(self.n) if x0 is None else x0), dtype=self.dtype)
        hist = np.zeros(max(cfg.max_iterations, 0) + 4, dtype=np.float64)
        res = Result()
        check(self.lib.ssa_solve(self.h, method, b.ctypes.data, x.ctypes.data, 0, C.byref(cfg),
                                 C.byref(res), hist.ctypes.data))
        return x, res, hist[:res.history_len].copy()

    def profile(self, what: int, reps: int = 20):
        """(mean microseconds per call, algorithmic bytes per call), timed with
        HIP events on the library's stream."""
        us, by = C.c_double(), C.c_double()
        check(self.lib.ssa_ic_profile(self.h, what, reps, C.byref(us), C.byref(by)))
        return us.value, by.value

    def solve_device(self, method: int, b_ptr: int, x_ptr: int, cfg: Config,
                     history: np.ndarray | None = None) -> Result:
        res = Result()
        hp = history.ctypes.data if history is not None else None
        check(self.lib.ssa_solve(self.h, method, C.c_void_p(b_ptr), C.c_void_p(x_ptr), 1,
                                 C.byref(cfg), C.byref(res), hp))
        return res

    # ---- several right-hand sides (rows of a (nrhs, n) array) -------------
    def _block(self, X):
        X = np.ascontiguousarray(X, dtype=self.dtype)
        if X.ndim != 2 or X.shape[1] != self.n or X.shape[0] < 1:
            raise ValueError(f"(nrhs, {self.n}) array expected, got {X.shape}")
        return X

    def solve_multi(self, method: int, B, X0, cfg: Config):
        """ssa_solve_multi: X (nrhs, n) and one Result per right-hand side."""
        B = self._block(B)
        X = np.array(self._block(np.zeros_like(B) if X0 is None else X0), dtype=self.dtype)
        res = (Result * B.shape[0])()
        check(self.lib.ssa_solve_multi(self.h, method, B.shape[0], B.ctypes.data, X.ctypes.data, 0,
                                       C.byref(cfg), res))
        return X, list(res)

    def solve_multi_device(self, method: int, nrhs: int, b_ptr: int, x_ptr: int, cfg: Config):
        res = (Result * nrhs)()
        check(self.lib.ssa_solve_multi(self.h, method, nrhs, C.c_void_p(b_ptr), C.c_void_p(x_ptr), 1,
                                       C.byref(cfg), res))
        return list(res)

    def apply_multi(self, X) -> np.ndarray:
        X = self._block(X)
        Y = np.empty_like(X)
        check(self.lib.ssa_ic_apply_multi(self.h, X.shape[0], X.ctypes.data, Y.ctypes.data, 0))
        return Y

    def spmv_multi(self, X) -> np.ndarray:
        X = self._block(X)
        Y = np.empty_like(X)
        check(self.lib.ssa_spmv_multi(self.h, X.shape[0], X.ctypes.data, Y.ctypes.data, 0))
        return Y

    def profile_multi(self, what: int, nrhs: int, reps: int = 20):
        """profile() for a block of nrhs right-hand sides."""
        us, by = C.c_double(), C.c_double()
        check(self.lib.ssa_ic_profile_multi(self.h, what, nrhs, reps, C.byref(us), C.byref(by)))
        return us.value, by.value

    def trace(self, what: int) -> None:
        """ssa_ic_trace: time every launch of one kernel class inside the
        following solves (TRACE_OFF stops); the totals restart from zero."""
        check(self.lib.ssa_ic_trace(self.h, what))

    def trace_read(self):
        """(launches, total microseconds, total algorithmic bytes) timed so far."""
        k, us, by = C.c_longlong(), C.c_double(), C.c_double()
        check(self.lib.ssa_ic_trace_read(self.h, C.byref(k), C.byref(us), C.byref(by)))
        return k.value, us.value, by.value


class BatchHandle(IcHandle):
    """Several independent systems in one handle (ssa_batch_create).  Vectors
    hold the systems back to back, each in its own ordering."""

    def __init__(self, systems, cfg: Config, ctx: Context | None = None, host_only: bool = False,
                 block_sizes=None, num_colors=None):
        """block_sizes / num_colors: an ABMC block size / colour count per
        system (ssa_batch_create_each) instead of cfg's."""
        self.ctx = None if host_only else (ctx or context(0))
        self.lib = load()
        mats = list(systems)
        if not mats:
            raise ValueError("no systems")
        self.is_complex = bool(mats[0].is_complex)
        if any(bool(A.is_complex) != self.is_complex for A in mats):
            raise TypeError("all systems must have the same scalar type")
        self.dtype = np.complex128 if self.is_complex else np.float64
        self.sizes = [int(A.n) for A in mats]
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        self.n = int(self.offsets[-1])
        self.count = len(mats)
        self._keep = []
        rps, cis, vals = [], [], []
        for A in mats:
            rp = np.ascontiguousarray(A.row_ptr, dtype=np.int32)
            ci = np.ascontiguousarray(A.col_idx, dtype=np.int32)
            v = np.ascontiguousarray(A.values, dtype=self.dtype)
            self._keep += [rp, ci, v]
            rps.append(rp.ctypes.data)
            cis.append(ci.ctypes.data)
            vals.append(v.ctypes.data)
        ns = np.array(self.sizes, dtype=np.int32)
        arr = C.c_void_p * self.count
        h = C.c_void_p()
        if block_sizes is None and num_colors is None:
            check(self.lib.ssa_batch_create(self.ctx.h if self.ctx else None,
                                            SSA_C128 if self.is_complex else SSA_F64, self.count,
                                            ns.ctypes.data, arr(*rps), arr(*cis), arr(*vals),
                                            C.byref(cfg), C.byref(h)))
        else:
            def per_system(v):
                if v is None:
                    return None, None
                a = np.ascontiguousarray(v, dtype=np.int32)
                if a.shape != (self.count,):
                    raise ValueError("one block size / colour count per system expected")
                return a, a.ctypes.data
            bs, bsp = per_system(block_sizes)
            nc, ncp = per_system(num_colors)
            check(self.lib.ssa_batch_create_each(self.ctx.h if self.ctx else None,
                                                 SSA_C128 if self.is_complex else SSA_F64, self.count,
                                                 ns.ctypes.data, arr(*rps), arr(*cis), arr(*vals),
                                                 C.byref(cfg), bsp, ncp, C.byref(h)))
        self.h = h
        self._keep = []

    def split(self, v):
        """Per-system views of a concatenated vector."""
        return [v[self.offsets[k]:self.offsets[k + 1]] for k in range(self.count)]

    def solve_batch(self, method: int, b, x0, cfg: Config):
        """(x, [Result per system]) for concatenated b / x0 (x0 None: zeros)."""
        b = self._vec(b)
        x = np.array(self._vec(np.zeros(self.n) if x0 is None else x0), dtype=self.dtype)
        res = (Result * self.count)()
        check(self.lib.ssa_batch_solve(self.h, method, b.ctypes.data, x.ctypes.data, 0,
                                       C.byref(cfg), res))
        return x, list(res)

    def solve_batch_device(self, method: int, b_ptr: int, x_ptr: int, cfg: Config):
        res = (Result * self.count)()
        check(self.lib.ssa_batch_solve(self.h, method, C.c_void_p(b_ptr), C.c_void_p(x_ptr), 1,
                                       C.byref(cfg), res))
        return list(res)

    def systems(self) -> int:
        c = C.c_int()
        check(self.lib.ssa_ic_systems(self.h, C.byref(c)))
        return c.value
