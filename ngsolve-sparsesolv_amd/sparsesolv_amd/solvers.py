"""Python drop-in for the sparsesolv_ngsolve factories on this path.

Mirrors ngsolve/sparsesolv_python_export.hpp: ``SparseSolvSolver`` (:242-329),
``ICPreconditioner`` (:176-207), ``SGSPreconditioner`` (:82-88, :209-237) and
``COCRSolver`` (:577-632), with the same
keyword arguments, defaults, properties and ``SparseSolvResult`` fields.  The
matrices are CSR (``problems.CSR`` or any scipy sparse matrix) instead of
NGSolve ``SparseMatrix``; ``freedofs`` is a boolean array instead of a
``BitArray``; vectors are numpy arrays.  Real/complex dispatch follows
``mat.IsComplex()`` as in the reference factories, returning the typed
``*D`` / ``*C`` classes.

Every solve and preconditioner application runs on the GPU through
libsparsesolv_amd.so.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import _lib
from .problems import CSR

__all__ = ["SparseSolvSolver", "ICPreconditioner", "SGSPreconditioner", "COCRSolver",
           "SparseSolvResult", "SparseSolvSolverD", "SparseSolvSolverC", "ICPreconditionerD",
           "ICPreconditionerC", "SGSPreconditionerD", "SGSPreconditionerC", "COCRSolverD",
           "COCRSolverC"]


class SparseSolvResult:
    """sparsesolv::SolverResult as exported to Python (python_export.hpp:29-49)."""

    def __init__(self, converged=False, iterations=0, final_residual=0.0,
                 residual_history=None):
        self.converged = bool(converged)
        self.iterations = int(iterations)
        self.final_residual = float(final_residual)
        self.residual_history = list(residual_history if residual_history is not None else [])

    def __repr__(self):
        return (f"SparseSolvResult(converged={'True' if self.converged else 'False'}, "
                f"iterations={self.iterations}, residual={self.final_residual:f})")


def _as_csr(mat) -> CSR:
    if isinstance(mat, CSR):
        return mat
    if sp.issparse(mat):
        return CSR.from_scipy(mat)
    raise TypeError("expected a sparse matrix (scipy.sparse or problems.CSR)")


def _words_hash(a: np.ndarray) -> int:
    """A position-weighted sum of an array's 32-bit words (wrapping 64-bit
    arithmetic)."""
    w = np.ascontiguousarray(a).view(np.uint32).astype(np.uint64)
    k = np.arange(1, w.size + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        return int(np.sum((w + np.uint64(0x632BE59BD9B4E019)) * k, dtype=np.uint64))


def _fingerprint(A: CSR) -> tuple:
    """Content identity of a matrix (pattern and values): whether an operator
    is the one a handle holds, or changed in place since."""
    return (A.n, _words_hash(A.row_ptr), _words_hash(A.col_idx), _words_hash(A.values))


def _with_freedofs(A: CSR, freedofs) -> CSR:
    """BuildSparseMatrixView (ngsolve/sparsesolv_precond.hpp:35-81): constrained
    rows become identity rows, constrained columns of free rows are zeroed;
    the sparsity pattern is kept."""
    if freedofs is None:
        return A
    free = np.asarray(freedofs, dtype=bool)
    if free.shape != (A.n,):
        raise ValueError("freedofs must have one entry per row")
    rows = np.repeat(np.arange(A.n), np.diff(A.row_ptr))
    v = A.values.copy()
    fixed_row = ~free[rows]
    v[fixed_row] = 0
    v[fixed_row & (A.col_idx == rows)] = 1
    v[free[rows] & ~free[A.col_idx]] = 0
    return CSR(A.n, A.row_ptr, A.col_idx, v)


class _GpuConfigMixin:
    """Fields of sparsesolv::SolverConfig exposed as properties."""

    def _cfg(self, **extra) -> _lib.Config:
        c = self._config
        kw = dict(c)
        kw.update(extra)
        return _lib.default_config(**kw)


# ---------------------------------------------------------------------------
# Preconditioners
# ---------------------------------------------------------------------------
class _DevicePreconditioner:
    """SparseSolvPrecondBase (ngsolve/sparsesolv_precond.hpp:88-176): the
    matrix with its freedofs, Update(), Mult / MultAdd; the factor lives in a
    device handle built from config()."""
    _dtype = np.float64

    def __init__(self, mat, freedofs=None):
        self.mat = _as_csr(mat)
        self.freedofs = None if freedofs is None else np.array(freedofs, dtype=bool)
        self._handle = None
        self._stale = True

    def config(self) -> _lib.Config:
        raise NotImplementedError

    def Update(self):
        """Rebuild the factor with the current properties (Update(), :188-191)."""
        if self._handle is not None:
            self._handle.close()
        M = _with_freedofs(self.mat, self.freedofs)
        self._handle = _lib.IcHandle(M, self.config())
        self._operator = _fingerprint(M)   # the operator the handle holds
        self._stale = False

    @property
    def handle(self) -> _lib.IcHandle:
        if self._handle is None or self._stale:
            self.Update()
        return self._handle

    def _apply(self, x):
        x = np.asarray(x, dtype=self._dtype)
        # a (k, n) array: k vectors in one block application
        return self.handle.apply_multi(x) if x.ndim == 2 else self.handle.apply(x)

    def Mult(self, x, y=None):
        """y = M^{-1} x, constrained entries zeroed (Mult, :123-136)."""
        out = self._apply(x)
        if self.freedofs is not None:
            out[..., ~self.freedofs] = 0
        if y is not None:
            y[...] = out
            return y
        return out

    def MultAdd(self, s, x, y):
        """y += s * M^{-1} x on free entries (MultAdd, :138-160)."""
        out = self._apply(x)
        if self.freedofs is not None:
            out[..., ~self.freedofs] = 0
        y += s * out
        return y

    def __mul__(self, x):
        return self.Mult(x)

    @property
    def height(self):
        return self.mat.n

    @property
    def width(self):
        return self.mat.n


class _ICPreconditioner(_DevicePreconditioner):
    """SparseSolvICPreconditioner (ngsolve/sparsesolv_precond.hpp:178-224)."""

    def __init__(self, mat, freedofs=None, shift: float = 1.05):
        super().__init__(mat, freedofs)
        self._shift = float(shift)
        self._use_abmc = False
        self._abmc_block_size = 4
        self._abmc_num_colors = 4
        self._diagonal_scaling = False

    # properties (python_export.hpp:63-80)
    shift = property(lambda s: s._shift, lambda s, v: s._set("_shift", float(v)))
    use_abmc = property(lambda s: s._use_abmc, lambda s, v: s._set("_use_abmc", bool(v)))
    abmc_block_size = property(lambda s: s._abmc_block_size,
                               lambda s, v: s._set("_abmc_block_size", int(v)))
    abmc_num_colors = property(lambda s: s._abmc_num_colors,
                               lambda s, v: s._set("_abmc_num_colors", int(v)))
    diagonal_scaling = property(lambda s: s._diagonal_scaling,
                                lambda s, v: s._set("_diagonal_scaling", bool(v)))

    def _set(self, name, value):
        setattr(self, name, value)

    def config(self) -> _lib.Config:
        return _lib.default_config(shift_parameter=self._shift, use_abmc=self._use_abmc,
                                   abmc_block_size=self._abmc_block_size,
                                   abmc_num_colors=self._abmc_num_colors,
                                   diagonal_scaling=self._diagonal_scaling)


class ICPreconditionerD(_ICPreconditioner):
    _dtype = np.float64


class ICPreconditionerC(_ICPreconditioner):
    _dtype = np.complex128


def ICPreconditioner(mat, freedofs=None, shift: float = 1.05):
    """Incomplete Cholesky (IC) preconditioner; real/complex auto-detected."""
    A = _as_csr(mat)
    cls = ICPreconditionerC if A.is_complex else ICPreconditionerD
    return cls(A, freedofs, shift)


class _SGSPreconditioner(_DevicePreconditioner):
    """SparseSolvSGSPreconditioner (ngsolve/sparsesolv_precond.hpp:229-250):
    SGSPreconditioner (preconditioners/sgs_preconditioner.hpp:47-203), whose
    apply (D + L) D^-1 (D + U) runs as the two level-schedule sweeps of A's
    triangles in natural order."""

    def config(self) -> _lib.Config:
        return _lib.default_config(preconditioner=_lib.PRECOND_SGS)


class SGSPreconditionerD(_SGSPreconditioner):
    _dtype = np.float64


class SGSPreconditionerC(_SGSPreconditioner):
    _dtype = np.complex128


def SGSPreconditioner(mat, freedofs=None):
    """Symmetric Gauss-Seidel (SGS) preconditioner; real/complex auto-detected,
    set up at once (the factory calls Update, python_export.hpp:209-230)."""
    A = _as_csr(mat)
    cls = SGSPreconditionerC if A.is_complex else SGSPreconditionerD
    p = cls(A, freedofs)
    p.Update()
    return p


# ---------------------------------------------------------------------------
# SparseSolvSolver
# ---------------------------------------------------------------------------
_METHODS = {"ICCG": _lib.METHOD_ICCG, "iccg": _lib.METHOD_ICCG,
            "CG": _lib.METHOD_CG, "cg": _lib.METHOD_CG,
            "COCR": _lib.METHOD_COCR, "cocr": _lib.METHOD_COCR,
            "SGSMRTR": _lib.METHOD_SGSMRTR, "sgsmrtr": _lib.METHOD_SGSMRTR}


class _SparseSolvSolver:
    """SparseSolvSolver<SCAL> (ngsolve/sparsesolv_solvers.hpp:34-239)."""
    _dtype = np.float64

    def __init__(self, mat, method="ICCG", freedofs=None, tol=1e-10, maxiter=1000,
                 shift=1.05, save_best_result=True, save_residual_history=False,
                 printrates=False):
        self.mat = _as_csr(mat)
        self.freedofs = None if freedofs is None else np.array(freedofs, dtype=bool)
        self.method = method
        self.tol = float(tol)
        self.maxiter = int(maxiter)
        self.shift = float(shift)
        self.save_best_result = bool(save_best_result)
        self.save_residual_history = bool(save_residual_history)
        self.printrates = bool(printrates)
        self.auto_shift = False
        self.diagonal_scaling = False
        self.divergence_check = False
        self.divergence_threshold = 1000.0
        self.divergence_count = 100
        self.conjugate = False
        self.use_abmc = False
        self.abmc_block_size = 4
        self.abmc_num_colors = 4
        self.abmc_reorder_spmv = False
        self.abmc_use_rcm = False
        # GPU-only knobs (no reference counterpart)
        self.precision = _lib.PRECISION_FP64
        self.use_graph = True
        self._last_result = SparseSolvResult()
        self._cache_key = None
        self._handle = None

    @property
    def mixed_precision(self) -> bool:
        """IC factor and sweeps in float / complex<float> inside the FP64
        Krylov loop (GPU extension; FP64 is the reference behaviour)."""
        return self.precision == _lib.PRECISION_MIXED

    @mixed_precision.setter
    def mixed_precision(self, on: bool):
        self.precision = _lib.PRECISION_MIXED if on else _lib.PRECISION_FP64

    @property
    def last_result(self) -> SparseSolvResult:
        return self._last_result

    def _config(self) -> _lib.Config:
        return _lib.default_config(
            tolerance=self.tol, max_iterations=self.maxiter, shift_parameter=self.shift,
            save_best_result=self.save_best_result,
            save_residual_history=self.save_residual_history or self.printrates,
            auto_shift=self.auto_shift, diagonal_scaling=self.diagonal_scaling,
            divergence_check=1 if self.divergence_check else 0,
            divergence_threshold=self.divergence_threshold,
            divergence_count=self.divergence_count, conjugate=self.conjugate,
            use_abmc=self.use_abmc, abmc_block_size=self.abmc_block_size,
            abmc_num_colors=self.abmc_num_colors, abmc_reorder_spmv=self.abmc_reorder_spmv,
            abmc_use_rcm=self.abmc_use_rcm, precision=self.precision,
            use_graph=1 if self.use_graph else 0,
            # solve_sgsmrtr builds its own split factor (sparsesolv.hpp:123-137)
            preconditioner=_lib.PRECOND_SGSMRTR if _METHODS.get(self.method) == _lib.METHOD_SGSMRTR
            else _lib.PRECOND_IC)

    def _method_code(self) -> int:
        if self.method not in _METHODS:
            raise RuntimeError(f"SparseSolvSolver: Unknown method '{self.method}'. "
                               "Available: ICCG, SGSMRTR, CG, COCR")
        return _METHODS[self.method]

    def _get_handle(self, cfg: _lib.Config) -> _lib.IcHandle:
        # the reference rebuilds the preconditioner on every Solve
        # (sparsesolv.hpp:52-54); the factor depends only on the matrix and
        # these fields, so the GPU copy is reused while they are unchanged
        key = (cfg.shift_parameter, cfg.auto_shift, cfg.diagonal_scaling, cfg.use_abmc,
               cfg.abmc_block_size, cfg.abmc_num_colors, cfg.abmc_use_rcm, cfg.precision,
               cfg.preconditioner)
        if self._handle is None or key != self._cache_key:
            if self._handle is not None:
                self._handle.close()
            self._handle = _lib.IcHandle(_with_freedofs(self.mat, self.freedofs), cfg)
            self._cache_key = key
        return self._handle

    def Solve(self, rhs, sol) -> SparseSolvResult:
        """Solve A sol = rhs with sol as the initial guess (Solve, :73-160)."""
        code = self._method_code()
        cfg = self._config()
        b = np.asarray(rhs, dtype=self._dtype)
        x0 = np.asarray(sol, dtype=self._dtype)
        if self.freedofs is not None:
            b = np.where(self.freedofs, b, 0)
            x0 = np.where(self.freedofs, x0, 0)
        if code in (_lib.METHOD_CG, _lib.METHOD_COCR) and not cfg.use_abmc:
            # no preconditioner: the ordering is only the operator's layout
            cfg = _lib.default_config(**{k: getattr(cfg, k) for k, _ in _lib.Config._fields_})
            cfg.use_abmc = 1
        h = self._get_handle(cfg)
        x, res, hist = h.solve(code, b, x0, cfg)
        if self.freedofs is not None:
            sol[self.freedofs] = x[self.freedofs]
        else:
            sol[...] = x
        result = SparseSolvResult(res.converged, res.iterations, res.final_residual,
                                  hist if cfg.save_residual_history else [])
        if self.printrates and result.residual_history:
            # doubles as std::cout writes them: 6 significant digits, %g
            for i, r in enumerate(result.residual_history):
                print(f"{self.method} iteration {i}, residual = {r:g}")
            state = "converged in" if result.converged else "NOT converged after"
            print(f"{self.method} {state} {result.iterations} iterations, "
                  f"residual = {result.final_residual:g}")
        self._last_result = result
        return result

    def SolveBlock(self, rhs, sol) -> list:
        """Several right-hand sides at once: rows of rhs and sol (shape
        (k, n)), sol holding the initial guesses; row i of the result is what
        Solve(rhs[i], sol[i]) gives, bit for bit (residual histories are not
        kept for blocks).  GPU extension (the reference solves one vector per
        call): the block's SpMV and sweeps load each matrix and factor entry
        once for up to 16 rows."""
        code = self._method_code()
        if code not in (_lib.METHOD_ICCG, _lib.METHOD_ICCOCR, _lib.METHOD_CG, _lib.METHOD_COCR):
            raise RuntimeError(f"SparseSolvSolver.SolveBlock: method {self.method} runs one vector")
        cfg = self._config()
        B = np.array(rhs, dtype=self._dtype, ndmin=2)
        if not isinstance(sol, np.ndarray) or sol.ndim != 2:
            raise ValueError(f"SolveBlock: sol must be a (k, {self.mat.n}) array (it receives the solutions)")
        X0 = np.array(sol, dtype=self._dtype)
        if B.shape != X0.shape or B.shape[1] != self.mat.n:
            raise ValueError(f"SolveBlock: rhs and sol must both be (k, {self.mat.n})")
        if self.freedofs is not None:
            B = np.where(self.freedofs, B, 0)
            X0 = np.where(self.freedofs, X0, 0)
        if code in (_lib.METHOD_CG, _lib.METHOD_COCR) and not cfg.use_abmc:
            cfg = _lib.default_config(**{k: getattr(cfg, k) for k, _ in _lib.Config._fields_})
            cfg.use_abmc = 1
        h = self._get_handle(cfg)
        X, res = h.solve_multi(code, B, X0, cfg)
        if self.freedofs is not None:
            sol[:, self.freedofs] = X[:, self.freedofs]
        else:
            sol[...] = X
        results = [SparseSolvResult(r.converged, r.iterations, r.final_residual, []) for r in res]
        self._last_result = results[-1]
        return results

    def Mult(self, x, y=None):
        """y = A^{-1} x with a zero initial guess (Mult, :60-63); a (k, n)
        array is solved as a block of k right-hand sides."""
        x = np.asarray(x)
        out = np.zeros(x.shape if x.ndim == 2 else self.mat.n, dtype=self._dtype)
        if x.ndim == 2:
            self.SolveBlock(x, out)
        else:
            self.Solve(x, out)
        if y is not None:
            y[...] = out
            return y
        return out

    def MultAdd(self, s, x, y):
        tmp = self.Mult(x)
        y += s * tmp
        return y

    def __mul__(self, x):
        return self.Mult(x)


class SparseSolvSolverD(_SparseSolvSolver):
    _dtype = np.float64


class SparseSolvSolverC(_SparseSolvSolver):
    _dtype = np.complex128


def SparseSolvSolver(mat, method="ICCG", freedofs=None, tol=1e-10, maxiter=1000, shift=1.05,
                     save_best_result=True, save_residual_history=False, printrates=False,
                     conjugate=False, use_abmc=False, abmc_block_size=4, abmc_num_colors=4,
                     abmc_reorder_spmv=False, abmc_use_rcm=False):
    """Iterative solver (ICCG / SGSMRTR / CG / COCR); real/complex auto-detected."""
    A = _as_csr(mat)
    cls = SparseSolvSolverC if A.is_complex else SparseSolvSolverD
    s = cls(A, method, freedofs, tol, maxiter, shift, save_best_result,
            save_residual_history, printrates)
    s.conjugate = bool(conjugate)
    s.use_abmc = bool(use_abmc)
    s.abmc_block_size = int(abmc_block_size)
    s.abmc_num_colors = int(abmc_num_colors)
    s.abmc_reorder_spmv = bool(abmc_reorder_spmv)
    s.abmc_use_rcm = bool(abmc_use_rcm)
    return s


# ---------------------------------------------------------------------------
# COCRSolver (external preconditioner)
# ---------------------------------------------------------------------------
class _COCRSolver:
    """COCRSolverNGS<SCAL> (ngsolve/sparsesolv_solvers.hpp:256-396)."""
    _dtype = np.complex128

    def __init__(self, mat, pre, freedofs=None, maxiter=500, tol=1e-8, printrates=False):
        self.mat = _as_csr(mat)
        if not isinstance(pre, _DevicePreconditioner):
            raise TypeError("COCRSolver: on the GPU the preconditioner must be an "
                            "ICPreconditioner or SGSPreconditioner from this module")
        self.pre = pre
        self.freedofs = None if freedofs is None else np.array(freedofs, dtype=bool)
        self.maxiter = int(maxiter)
        self.tol = float(tol)
        self.printrates = bool(printrates)
        self._iterations = 0

    @property
    def iterations(self) -> int:
        return self._iterations

    def GetIterations(self) -> int:
        return self._iterations

    def _handle(self, fd) -> _lib.IcHandle:
        """pre's handle with this solve's operator (this solver's matrix with
        its freedofs) as the handle's operator: COCRSolverNGS multiplies by
        the live matrix at every Mult (ngsolve/sparsesolv_solvers.hpp:273-369)
        and applies pre's factor as set up.  The operator is uploaded
        (ssa_ic_set_operator) only when it differs from the one the handle
        holds (another matrix, or values changed in place); pre's own
        applies never use it."""
        h = self.pre.handle
        op = _with_freedofs(self.mat, fd)
        fp = _fingerprint(op)
        if fp != self.pre._operator:
            h.set_operator(op)
            self.pre._operator = fp
        return h

    def Mult(self, rhs, sol=None):
        """sol = A^{-1} rhs from a zero guess; vectors are projected with this
        solver's freedofs only, as COCRSolverNGS does (:379-384)."""
        b = np.asarray(rhs, dtype=self._dtype)
        fd = self.freedofs
        if fd is None and self.pre.freedofs is not None:
            # the reference then projects nothing in the loop while its
            # preconditioner zeroes the constrained outputs; the GPU loop
            # applies the factor without that mask
            raise ValueError("COCRSolver: the preconditioner has freedofs; pass the same "
                             "freedofs to COCRSolver")
        h = self._handle(fd)
        if fd is not None:
            b = np.where(fd, b, 0)
        cfg = _lib.default_config(tolerance=self.tol, max_iterations=self.maxiter,
                                  save_best_result=0, save_residual_history=self.printrates,
                                  use_abmc=1)
        x, res, hist = h.solve(_lib.METHOD_NGS_ICCOCR, b, None, cfg)
        self._iterations = res.iterations
        if self.printrates:
            for i, r in enumerate(hist):
                print(f"COCR iter {i}: res = {r:g}")
        if fd is not None:
            x[~fd] = 0
        if sol is not None:
            sol[...] = x
            return sol
        return x

    def __mul__(self, rhs):
        return self.Mult(rhs)


class COCRSolverD(_COCRSolver):
    _dtype = np.float64


class COCRSolverC(_COCRSolver):
    _dtype = np.complex128


def COCRSolver(mat, pre, freedofs=None, maxiter=500, tol=1e-8, printrates=False):
    """COCR for complex-symmetric A^T = A with an external IC preconditioner."""
    A = _as_csr(mat)
    cls = COCRSolverC if A.is_complex else COCRSolverD
    return cls(A, pre, freedofs, maxiter, tol, printrates)
