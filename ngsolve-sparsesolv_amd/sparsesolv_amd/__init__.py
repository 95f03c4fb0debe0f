"""MI355X-native ABMC-IC(0) Krylov path (ICCG, auto-shift ICCG, COCR) with the
SGS preconditioner and SGS-MRTR on the same sweeps.

Drop-in for the ICCG / SGSMRTR / COCR part of ksugahar/ngsolve-sparsesolv's
Python module: ``SparseSolvSolver``, ``ICPreconditioner``,
``SGSPreconditioner``, ``COCRSolver`` and ``SparseSolvResult`` with the
reference's names, arguments and defaults, running on the GPU through
libsparsesolv_amd.so.
"""
from . import problems
from ._lib import load as load_library
from .solvers import (COCRSolver, COCRSolverC, COCRSolverD, ICPreconditioner,
                      ICPreconditionerC, ICPreconditionerD, SGSPreconditioner,
                      SGSPreconditionerC, SGSPreconditionerD, SparseSolvResult,
                      SparseSolvSolver, SparseSolvSolverC, SparseSolvSolverD)

__version__ = "0.1.0"

__all__ = ["COCRSolver", "COCRSolverC", "COCRSolverD", "ICPreconditioner",
           "ICPreconditionerC", "ICPreconditionerD", "SGSPreconditioner",
           "SGSPreconditionerC", "SGSPreconditionerD", "SparseSolvResult",
           "SparseSolvSolver", "SparseSolvSolverC", "SparseSolvSolverD", "problems",
           "load_library"]
