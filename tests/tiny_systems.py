"""Tiny and ragged systems shared by the CPU and GPU edge-case tests: one and
two rows, and row counts on either side of the 64-row slice and the
256-row workgroup, from 5-point Laplacians on rectangular grids (SPD; a
complex-symmetric variant adds 0.3i on the diagonal)."""
import numpy as np
import scipy.sparse as sp

from sparsesolv_amd import problems as P

# name -> grid (a, b): n = a*b rows
GRIDS = {
    "n1": (1, 1), "n2": (1, 2), "n63": (7, 9), "n64": (8, 8), "n65": (5, 13),
    "n255": (15, 17), "n257": (1, 257),
}


def grid_laplacian(a, b, complex_=False):
    """-Laplace on an a x b grid, diagonal 4.2 (so 1x1 and 1x2 are SPD too)."""
    def chain(k):
        return sp.diags([-np.ones(k - 1), np.zeros(k), -np.ones(k - 1)], [-1, 0, 1]) if k > 1 \
            else sp.csr_matrix((1, 1))
    A = sp.kron(sp.identity(a), chain(b)) + sp.kron(chain(a), sp.identity(b))
    A = (A + 4.2 * sp.identity(a * b)).tocsr()
    if complex_:
        A = (A.astype(complex) + 0.3j * sp.identity(a * b)).tocsr()
    A.sort_indices()
    return P.CSR.from_scipy(A)


def rhs(n, seed, complex_=False):
    rng = np.random.default_rng(seed)
    b = rng.uniform(-1.0, 1.0, n)
    if complex_:
        b = b + 1j * rng.uniform(-1.0, 1.0, n)
    return b
