"""Synthetic test systems for the ABMC-IC(0) Krylov path.

The reference ships no matrices; its tests assemble NGSolve finite-element
systems (tests/test_sparsesolv.py).  The systems here are structured-grid
stand-ins of the same kinds, generated deterministically so that every
consumer (tests, bench, smoke) sees identical inputs:

* ``laplace7(m)`` -- 7-point Laplacian (diag 6, off-diagonals -1) on an m^3
  interior grid, Dirichlet boundary eliminated.
* ``laplace7_varcoef(m, coef)`` -- cell-centred 7-point operator with
  per-cell coefficients; face transmissibility is the harmonic mean of the
  two cells (a Dirichlet boundary face uses 2*k_cell).
* ``curlcurl(m)`` -- A = C^T diag(nu_f) C + eps*diag(nu_e) with lowest-order
  Cartesian edge unknowns on an m^3 hex grid, C the signed face-edge
  incidence, boundary edges eliminated.  nu = 1 in air and nu_iron in a
  central cube of side round(m*frac^(1/3)); face values average the 1-2
  adjacent cells, edge values the 1-4 adjacent cells.
* ``eddy(m)`` -- complex-symmetric A = K + j*M_sigma, K = curlcurl(m) and
  M_sigma the lumped edge mass weighted by a conductivity of ``sigma`` in a
  central box of volume fraction ``cond_frac``:
  M_sigma[e] = sum over the cells c adjacent to e of sigma_c / 4.

All matrices are returned as sorted, duplicate-free int32 CSR with the
diagonal stored (what NGSolve's SparseMatrix holds after CreateFromCOO).

Random numbers come from splitmix64, so the streams are reproducible in any
language: u_k = (mix(seed + (k+1)*0x9E3779B97F4A7C15) >> 11) * 2^-53.
"""
from __future__ import annotations

import dataclasses
import numpy as np
import scipy.sparse as sp

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """The splitmix64 stream: outputs offset .. offset+count-1 of ``seed``."""
    k = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, count: int, lo: float = 0.0, hi: float = 1.0,
            offset: int = 0) -> np.ndarray:
    u = (splitmix64(seed, count, offset) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return lo + (hi - lo) * u


def gaussian(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """Box-Muller on consecutive uniform pairs (u1 in (0,1], u2 in [0,1))."""
    npair = (count + 1) // 2
    u = uniform(seed, 2 * npair, offset=2 * offset)
    u1 = 1.0 - u[0::2]
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    g = np.empty(2 * npair)
    g[0::2] = r * np.cos(2.0 * np.pi * u2)
    g[1::2] = r * np.sin(2.0 * np.pi * u2)
    return g[:count]


@dataclasses.dataclass
class CSR:
    """int32 CSR matrix with sorted columns; ``values`` is float64 or complex128."""
    n: int
    row_ptr: np.ndarray
    col_idx: np.ndarray
    values: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.row_ptr[-1])

    @property
    def is_complex(self) -> bool:
        return np.iscomplexobj(self.values)

    def to_scipy(self) -> sp.csr_matrix:
        return sp.csr_matrix((self.values, self.col_idx, self.row_ptr),
                             shape=(self.n, self.n))

    @staticmethod
    def from_scipy(a) -> "CSR":
        a = sp.csr_matrix(a)
        a.sum_duplicates()
        a.sort_indices()
        return CSR(a.shape[0], a.indptr.astype(np.int32), a.indices.astype(np.int32),
                   np.ascontiguousarray(a.data))


def laplace7(m: int) -> CSR:
    """7-point Laplacian on an m^3 interior grid (n = m^3, nnz = 7n - 6m^2)."""
    n = m ** 3
    idx = np.arange(n).reshape(m, m, m)      # [k, j, i]
    rows = [idx.ravel()]
    cols = [idx.ravel()]
    vals = [np.full(n, 6.0)]
    for axis in range(3):
        a = np.take(idx, np.arange(m - 1), axis=axis).ravel()
        b = np.take(idx, np.arange(1, m), axis=axis).ravel()
        rows += [a, b]
        cols += [b, a]
        vals += [np.full(a.size, -1.0), np.full(a.size, -1.0)]
    A = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n))
    return CSR.from_scipy(A)


def laplace7_varcoef(m: int, coef: np.ndarray) -> CSR:
    """Cell-centred 7-point operator with per-cell coefficients ``coef`` (m^3)."""
    n = m ** 3
    k = np.asarray(coef, dtype=np.float64).reshape(m, m, m)
    idx = np.arange(n).reshape(m, m, m)
    diag = np.zeros((m, m, m))
    rows, cols, vals = [], [], []
    for axis in range(3):
        ka = np.take(k, np.arange(m - 1), axis=axis)
        kb = np.take(k, np.arange(1, m), axis=axis)
        t = 2.0 * ka * kb / (ka + kb)
        sl_a = [slice(None)] * 3
        sl_b = [slice(None)] * 3
        sl_a[axis] = slice(0, m - 1)
        sl_b[axis] = slice(1, m)
        diag[tuple(sl_a)] += t
        diag[tuple(sl_b)] += t
        a = idx[tuple(sl_a)].ravel()
        b = idx[tuple(sl_b)].ravel()
        rows += [a, b]
        cols += [b, a]
        vals += [-t.ravel(), -t.ravel()]
        for end in (0, m - 1):                     # Dirichlet faces
            sl = [slice(None)] * 3
            sl[axis] = slice(end, end + 1)
            diag[tuple(sl)] += 2.0 * k[tuple(sl)]
    rows.append(idx.ravel())
    cols.append(idx.ravel())
    vals.append(diag.ravel())
    A = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(n, n))
    return CSR.from_scipy(A)


# ---------------------------------------------------------------------------
# Edge-element curl-curl operator on a structured hex grid
# ---------------------------------------------------------------------------

class HexGrid:
    """Edge / face numbering of an m^3 cell grid with (m+1)^3 nodes.

    x-edge (i,j,k): i<m, j,k<=m;  y-edge: j<m;  z-edge: k<m.
    Face families: 'z' faces (normal z, spanned by x and y edges), 'x' faces
    (normal x) and 'y' faces (normal y).
    """

    def __init__(self, m: int):
        self.m = m
        m1 = m + 1
        self.nx = m * m1 * m1
        self.nedges = 3 * self.nx
        self.nfaces = 3 * m * m * m1

    # edge index helpers -------------------------------------------------
    def ex(self, i, j, k):
        m, m1 = self.m, self.m + 1
        return i + m * (j + m1 * k)

    def ey(self, i, j, k):
        m, m1 = self.m, self.m + 1
        return self.nx + i + m1 * (j + m * k)

    def ez(self, i, j, k):
        m1 = self.m + 1
        return 2 * self.nx + i + m1 * (j + m1 * k)

    def incidence(self) -> sp.csr_matrix:
        """Signed face-edge incidence C (nfaces x nedges), 4 entries per face."""
        m = self.m
        r = np.arange(m)
        r1 = np.arange(m + 1)
        rows, cols, vals = [], [], []
        off = 0
        # z faces at (i,j,k): i,j<m, k<=m : x(i,j,k) + y(i+1,j,k) - x(i,j+1,k) - y(i,j,k)
        K, J, I = np.meshgrid(r1, r, r, indexing="ij")
        f = off + np.arange(I.size); off += I.size
        for e, s in ((self.ex(I, J, K), 1.0), (self.ey(I + 1, J, K), 1.0),
                     (self.ex(I, J + 1, K), -1.0), (self.ey(I, J, K), -1.0)):
            rows.append(f); cols.append(e.ravel()); vals.append(np.full(f.size, s))
        # x faces at (i,j,k): i<=m, j,k<m : y(i,j,k) + z(i,j+1,k) - y(i,j,k+1) - z(i,j,k)
        K, J, I = np.meshgrid(r, r, r1, indexing="ij")
        f = off + np.arange(I.size); off += I.size
        for e, s in ((self.ey(I, J, K), 1.0), (self.ez(I, J + 1, K), 1.0),
                     (self.ey(I, J, K + 1), -1.0), (self.ez(I, J, K), -1.0)):
            rows.append(f); cols.append(e.ravel()); vals.append(np.full(f.size, s))
        # y faces at (i,j,k): j<=m, i,k<m : z(i,j,k) + x(i,j,k+1) - z(i+1,j,k) - x(i,j,k)
        K, J, I = np.meshgrid(r, r1, r, indexing="ij")
        f = off + np.arange(I.size); off += I.size
        for e, s in ((self.ez(I, J, K), 1.0), (self.ex(I, J, K + 1), 1.0),
                     (self.ez(I + 1, J, K), -1.0), (self.ex(I, J, K), -1.0)):
            rows.append(f); cols.append(e.ravel()); vals.append(np.full(f.size, s))
        assert off == self.nfaces
        return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                             shape=(self.nfaces, self.nedges))

    def face_average(self, cell: np.ndarray) -> np.ndarray:
        """Mean of the 1-2 cells adjacent to each face (cell indexed [k,j,i])."""
        m = self.m
        out = []
        for axis in (0, 2, 1):            # z faces, x faces, y faces (numbering order)
            pad = np.zeros([m + 1 if a == axis else m for a in range(3)])
            cnt = np.zeros_like(pad)
            sl_lo = [slice(None)] * 3
            sl_hi = [slice(None)] * 3
            sl_lo[axis] = slice(0, m)
            sl_hi[axis] = slice(1, m + 1)
            pad[tuple(sl_lo)] += cell; cnt[tuple(sl_lo)] += 1
            pad[tuple(sl_hi)] += cell; cnt[tuple(sl_hi)] += 1
            out.append((pad / cnt).ravel())
        return np.concatenate(out)

    def edge_sum(self, cell: np.ndarray, mean: bool) -> np.ndarray:
        """Sum (or mean) over the 1-4 cells adjacent to each edge."""
        m = self.m
        out = []
        for axis in (2, 1, 0):            # x edges (along i), y edges (j), z edges (k)
            shape = [m + 1] * 3
            shape[axis] = m
            acc = np.zeros(shape)
            cnt = np.zeros(shape)
            others = [a for a in range(3) if a != axis]
            for da in (0, 1):
                for db in (0, 1):
                    sl = [slice(None)] * 3
                    sl[axis] = slice(0, m)
                    sl[others[0]] = slice(da, da + m)
                    sl[others[1]] = slice(db, db + m)
                    acc[tuple(sl)] += cell
                    cnt[tuple(sl)] += 1
            out.append((acc / cnt if mean else acc).ravel())
        return np.concatenate(out)

    def interior_edges(self) -> np.ndarray:
        """Indices of the edges not on the boundary surface."""
        m = self.m
        keep = []
        for axis in (2, 1, 0):
            shape = [m + 1] * 3
            shape[axis] = m
            mask = np.ones(shape, dtype=bool)
            for a in range(3):
                if a == axis:
                    continue
                sl = [slice(None)] * 3
                sl[a] = 0
                mask[tuple(sl)] = False
                sl[a] = m
                mask[tuple(sl)] = False
            keep.append(mask.ravel())
        return np.nonzero(np.concatenate(keep))[0]


def central_box(m: int, frac: float) -> np.ndarray:
    """Boolean cell mask [k,j,i] of a centred cube with volume fraction ~frac."""
    side = int(round(m * frac ** (1.0 / 3.0)))
    lo = (m - side) // 2
    mask = np.zeros((m, m, m), dtype=bool)
    mask[lo:lo + side, lo:lo + side, lo:lo + side] = True
    return mask


@dataclasses.dataclass
class System:
    A: CSR
    b: np.ndarray
    name: str
    info: dict


def curlcurl(m: int, iron_frac: float = 0.2, nu_iron: float = 1e-3, eps: float = 1e-6,
             rhs_seed: int = 1, return_parts: bool = False):
    """Real curl-curl system (boundary edges eliminated) and RHS C^T g."""
    g = HexGrid(m)
    iron = central_box(m, iron_frac)
    nu_cell = np.where(iron, nu_iron, 1.0)
    C = g.incidence()
    nu_f = g.face_average(nu_cell)
    nu_e = g.edge_sum(nu_cell, mean=True)
    K = (C.T @ sp.diags(nu_f) @ C + sp.diags(eps * nu_e)).tocsr()
    inner = g.interior_edges()
    Kii = K[inner][:, inner]
    gf = gaussian(rhs_seed, g.nfaces)
    b = (C.T @ gf)[inner]
    A = CSR.from_scipy(Kii)
    if return_parts:
        return A, b, (g, inner, nu_cell)
    return System(A, b, f"curlcurl{m}", {"m": m, "iron_side": int(iron.sum() ** (1 / 3) + 0.5)})


def eddy(m: int, iron_frac: float = 0.2, nu_iron: float = 1e-3, eps: float = 1e-6,
         cond_frac: float = 0.15, sigma: float = 10.0, rhs_seed: int = 2) -> System:
    """Complex-symmetric A = K + j*M_sigma with a complex Gaussian RHS."""
    K, _, (g, inner, _) = curlcurl(m, iron_frac, nu_iron, eps, return_parts=True)
    cond = central_box(m, cond_frac)
    sig_cell = np.where(cond, sigma, 0.0)
    msig = (g.edge_sum(sig_cell, mean=False) / 4.0)[inner]
    A = K.to_scipy().astype(np.complex128) + sp.diags(1j * msig)
    n = A.shape[0]
    z = gaussian(rhs_seed, 2 * n)
    b = z[0::2] + 1j * z[1::2]
    return System(CSR.from_scipy(A), b, f"eddy{m}", {"m": m})


def complex_gaussian_block(n: int, count: int, seed: int = 2) -> np.ndarray:
    """`count` complex Gaussian vectors of length n, (count, n): vector k is
    outputs 2nk .. 2n(k+1)-1 of the Gaussian stream of `seed`, paired as
    (re, im); vector 0 is eddy()'s right-hand side for the same seed."""
    out = np.empty((count, n), dtype=np.complex128)
    for k in range(count):
        z = gaussian(seed, 2 * n, offset=2 * n * k)
        out[k] = z[0::2] + 1j * z[1::2]
    return out


def poisson(m: int, rhs_seed: int = 0) -> System:
    """7-point Laplacian with a uniform[-1,1] RHS."""
    A = laplace7(m)
    return System(A, uniform(rhs_seed, A.n, -1.0, 1.0), f"laplace{m}", {"m": m})


def batch_sizes(count: int = 64, lo: int = 16, hi: int = 96, seed: int = 3) -> np.ndarray:
    u = uniform(seed, count)
    return (lo + np.floor(u * (hi - lo + 1))).astype(int)


def varcoef_batch(count: int = 64, lo: int = 16, hi: int = 96, seed: int = 3):
    """Independent variable-coefficient Laplacians, sizes m^3 with m~U{lo..hi},
    cell coefficients log-uniform in [1e-2, 1e2], one uniform RHS each."""
    sizes = batch_sizes(count, lo, hi, seed)
    systems = []
    off = count
    for s, m in enumerate(sizes):
        n = int(m) ** 3
        coef = 10.0 ** uniform(seed, n, -2.0, 2.0, offset=off)
        off += n
        A = laplace7_varcoef(int(m), coef)
        b = uniform(seed, n, -1.0, 1.0, offset=off)
        off += n
        systems.append(System(A, b, f"varcoef{m}", {"m": int(m), "index": s}))
    return systems
