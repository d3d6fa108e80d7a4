"""Model problems for the GCG eigensolver, in CSR form.

These generate the matrices of the benchmark configurations:

* ``fd_laplacian_2d(N)`` -- the 5-point finite-difference Laplacian on an
  N x N interior grid of the unit square, Dirichlet boundary, h = 1/(N+1),
  ``A = (1/h^2) (T (x) I + I (x) T)`` with ``T = tridiag(-1, 2, -1)``,
  lexicographic ordering.  Exact eigenvalues
  ``(4/h^2) (sin^2(i pi h / 2) + sin^2(j pi h / 2))``, the formula the
  reference's own HYPRE test prints (test/hypre/laplace.c:31-43).
* ``q1_fem_3d(E, kappa)`` -- trilinear (Q1) hexahedral finite elements on the
  uniform E^3 mesh of the unit cube, Dirichlet nodes eliminated, so
  n = (E-1)^3 interior nodes in lexicographic order.  Stiffness
  ``K = sum_e kappa_e h Khat`` and consistent mass ``M = sum_e h^3 Mhat``.
  The element matrices are exact rationals: ``36 Khat`` and ``216 Mhat`` are
  integers, so the assembly adds integer numerators times kappa_e and scales
  once.  Face-neighbour entries of K are exactly zero for every kappa
  (each element contributes 0 to them) and are dropped.  With kappa = 1 the
  spectrum is the Kronecker sum ``mu_i + mu_j + mu_k``,
  ``mu_j = (6/h^2) (1 - cos t_j) / (2 + cos t_j)``, ``t_j = j pi / E``.

The assembly is vectorised per stencil offset (27 offsets x 8 surrounding
elements), so a 128^3 mesh assembles in seconds without forming element
lists.
"""
from __future__ import annotations

import dataclasses
import itertools

import numpy as np


@dataclasses.dataclass
class CSR:
    """Compressed sparse rows with int32 indices and float64 values."""

    nrows: int
    ncols: int
    rowptr: np.ndarray
    colidx: np.ndarray
    vals: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.colidx.shape[0])

    def to_scipy(self):
        import scipy.sparse as sp

        return sp.csr_matrix((self.vals, self.colidx, self.rowptr), shape=(self.nrows, self.ncols))

    @staticmethod
    def from_scipy(m) -> "CSR":
        m = m.tocsr()
        m.sort_indices()
        return CSR(m.shape[0], m.shape[1], m.indptr.astype(np.int32), m.indices.astype(np.int32),
                   m.data.astype(np.float64))

    def write_bin(self, path: str) -> None:
        """Raw layout read by the reference driver: int32 nrows, ncols, nnz,
        int32 rowptr, int32 colidx, float64 vals."""
        with open(path, "wb") as f:
            np.array([self.nrows, self.ncols, self.nnz], dtype=np.int32).tofile(f)
            self.rowptr.astype(np.int32).tofile(f)
            self.colidx.astype(np.int32).tofile(f)
            self.vals.astype(np.float64).tofile(f)

    @staticmethod
    def read_text(path: str) -> "CSR":
        """The reference CSR text format (app/csr/gcge_app_csr.c:185-236):
        nrow, ncol, nnz, then ia[nrow+1], ja[nnz], aa[nnz], 0-based."""
        data = np.loadtxt(path, dtype=np.float64)
        nrow, ncol, nnz = int(data[0]), int(data[1]), int(data[2])
        ia = data[3:3 + nrow + 1].astype(np.int32)
        ja = data[3 + nrow + 1:3 + nrow + 1 + nnz].astype(np.int32)
        aa = data[3 + nrow + 1 + nnz:3 + nrow + 1 + 2 * nnz].copy()
        return CSR(nrow, ncol, ia, ja, aa)

    PETSC_MAT_CLASSID = 1211216

    @staticmethod
    def read_petsc(path: str) -> "CSR":
        """A PETSc binary matrix file, the input of the reference's SLEPc
        driver (app/slepc/gcge_app_slepc.c:25-36, MatLoad): big-endian
        int32 class id 1211216, rows, columns, nonzeros, the row lengths,
        the column indices, then float64 values."""
        with open(path, "rb") as f:
            data = f.read()
        head = np.frombuffer(data, dtype=">i4", count=4)
        if head[0] != CSR.PETSC_MAT_CLASSID:
            raise ValueError(f"{path}: not a PETSc binary matrix (class id {head[0]})")
        m, n, nz = (int(v) for v in head[1:])
        if len(data) != 16 + 4 * m + 12 * nz:
            raise ValueError(f"{path}: size {len(data)} does not match {m} rows and {nz} nonzeros")
        lens = np.frombuffer(data, dtype=">i4", count=m, offset=16).astype(np.int64)
        cols = np.frombuffer(data, dtype=">i4", count=nz, offset=16 + 4 * m).astype(np.int32)
        vals = np.frombuffer(data, dtype=">f8", count=nz, offset=16 + 4 * m + 4 * nz).astype(np.float64)
        rowptr = np.zeros(m + 1, dtype=np.int64)
        np.cumsum(lens, out=rowptr[1:])
        if rowptr[-1] != nz:
            raise ValueError(f"{path}: row lengths sum to {rowptr[-1]}, not {nz}")
        return CSR(m, n, rowptr.astype(np.int32), cols, vals)

    def write_petsc(self, path: str) -> None:
        """The same format (readable by PETSc's MatLoad)."""
        with open(path, "wb") as f:
            f.write(np.array([CSR.PETSC_MAT_CLASSID, self.nrows, self.ncols, self.nnz], ">i4").tobytes())
            f.write(np.diff(self.rowptr).astype(">i4").tobytes())
            f.write(np.asarray(self.colidx).astype(">i4").tobytes())
            f.write(np.asarray(self.vals).astype(">f8").tobytes())

    def write_text(self, path: str) -> None:
        with open(path, "w") as f:
            f.write(f"{self.nrows}\n{self.ncols}\n{self.nnz}\n")
            f.write("\n".join(str(int(v)) for v in self.rowptr) + "\n")
            f.write("\n".join(str(int(v)) for v in self.colidx) + "\n")
            f.write("\n".join(repr(float(v)) for v in self.vals) + "\n")


# ---------------------------------------------------------------------------
# 2D finite differences

def fd_laplacian_2d(N: int) -> CSR:
    h = 1.0 / (N + 1)
    n = N * N
    idx = np.arange(n, dtype=np.int64)
    ix, iy = idx % N, idx // N
    rows, cols, vals = [idx], [idx], [np.full(n, 4.0)]
    for dx, dy in ((-1, 0), (1, 0), (0, -1), (0, 1)):
        ok = (ix + dx >= 0) & (ix + dx < N) & (iy + dy >= 0) & (iy + dy < N)
        rows.append(idx[ok])
        cols.append(idx[ok] + dx + N * dy)
        vals.append(np.full(int(ok.sum()), -1.0))
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    v = np.concatenate(vals) / (h * h)
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(rowptr, r + 1, 1)
    rowptr = np.cumsum(rowptr)
    return CSR(n, n, rowptr.astype(np.int32), c.astype(np.int32), v)


def fd_laplacian_2d_eigs(N: int, k: int) -> np.ndarray:
    h = 1.0 / (N + 1)
    j = np.arange(1, N + 1)
    s = 4.0 / h**2 * np.sin(j * np.pi * h / 2) ** 2
    return np.sort((s[:, None] + s[None, :]).ravel())[:k]


# ---------------------------------------------------------------------------
# 3D Q1 finite elements

def _element_numerators():
    """36*Khat and 216*Mhat for the unit cube, local node a = (ax, ay, az)."""
    corners = list(itertools.product((0, 1), repeat=3))  # (ax, ay, az)
    K36 = np.zeros((8, 8), dtype=np.int64)
    M216 = np.zeros((8, 8), dtype=np.int64)
    for p, a in enumerate(corners):
        for q, b in enumerate(corners):
            same = [a[d] == b[d] for d in range(3)]
            k1 = [1 if s else -1 for s in same]
            m6 = [2 if s else 1 for s in same]
            K36[p, q] = (k1[0] * m6[1] * m6[2] + m6[0] * k1[1] * m6[2] + m6[0] * m6[1] * k1[2])
            M216[p, q] = m6[0] * m6[1] * m6[2]
    return corners, K36, M216


def q1_fem_3d(E: int, kappa: np.ndarray | None = None, drop_zeros: bool = True):
    """Return (K, M) for the E^3 mesh; kappa has E^3 entries in element
    order ex + E*ey + E^2*ez (None: kappa = 1)."""
    N = E - 1
    h = 1.0 / E
    n = N ** 3
    if kappa is None:
        kappa = np.ones(E ** 3)
    kappa = np.asarray(kappa, dtype=np.float64).reshape(E, E, E)  # [ez, ey, ex]
    corners, K36, M216 = _element_numerators()
    corner_index = {c: i for i, c in enumerate(corners)}
    g = np.arange(1, E)  # interior grid coordinates
    # node grid coordinates, node index x-fastest
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    offsets = list(itertools.product((-1, 0, 1), repeat=3))  # (dz, dy, dx)
    kv = np.zeros((27, n))
    mv = np.zeros((27, n))
    valid = np.zeros((27, n), dtype=bool)
    for o, (dz, dy, dx) in enumerate(offsets):
        xj, yj, zj = x + dx, y + dy, z + dz
        ok = (xj >= 1) & (xj <= N) & (yj >= 1) & (yj <= N) & (zj >= 1) & (zj <= N)
        ksum = np.zeros_like(x, dtype=np.float64)
        msum = 0
        for sx, sy, sz in itertools.product((0, 1), repeat=3):
            # element (x-1+sx, ...) in which node i has local coordinate (1-sx, ...)
            a = (1 - sx, 1 - sy, 1 - sz)
            b = (a[0] + dx, a[1] + dy, a[2] + dz)
            if not all(v in (0, 1) for v in b):
                continue
            p, q = corner_index[a], corner_index[b]
            ex, ey, ez = x - 1 + sx, y - 1 + sy, z - 1 + sz
            ksum = ksum + K36[p, q] * kappa[ez, ey, ex]
            msum = msum + M216[p, q]
        kv[o] = (ksum * (h / 36.0)).ravel()
        mv[o] = np.full(n, msum * (h ** 3 / 216.0))
        valid[o] = ok.ravel()
    node = np.arange(n, dtype=np.int64)
    cols = np.stack([node + dx + N * dy + N * N * dz for (dz, dy, dx) in offsets])  # sorted per row
    K = _stencil_to_csr(n, cols, kv, valid & ((kv != 0.0) if drop_zeros else True))
    M = _stencil_to_csr(n, cols, mv, valid)
    return K, M


def _stencil_to_csr(n, cols, vals, keep):
    counts = keep.sum(axis=0)
    rowptr = np.zeros(n + 1, dtype=np.int64)
    rowptr[1:] = np.cumsum(counts)
    colidx = cols.T[keep.T].astype(np.int32)
    v = vals.T[keep.T].astype(np.float64)
    return CSR(n, n, rowptr.astype(np.int32), colidx, v)


def q1_eigs(E: int, k: int) -> np.ndarray:
    """Exact smallest k eigenvalues of the kappa = 1 Q1 pencil (K, M)."""
    h = 1.0 / E
    t = np.arange(1, E) * np.pi / E
    mu = 6.0 / h**2 * (1 - np.cos(t)) / (2 + np.cos(t))
    mu = np.sort(mu)
    mu = mu[:min(len(mu), k)]  # the k smallest sums use at most k modes per axis
    lam = (mu[:, None, None] + mu[None, :, None] + mu[None, None, :]).ravel()
    return np.sort(lam)[:k]


def lognormal_kappa(E: int, seed: int) -> np.ndarray:
    """kappa_e = exp(z_e), z_e ~ N(0, 1) i.i.d., element order ex + E ey + E^2 ez."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.exp(rng.standard_normal(E ** 3))


def symmetric_permutation(A: CSR, perm: np.ndarray, sort_rows: bool = True) -> CSR:
    """P A P^T with row i of the result = row perm[i] of A (and column j =
    column perm[j]); each row's entries are gathered in one pass and the
    columns renumbered, then sorted within the row unless sort_rows is
    False (CSR does not require sorted rows; the sort is the costly step at
    10^8 entries)."""
    n = A.nrows
    perm = np.asarray(perm, dtype=np.int64)
    inv = np.empty(n, dtype=np.int64)
    inv[perm] = np.arange(n)
    rp = A.rowptr.astype(np.int64)
    counts = np.diff(rp)[perm]
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=rowptr[1:])
    src = np.repeat(rp[perm] - rowptr[:-1], counts) + np.arange(rowptr[-1])
    cols = inv[A.colidx[src]]
    vals = A.vals[src]
    del src
    if sort_rows:
        rows = np.repeat(np.arange(n, dtype=np.int64), counts)
        order = np.lexsort((cols, rows))
        cols, vals = cols[order], vals[order]
    return CSR(n, A.ncols, rowptr.astype(np.int32), cols.astype(np.int32), vals)
