"""ORACLE -- test infrastructure only.

numpy restatement of the eigenpair adjoints (the formulas of
gcge_old_amd/adjoint.py), pinned in tests/test_adjoint.py by Euler
homogeneity (sum_e kappa_e dlambda/dkappa_e = lambda, K linear in kappa, M
independent of it) and by central finite differences of dense eigensolves.
"""
from __future__ import annotations

import itertools

import numpy as np
import scipy.linalg as sla

from gcge_old_amd import problems


def _corner_nodes(E: int) -> np.ndarray:
    """(E^3, 8) interior node index of each element corner, -1 on the
    boundary; corners p = 4 ax + 2 ay + az."""
    N = E - 1
    e = np.arange(E ** 3)
    ex, ey, ez = e % E, (e // E) % E, e // (E * E)
    out = np.empty((E ** 3, 8), dtype=np.int64)
    for p, (ax, ay, az) in enumerate(itertools.product((0, 1), repeat=3)):
        ix, iy, iz = ex + ax, ey + ay, ez + az
        ok = (ix >= 1) & (ix <= N) & (iy >= 1) & (iy <= N) & (iz >= 1) & (iz <= N)
        out[:, p] = np.where(ok, (ix - 1) + N * (iy - 1) + N * N * (iz - 1), -1)
    return out


def q1_element_gradient(E: int, X: np.ndarray, w, Y: np.ndarray | None = None) -> np.ndarray:
    """sum_c w_c x_c|e^T (h Khat) y_c|e for every element."""
    Y = X if Y is None else Y
    _, K36, _ = problems._element_numerators()
    nodes = _corner_nodes(E)
    Xp = np.vstack([X, np.zeros((1, X.shape[1]))])  # index -1 -> zero row
    Yp = np.vstack([Y, np.zeros((1, Y.shape[1]))])
    xe = Xp[nodes]  # (E^3, 8, k)
    ye = Yp[nodes]
    q = np.einsum("epc,pq,eqc->ec", xe, K36.astype(np.float64), ye)
    return (1.0 / E) / 36.0 * (q @ np.asarray(w, dtype=np.float64))


def pattern_gradient(A, X: np.ndarray, w, Y: np.ndarray | None = None) -> np.ndarray:
    """sum_c w_c (X[i,c] Y[j,c] + Y[i,c] X[j,c]) / 2 at the CSR entries (i, j)."""
    Y = X if Y is None else Y
    rows = np.repeat(np.arange(A.nrows), np.diff(A.rowptr))
    cols = A.colidx
    w = np.asarray(w, dtype=np.float64)
    return 0.5 * ((X[rows] * Y[cols] + Y[rows] * X[cols]) @ w)


def dense_eigs(K, M, k: int):
    """Smallest k B-normalised eigenpairs of the dense pencil."""
    lam, X = sla.eigh(K.to_scipy().toarray(), M.to_scipy().toarray(), subset_by_index=[0, k - 1])
    return lam, X


def eigvec_adjoint(K, M, x, lam, gx):
    """y with (K - lam M) y = (I - M x x^T) gx, x^T M y = 0 (dense), and
    the Q1-independent pieces: returns y."""
    Kd, Md = K.to_scipy().toarray(), M.to_scipy().toarray()
    rhs = gx - Md @ x * (x @ gx)
    y = np.linalg.lstsq(Kd - lam * Md, rhs, rcond=None)[0]
    return y - x * (x @ (Md @ y))
