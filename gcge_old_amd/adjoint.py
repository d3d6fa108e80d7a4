"""Gradients of losses on computed eigenpairs (no counterpart in the reference).

For B-normalised eigenpairs ``A x_c = lambda_c B x_c`` of symmetric matrices:

* eigenvalue losses ``L(lambda)``: with ``gbar_c = dL/dlambda_c``,
  ``dL/dA = sum_c gbar_c x_c x_c^T`` and ``dL/dB = -sum_c gbar_c lambda_c
  x_c x_c^T``, formed on the sparsity pattern (``pattern_gradients``); for the
  Q1 stiffness ``K = sum_e kappa_e h Khat`` of ``problems.q1_fem_3d``,
  ``dL/dkappa_e = sum_c gbar_c x_c|e^T (h Khat) x_c|e``
  (``q1_kappa_gradient``);
* eigenvector losses ``L(x)`` of the smallest pair (simple): with ``g = dL/dx`` solve
  ``(A - lambda B) y = (I - B x x^T) g`` with ``x^T B y = 0``; then
  ``dL/dA = -(y x^T + x y^T)/2`` on the pattern and
  ``dL/dkappa_e = -y|e^T (h Khat) x|e`` (``q1_eigvec_kappa_gradient``; the
  solve is the device block CG on the shifted matrix, started from 0).

All arithmetic runs in the HIP library (``GCGE_HIP_PatternGradient``,
``GCGE_HIP_Q1ElementGradient``, ``GCGE_HIP_MultiLinearSolve``); this module
moves data and sets parameters.  ``Q1Eigenvalues`` wraps the solve and the
eigenvalue adjoint as a ``torch.autograd.Function`` (float64).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native
from . import DeviceMatrix, DeviceMultiVec, problems, solve


def _as_dev_mv(X) -> DeviceMultiVec:
    if isinstance(X, DeviceMultiVec):
        return X
    return DeviceMultiVec.from_host(np.asfortranarray(X, dtype=np.float64))


def pattern_gradients(A: DeviceMatrix, B: DeviceMatrix | None, X, lam, gbar):
    """(dL/dA, dL/dB) on the stored entries (CSR order of the matrices) for
    the eigenvalue loss with dL/dlambda = gbar; dL/dB is None when B is."""
    Xd = _as_dev_mv(X)
    k = len(gbar)
    lib = _native.hip()
    w = np.ascontiguousarray(gbar, dtype=np.float64)
    gA = np.zeros(A.nnz)
    lib.GCGE_HIP_PatternGradient(A.handle, Xd.handle, 0, None, 0, k, _native.as_dptr(w),
                                 _native.as_dptr(gA))
    gB = None
    if B is not None:
        wb = np.ascontiguousarray(-np.asarray(gbar) * np.asarray(lam), dtype=np.float64)
        gB = np.zeros(B.nnz)
        lib.GCGE_HIP_PatternGradient(B.handle, Xd.handle, 0, None, 0, k, _native.as_dptr(wb),
                                     _native.as_dptr(gB))
    return gA, gB


def q1_kappa_gradient(E: int, X, gbar, Y=None, node_row=None) -> np.ndarray:
    """sum_c gbar_c x_c|e^T (h Khat) y_c|e for every element of the E^3 mesh
    (Y = X when None): dL/dkappa of an eigenvalue loss.  node_row: for blocks
    of a symmetrically permuted pencil, the row that holds each node (the
    inverse of the permutation of problems.symmetric_permutation)."""
    Xd = _as_dev_mv(X)
    Yd = None if Y is None else _as_dev_mv(Y)
    k = len(gbar)
    w = np.ascontiguousarray(gbar, dtype=np.float64)
    g = np.zeros(E ** 3)
    lib = _native.hip()
    if node_row is None:
        lib.GCGE_HIP_Q1ElementGradient(E, Xd.handle, 0, None if Yd is None else Yd.handle, 0,
                                       k, _native.as_dptr(w), _native.as_dptr(g))
    else:
        rows = np.ascontiguousarray(node_row, dtype=np.int32)
        lib.GCGE_HIP_Q1ElementGradientRows(E, Xd.handle, 0, None if Yd is None else Yd.handle, 0,
                                           k, _native.as_dptr(w), _native.as_iptr(rows),
                                           _native.as_dptr(g))
    return g


def eigvec_adjoint_solve(A: DeviceMatrix, B: DeviceMatrix, x: np.ndarray, lam: float,
                         gx: np.ndarray, rate: float = 1e-12, max_it: int = 20000,
                         index: int = 0) -> np.ndarray:
    """y with (A - lam B) y = (I - B x x^T) gx and x^T B y = 0 (device block
    CG on the shifted matrix from y = 0, then the component along x removed).

    Only the smallest eigenpair (``index`` 0) is supported: conjugate
    gradients need A - lam B positive semidefinite, which holds for the
    smallest eigenvalue lam_1 of a pencil with B positive definite; for a
    higher eigenvalue the shifted matrix is indefinite and CG can break down,
    so any other index raises ValueError.  The linear-solver parameters in
    force before the call are restored afterwards."""
    if index != 0:
        raise ValueError("the eigenvector adjoint is solved by CG on A - lam B, which is "
                         "indefinite unless lam is the smallest eigenvalue (index 0)")
    lib = _native.hip()
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    X = DeviceMultiVec.from_host(x[:, None])
    BX = DeviceMultiVec(n, 1)
    lib.GCGE_HIP_SpMM(B.handle, X.handle, 0, BX.handle, 0, 1)
    bx = BX.get()[:, 0]
    rhs = gx - bx * (x @ gx)
    R = DeviceMultiVec.from_host(rhs[:, None])
    Y = DeviceMultiVec.from_host(np.zeros((n, 1)))
    saved = (np.zeros(1, dtype=np.int32), np.zeros(1), np.zeros(1, dtype=np.int32), np.zeros(1),
             C.c_void_p())
    lib.GCGE_HIP_GetLinearSolverParams(_native.as_iptr(saved[0]), _native.as_dptr(saved[1]),
                                       _native.as_iptr(saved[2]), _native.as_dptr(saved[3]),
                                       C.byref(saved[4]))
    lib.GCGE_HIP_SetLinearSolverParams(max_it, rate, 1, -float(lam), B.handle)
    try:
        lib.GCGE_HIP_MultiLinearSolve(A.handle, R.handle, 0, Y.handle, 0, 1)
    finally:
        lib.GCGE_HIP_SetLinearSolverParams(int(saved[0][0]), float(saved[1][0]), int(saved[2][0]),
                                           float(saved[3][0]), saved[4].value)
    y = Y.get()[:, 0]
    return y - x * (bx @ y)


def q1_eigvec_kappa_gradient(E: int, K: DeviceMatrix, M: DeviceMatrix, x: np.ndarray, lam: float,
                             gx: np.ndarray, **solve_args):
    """dL/dkappa for a loss on the B-normalised eigenvector x of the smallest
    eigenvalue lam (simple), gx = dL/dx; returns (gradient, y).  Other
    eigenpairs are refused (see eigvec_adjoint_solve)."""
    y = eigvec_adjoint_solve(K, M, x, lam, gx, **solve_args)
    g = q1_kappa_gradient(E, x[:, None], np.array([-1.0]), Y=y[:, None])
    return g, y


_q1_eigenvalues = None


def _make_q1_eigenvalues():
    import torch

    class Q1Eigenvalues(torch.autograd.Function):
        """lambda_1..nev of the Q1 pencil (K(kappa), M) on the E^3 mesh as a
        differentiable function of the per-element coefficients kappa
        (float64 tensor of E^3 entries, element order ex + E ey + E^2 ez)."""

        @staticmethod
        def forward(ctx, kappa, E, nev, flags, seed):
            K, M = problems.q1_fem_3d(E, kappa.detach().cpu().numpy())
            out = solve(K, M, nev, flags, seed=seed)
            if out["num_conv"] < nev:
                raise RuntimeError(f"eigensolver converged {out['num_conv']} of {nev} pairs")
            ctx.E = E
            ctx.evec = out["evec"]
            return torch.as_tensor(out["eval"], dtype=torch.float64)

        @staticmethod
        def backward(ctx, grad_lambda):
            g = q1_kappa_gradient(ctx.E, ctx.evec, grad_lambda.detach().cpu().numpy())
            return torch.as_tensor(g, dtype=torch.float64), None, None, None, None

    return Q1Eigenvalues


def __getattr__(name):
    # torch is imported only when the autograd wrapper is asked for
    global _q1_eigenvalues
    if name == "Q1Eigenvalues":
        if _q1_eigenvalues is None:
            _q1_eigenvalues = _make_q1_eigenvalues()
        return _q1_eigenvalues
    raise AttributeError(name)
