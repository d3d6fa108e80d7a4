"""The device block CG (the operations table's MultiLinearSolver) against the
reference GCGE_BCG fixtures and the CPU restatement (GPU).

float64: solutions and residual blocks agree with the reference to rounding
(1e-10 of the largest entry), NaN columns included.  float32 directions
(GCGE_HIP_MatSetPrecision): the run of the same iteration count stays within
float32 rounding of the float64 one and reduces the true residual as much.
"""
import os

import numpy as np
import pytest

import oracle
from gcge_old_amd import DeviceMatrix, DeviceMultiVec, _native, problems

from test_oracle import BCG_FIXTURES, bcg_matrices

pytestmark = pytest.mark.gpu


def device_block_cg(A, B, rhs, x0, max_it, rate, if_shift=0, shift=0.0, fp32=False,
                    diagonal=True, storage=None):
    lib = _native.hip()
    Ad = DeviceMatrix(A)
    Bd = DeviceMatrix(B) if B is not None else None
    if storage is None and not diagonal:
        storage = "sliced"
    if storage is not None:
        for d in (Ad, Bd):
            if d is not None:
                d.set_storage(storage)
    Ad.set_precision(fp32)
    if Bd is not None:
        Bd.set_precision(fp32)
    k = rhs.shape[1]
    R = DeviceMultiVec.from_host(rhs)
    X = DeviceMultiVec.from_host(x0)
    lib.GCGE_HIP_SetLinearSolverParams(max_it, rate, if_shift, shift,
                                       None if Bd is None else Bd.handle)
    lib.GCGE_HIP_MultiLinearSolve(Ad.handle, R.handle, 0, X.handle, 0, k)
    lib.GCGE_HIP_SetLinearSolverParams(20, 1e-2, 0, 0.0, None)
    return X.get(), R.get()


def _close(a, b, rtol):
    nan = np.isnan(b)
    np.testing.assert_array_equal(np.isnan(a), nan)
    scale = np.abs(b[~nan]).max() if (~nan).any() else 1.0
    assert np.abs(a[~nan] - b[~nan]).max() <= rtol * scale


@pytest.mark.parametrize("case", BCG_FIXTURES)
def test_device_block_cg_matches_reference_fixture(case, golden_dir):
    g = np.load(os.path.join(golden_dir, "bcg_runs.npz"))
    max_it, rate, if_shift, shift = g[case + "/params"]
    A, B = bcg_matrices(str(g[case + "/spec"]))
    print(f"block CG fixture {case}", flush=True)
    x, r = device_block_cg(A, B, g[case + "/rhs"], g[case + "/x0"], int(max_it), float(rate),
                           int(if_shift), float(shift))
    _close(x, g[case + "/x"], 1e-10)
    _close(r, g[case + "/r"], 1e-9)


@pytest.mark.parametrize("bookkeeping", ["device", "host"])
@pytest.mark.parametrize("diagonal", [True, False])
@pytest.mark.parametrize("shifted", [False, True])
def test_device_block_cg_matches_restatement(shifted, diagonal, bookkeeping, monkeypatch):
    """Q1 12^3 stiffness (and mass for the shifted solve), 8 columns of mixed
    scale so columns stop at different iterations; products through the
    diagonal storage of the matrices and through sliced ELLPACK; the column
    lists and CG scalars kept on the device (unshifted solves) or on the host
    (GCGE_HIP_BCG_HOST=1, and every shifted solve)."""
    if bookkeeping == "host":
        monkeypatch.setenv("GCGE_HIP_BCG_HOST", "1")
    K, M = problems.q1_fem_3d(12)
    rng = np.random.default_rng(31)
    b = rng.standard_normal((K.nrows, 8)) * np.logspace(0, -2, 8)
    x0 = 1e-3 * rng.standard_normal((K.nrows, 8))
    args = (20, 1e-2, 1, 40.0) if shifted else (20, 1e-2, 0, 0.0)
    print(f"block CG restatement shifted={shifted} diagonal={diagonal} {bookkeeping}", flush=True)
    x, r = device_block_cg(K, M, b, x0, *args, diagonal=diagonal)
    xo, ro = oracle.block_cg(K, M, b, x0, *args)
    _close(x, xo, 1e-9)
    _close(r, ro, 1e-8)


@pytest.mark.parametrize("diagonal", [True, False])
@pytest.mark.parametrize("case", ["fd64", "q1e10_shift_B"])
def test_device_block_cg_fp32_directions(case, diagonal):
    """14 CG steps with float32 directions and products against the float64
    restatement: solution within float32 rounding, true residual as small."""
    if case == "fd64":
        A, B, args = problems.fd_laplacian_2d(64), None, (15, 1e-30, 0, 0.0)
    else:
        A, B = problems.q1_fem_3d(10)
        args = (15, 1e-30, 1, 50.0)
    rng = np.random.default_rng(32)
    b = rng.standard_normal((A.nrows, 6))
    x0 = np.zeros_like(b)
    print(f"block CG float32 {case} diagonal={diagonal}", flush=True)
    x32, _ = device_block_cg(A, B, b, x0, *args, fp32=True, diagonal=diagonal)
    x64, _ = oracle.block_cg(A, B, b, x0, *args)
    rel = np.linalg.norm(x32 - x64, axis=0) / np.linalg.norm(x64, axis=0)
    assert np.all(rel < 1e-4), rel
    Am = A.to_scipy()
    Op = (lambda X: Am @ X) if args[2] == 0 else (lambda X: Am @ X + args[3] * (B.to_scipy() @ X))
    t32 = np.linalg.norm(b - Op(x32), axis=0)
    t64 = np.linalg.norm(b - Op(x64), axis=0)
    assert np.all(t32 <= 1.5 * t64 + 1e-5 * np.linalg.norm(b, axis=0))


@pytest.mark.parametrize("fp32", [False, True])
def test_device_bookkeeping_matches_host_bookkeeping(fp32, monkeypatch):
    """Same solve twice, lists and scalars on the device and on the host:
    identical arithmetic, so identical results (NaN pattern included: the
    third column starts converged, so later columns read rho1 entries that
    this call never wrote -- zeros here -- as the reference does)."""
    K, _ = problems.q1_fem_3d(9)
    rng = np.random.default_rng(34)
    b = rng.standard_normal((K.nrows, 7)) * np.logspace(0, -6, 7)
    b[:, 2] = 0.0
    x0 = 1e-2 * rng.standard_normal((K.nrows, 7))
    x0[:, 2] = 0.0
    out = {}
    for mode in ("device", "host"):
        if mode == "host":
            monkeypatch.setenv("GCGE_HIP_BCG_HOST", "1")
        print(f"block CG bookkeeping {mode} fp32={fp32}", flush=True)
        out[mode] = device_block_cg(K, None, b, x0, 20, 1e-3, fp32=fp32)
    for a, h in zip(out["device"], out["host"]):
        np.testing.assert_array_equal(a, h)


@pytest.mark.parametrize("cols", [20, 150])
@pytest.mark.parametrize("fp32", [False, True])
def test_device_block_cg_row_gathers_on_permuted_matrix(fp32, cols):
    """A randomly permuted Q1 pencil (row gathers with the p.w epilogue: 20
    columns, one pass per workgroup; 150 columns, three passes in one
    workgroup) against the restatement on the same matrices: float64 to
    rounding, float32 within float32 rounding."""
    K, M = problems.q1_fem_3d(9, problems.lognormal_kappa(9, 5))
    perm = np.random.default_rng(36).permutation(K.nrows)
    Kp = problems.symmetric_permutation(K, perm)
    rng = np.random.default_rng(37)
    b = rng.standard_normal((K.nrows, cols)) * np.logspace(0, -3, cols)
    x0 = 1e-3 * rng.standard_normal((K.nrows, cols))
    print(f"block CG row gathers fp32={fp32} cols={cols}", flush=True)
    x, _ = device_block_cg(Kp, None, b, x0, 12, 1e-3, fp32=fp32)
    xo, _ = oracle.block_cg(Kp, None, b, x0, 12, 1e-3)
    if fp32:
        rel = np.linalg.norm(x - xo, axis=0) / np.linalg.norm(xo, axis=0)
        assert np.all(rel < 1e-4), rel
    else:
        _close(x, xo, 1e-9)


def test_device_block_cg_fp32_large_beta():
    """Columns of very different scale (1 .. 1e-14): once the first columns
    stop, the reference bookkeeping hands later columns another column's
    scalars and beta grows by orders of magnitude per step.  float64 absorbs
    it; the scaled float32 directions must follow the float64 iteration (an
    unscaled float32 store overflows to inf/NaN on this case)."""
    K, _ = problems.q1_fem_3d(10)
    rng = np.random.default_rng(33)
    b = rng.standard_normal((K.nrows, 8)) * np.logspace(0, -14, 8)
    x0 = np.zeros_like(b)
    print("block CG float32 large beta", flush=True)
    x32, _ = device_block_cg(K, None, b, x0, 20, 1e-2, fp32=True)
    x64, _ = oracle.block_cg(K, None, b, x0, 20, 1e-2)
    assert np.isfinite(x32).all()
    rel = np.linalg.norm(x32 - x64, axis=0) / np.linalg.norm(x64, axis=0)
    assert np.all(rel < 1e-4), rel
