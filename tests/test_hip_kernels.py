"""Block kernels of the device backend against the oracle (GPU).

Tolerances: the device sums in a different order than the oracle's serial
loops, so entries are compared against a forward-error bound built from
absolute values -- |gpu - cpu| <= tol * (|A| |X|) for SpMM and
|gpu - cpu| <= tol * (|V|^T |W|) for the Gram product -- with tol far below
what a wrong index, a dropped term or a transposed operand would produce
(those give O(1) relative errors).  The random generator is integer
arithmetic and must match bit for bit.
"""
import numpy as np
import pytest

import oracle
from gcge_old_amd import DeviceMatrix, DeviceMultiVec, _native, problems

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


def _rng(seed=0):
    return np.random.default_rng(seed)


def _random_csr(n, m, density, seed, long_rows=(), empty_rows=()):
    import scipy.sparse as sp

    rng = _rng(seed)
    A = sp.random(n, m, density=density, random_state=rng, format="lil")
    for r in long_rows:
        A[r, :] = rng.standard_normal(m)
    for r in empty_rows:
        A[r, :] = 0
    return problems.CSR.from_scipy(A.tocsr())


def _banded_csr(n, offsets, keep, seed):
    """random entries on the given diagonals, each kept with probability
    `keep` (ragged rows, diagonals partly empty)"""
    import scipy.sparse as sp

    rng = _rng(seed)
    rows, cols = [], []
    for d in offsets:
        r = np.arange(max(0, -d), min(n, n - d))
        r = r[rng.uniform(size=len(r)) < keep]
        rows.append(r)
        cols.append(r + d)
    r, c = np.concatenate(rows), np.concatenate(cols)
    A = sp.csr_matrix((rng.standard_normal(len(r)), (r, c)), shape=(n, n))
    A.sort_indices()
    return problems.CSR.from_scipy(A)


MATS = {
    "fd45": lambda: problems.fd_laplacian_2d(45),
    "q1_lognormal_e8": lambda: problems.q1_fem_3d(8, problems.lognormal_kappa(8, 3))[0],
    "band130": lambda: _banded_csr(130, (-70, -69, -3, -1, 0, 1, 2, 5, 63, 64, 66), 0.8, 4),
    "band65": lambda: _banded_csr(65, (-64, -2, 0, 1, 64), 0.9, 5),
    "q1_e10": lambda: problems.q1_fem_3d(10)[0],
    "q1_mass_e9": lambda: problems.q1_fem_3d(9)[1],
    "ragged": lambda: _random_csr(1000, 1000, 0.01, 3, long_rows=(5, 700), empty_rows=(0, 64, 999)),
    "tiny": lambda: problems.CSR(3, 3, np.array([0, 2, 3, 5], np.int32),
                                 np.array([0, 2, 1, 0, 2], np.int32),
                                 np.array([2.0, -1.0, 3.0, -1.0, 4.0])),
}


@pytest.mark.parametrize("storage", ["default", "sliced", "rows"])
@pytest.mark.parametrize("name", sorted(MATS))
@pytest.mark.parametrize("k", [1, 2, 3, 5, 8, 10, 13, 17, 33, 70])
def test_spmm_matches_oracle(name, k, storage):
    """Every storage: the matrix's own choice (diagonal where it has it, row
    gathers for scattered columns), sliced ELLPACK with element gathers, and
    row gathers (from 16 columns; 70 columns take two column passes)."""
    A = MATS[name]()
    rng = _rng(k)
    X = rng.standard_normal((A.ncols, k + 3))
    dA = DeviceMatrix(A)
    if storage != "default":
        assert not dA.set_storage(storage)
    dX = DeviceMultiVec.from_host(X)
    dY = DeviceMultiVec(A.nrows, k + 2)
    _native.hip().GCGE_HIP_SpMM(dA.handle, dX.handle, 2, dY.handle, 1, k)
    got = dY.get(1, k)
    want = oracle.spmm(oracle.Matrix(A), X[:, 2:2 + k])
    bound = abs(A.to_scipy()) @ np.abs(X[:, 2:2 + k])
    assert np.all(np.abs(got - want) <= 64 * EPS * bound + 1e-300)


DIAGONAL_LAYOUT = {
    # matrix: (stored diagonals, runs of three offsets), (0, 0) = none
    "fd45": (5, 3),            # -45 | -1 0 1 | 45
    "q1_e10": (21, 9),         # kappa = 1: the +-1, +-N, +-N^2 couplings vanish
    "q1_mass_e9": (27, 9),
    "q1_lognormal_e8": (21, 9),  # those couplings are zero in every element
    "band130": (11, 6),        # -70 -69 | -3 -1 | 0 1 2 | 5 | 63 64 | 66 (runs start
                               # at the smallest offset not yet covered)
    "band65": (5, 4),          # -64 | -2 | 0 1 | 64
    "tiny": (3, 2),            # -2 0 | 2
    "ragged": (0, 0),          # random pattern: too many diagonals
}
SCATTERED = {"ragged"}  # matrices whose products gather whole rows by default


@pytest.mark.parametrize("name", sorted(DIAGONAL_LAYOUT))
def test_diagonal_storage_layout(name):
    info = DeviceMatrix(MATS[name]()).storage_info()
    want = DIAGONAL_LAYOUT[name]
    assert (info["ndiag"], info["runs"]) == want
    assert info["diagonal"] == (want[0] > 0)
    assert info["storage"] == ("diagonal" if want[0] else ("rows" if name in SCATTERED else "sliced"))


def test_no_diagonal_storage_for_permuted_or_unsorted_rows():
    K = problems.q1_fem_3d(12)[0]  # n = 1331: permuted rows span far more than 256 columns
    perm = _rng(9).permutation(K.nrows)
    Kp = problems.CSR.from_scipy(K.to_scipy()[perm][:, perm].tocsr())
    info = DeviceMatrix(Kp).storage_info()
    assert not info["diagonal"] and info["storage"] == "rows"
    # same entries, one row's columns listed in decreasing order
    rp, ci, v = K.rowptr.copy(), K.colidx.copy(), K.vals.copy()
    ci[rp[7]:rp[8]] = ci[rp[7]:rp[8]][::-1].copy()
    v[rp[7]:rp[8]] = v[rp[7]:rp[8]][::-1].copy()
    assert not DeviceMatrix(problems.CSR(K.nrows, K.ncols, rp, ci, v)).storage_info()["diagonal"]


@pytest.mark.parametrize("name", ["q1_e10", "q1_lognormal_e8", "band130"])
def test_spmm_storages_agree(name):
    """Same products in the same order: the two storages agree to a few ulps
    of the absolute-value product (different fused multiply-add grouping
    only where the compiler contracts differently)."""
    A = MATS[name]()
    k = 11
    X = _rng(7).standard_normal((A.ncols, k))
    dA = DeviceMatrix(A)
    dX = DeviceMultiVec.from_host(X)
    dY = DeviceMultiVec(A.nrows, k)
    _native.hip().GCGE_HIP_SpMM(dA.handle, dX.handle, 0, dY.handle, 0, k)
    y_dia = dY.get()
    dA.set_storage(False)
    _native.hip().GCGE_HIP_SpMM(dA.handle, dX.handle, 0, dY.handle, 0, k)
    y_sell = dY.get()
    bound = abs(A.to_scipy()) @ np.abs(X)
    assert np.all(np.abs(y_dia - y_sell) <= 4 * EPS * bound + 1e-300)


@pytest.mark.parametrize("n,m,k", [(1, 1, 1), (63, 16, 16), (64, 17, 1), (65, 1, 17),
                                   (1000, 70, 10), (4097, 130, 65), (100000, 12, 3),
                                   (100000, 1, 1), (250047, 2, 3), (5000, 8, 8), (777, 5, 5),
                                   (3, 4, 2), (1025, 1, 7), (300000, 8, 1),
                                   (250047, 80, 10), (3000, 200, 16), (777, 33, 9),
                                   (2000, 100, 100), (1500, 256, 200), (999, 96, 96),
                                   (20000, 300, 100), (5000, 64, 128), (3001, 250, 128),
                                   (17, 384, 112), (100003, 200, 100), (7001, 150, 80),
                                   (7001, 330, 90), (4099, 70, 64), (2051, 300, 65),
                                   (7001, 300, 48), (5003, 200, 8), (3001, 100, 24),
                                   (4097, 65, 1), (2049, 384, 17)])
def test_gram_matches_oracle(n, m, k):
    """Small register blocks (m, k <= 8), column blocks, matrix-core 128 x 128
    blocks and the whole-block LDS-staged kernel (64 <= m <= 384, k <= 128:
    one to three row tiles per wave, an instance for each count of 1 to 8
    column tiles, fewer rows than one stage, partial tiles)."""
    rng = _rng(n + m + k)
    V = rng.standard_normal((n, m + 1))
    W = rng.standard_normal((n, k + 2))
    dV, dW = DeviceMultiVec.from_host(V), DeviceMultiVec.from_host(W)
    G = np.zeros((m + 3, k), order="F")  # leading dimension larger than m
    _native.hip().GCGE_HIP_Gram(dV.handle, 1, m, dW.handle, 2, k, _native.as_dptr(G), m + 3)
    want = oracle.gram(V[:, 1:1 + m], W[:, 2:2 + k])
    bound = np.abs(V[:, 1:1 + m]).T @ np.abs(W[:, 2:2 + k])
    assert np.all(np.abs(G[:m] - want) <= 1e-13 * bound)
    assert np.all(G[m:] == 0.0)  # rows beyond m untouched


@pytest.mark.parametrize("m,k", [(1, 1), (5, 7), (90, 8), (90, 9), (300, 50), (70, 16), (3, 17),
                                 (20, 24), (21, 70), (6, 130), (300, 100), (41, 200), (37, 45),
                                 (64, 33), (33, 129)])
@pytest.mark.parametrize("beta", [0.0, 1.0, -0.5])
def test_linear_comb_matches_oracle(m, k, beta):
    """Vector kernel (k <= 8, 17..32), matrix-core wavefront kernel (9..16,
    and > 32 with m < 32: one to four 64-column groups, m not a multiple of
    the 4-column step) and the LDS-staged matrix-core kernel (k > 32, m >= 32:
    one or two 128-column groups, partial 16-column stages, a partial last
    block of 128 rows)."""
    n = 3001
    rng = _rng(m * k)
    V = rng.standard_normal((n, m + 2))
    C = rng.standard_normal((m + 4, k))  # leading dimension m + 4
    Y0 = rng.standard_normal((n, k + 1))
    if beta == 0.0:
        Y0[:, 1:] = np.nan  # must not be read when beta == 0
    dV, dY = DeviceMultiVec.from_host(V), DeviceMultiVec.from_host(Y0)
    Cf = np.asfortranarray(C)
    _native.hip().GCGE_HIP_LinearComb(dV.handle, 2, m, _native.as_dptr(Cf), m + 4, beta, dY.handle,
                                      1, k)
    got = dY.get(1, k)
    want = oracle.linear_comb(V[:, 2:2 + m], C[:m], beta, np.nan_to_num(Y0[:, 1:]))
    bound = np.abs(V[:, 2:2 + m]) @ np.abs(C[:m]) + abs(beta) * np.abs(np.nan_to_num(Y0[:, 1:]))
    assert np.all(np.isfinite(got))
    assert np.all(np.abs(got - want) <= 1e-13 * bound)
    np.testing.assert_array_equal(dY.get(0, 1), Y0[:, :1])  # column outside the range untouched


@pytest.mark.parametrize("m,k", [(5, 7), (90, 9), (3, 17), (21, 70), (300, 100), (64, 33)])
def test_fused_projection_subtraction_same_bits(m, k):
    """The projection pass of the svqb path folds W -= V d into one linear
    combination (coefficients -d, beta = 1, V and W columns of one block).
    Every combination kernel forms V(-d) = -(V d) exactly and adds it to W in
    one rounding, so W comes out bit for bit as from the two-step form (V d
    into a temporary, then W = -1 tmp + 1 W)."""
    n = 3001
    rng = _rng(31 * m + k)
    A = rng.standard_normal((n, m + k))
    d = np.asfortranarray(rng.standard_normal((m, k)))
    nd = np.asfortranarray(-d)
    dA = DeviceMultiVec.from_host(A)
    _native.hip().GCGE_HIP_LinearComb(dA.handle, 0, m, _native.as_dptr(nd), m, 1.0, dA.handle, m, k)
    fused = dA.get(m, k)
    dB, dT = DeviceMultiVec.from_host(A), DeviceMultiVec(n, k)
    _native.hip().GCGE_HIP_LinearComb(dB.handle, 0, m, _native.as_dptr(d), m, 0.0, dT.handle, 0, k)
    _native.hip().GCGE_HIP_Axpby(-1.0, dT.handle, 0, 1.0, dB.handle, m, k)
    np.testing.assert_array_equal(fused, dB.get(m, k))
    np.testing.assert_array_equal(dA.get(0, m), A[:, :m])


@pytest.mark.parametrize("a,b", [(0.0, 2.0), (1.0, 0.0), (-3.0, 0.0), (0.5, -2.0), (0.0, 0.0)])
def test_axpby_special_cases(a, b):
    n, k = 777, 4
    rng = _rng(7)
    X = rng.standard_normal((n, k))
    Y = rng.standard_normal((n, k))
    Xs, Ys = X.copy(), Y.copy()
    if a == 0.0:
        Xs[:] = np.nan  # x is not read when a == 0 (y = b*y, as the reference computes it)
    elif b == 0.0:
        Ys[:] = np.nan  # y is not read when b == 0 and a != 0
    dX, dY = DeviceMultiVec.from_host(Xs), DeviceMultiVec.from_host(Ys)
    _native.hip().GCGE_HIP_Axpby(a, dX.handle, 0, b, dY.handle, 0, k)
    got = dY.get()
    if a == 0.0:
        want = b * Y
    elif b == 0.0:
        want = a * X
    else:
        want = a * X + b * Y
    np.testing.assert_allclose(got, want, rtol=4 * EPS, atol=4 * EPS * np.abs(want).max())


def test_random_fill_bitwise():
    n, k = 5000, 6
    lib = _native.hip()
    lib.GCGE_HIP_SetSeed(1234)
    dX = DeviceMultiVec(n, k + 2)
    lib.GCGE_HIP_RandomFill(dX.handle, 2, k)  # draw 0, columns 2..7
    lib.GCGE_HIP_RandomFill(dX.handle, 0, 2)  # draw 1, columns 0..1
    got = dX.get()
    np.testing.assert_array_equal(got[:, 2:], oracle.random_fill(1234, 0, 2, k, n))
    np.testing.assert_array_equal(got[:, :2], oracle.random_fill(1234, 1, 0, 2, n))
    assert got.min() >= -1.0 and got.max() < 1.0
