"""Q1 matrices assembled on the device against the host assembly (GPU).

`problems.q1_fem_3d` is the host assembly (itself pinned by the closed-form
spectrum of the kappa = 1 pencil and the element-matrix identities in
tests/test_oracle.py); `problems.symmetric_permutation` renumbers it (pinned
against scipy in tests/test_adjoint.py).  The device matrix must equal it
bit for bit -- row pointers, columns, values -- and make the same storage
decisions as GCGE_HIP_MatCreateCSR of the host arrays, so the products of the
two matrices agree bit for bit as well.
"""
import numpy as np
import pytest

from gcge_old_amd import DeviceMatrix, DeviceMultiVec, _native, problems

pytestmark = pytest.mark.gpu


def _host(E, kappa, which, node_row):
    K, M = problems.q1_fem_3d(E, kappa)
    A = K if which == "stiffness" else M
    if node_row is not None:
        perm = np.empty_like(node_row)
        perm[node_row] = np.arange(len(node_row))  # row i holds node perm[i]
        A = problems.symmetric_permutation(A, perm)
    return A


def _node_row(kind, n, seed=0):
    if kind == "natural":
        return None
    if kind == "identity":
        return np.arange(n, dtype=np.int32)
    if kind == "reversed":
        return np.arange(n - 1, -1, -1, dtype=np.int32)
    return np.random.default_rng(seed).permutation(n).astype(np.int32)


def _kappa(kind, E, seed=1):
    if kind == "one":
        return None
    k = problems.lognormal_kappa(E, seed)
    if kind == "zeros":  # some elements without stiffness: more exact zeros
        k[::3] = 0.0
    return k


def _same_bits(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


def _spmm(dA, X):
    n, k = X.shape
    dX = DeviceMultiVec.from_host(X)
    dY = DeviceMultiVec(n, k)
    _native.hip().GCGE_HIP_SpMM(dA.handle, dX.handle, 0, dY.handle, 0, k)
    return dY.get()


CASES = [
    (2, "one", "natural"), (3, "one", "natural"), (3, "lognormal", "random"),
    (4, "lognormal", "natural"), (7, "one", "random"), (7, "zeros", "natural"),
    (9, "lognormal", "random"), (9, "lognormal", "reversed"), (9, "one", "identity"),
    (12, "lognormal", "natural"), (12, "zeros", "random"),
]


@pytest.mark.parametrize("which", ["stiffness", "mass"])
@pytest.mark.parametrize("E,kappa,order", CASES)
def test_device_assembly_equals_host(E, kappa, order, which):
    n = (E - 1) ** 3
    kap = _kappa(kappa, E)
    nr = _node_row(order, n)
    A = _host(E, kap, which, nr)
    dA = DeviceMatrix.q1(E, kap, which, nr)
    got = dA.to_csr()
    assert got.nnz == A.nnz
    assert np.array_equal(got.rowptr, A.rowptr)
    assert np.array_equal(got.colidx, A.colidx)
    assert _same_bits(got.vals, A.vals)
    hA = DeviceMatrix(A)
    assert dA.storage_info() == hA.storage_info()
    for k in (1, 10, 24):
        X = np.random.default_rng(k).standard_normal((n, k))
        assert _same_bits(_spmm(dA, X), _spmm(hA, X))


def test_device_assembly_storage_decisions():
    """natural order: diagonal storage (21 diagonals of K, 27 of M, 9 runs);
    reversed order: diagonal too (offsets negated); a random order of a
    larger mesh: scattered columns, row gathers"""
    E = 12
    info = DeviceMatrix.q1(E, None, "stiffness").storage_info()
    assert (info["storage"], info["ndiag"], info["runs"]) == ("diagonal", 21, 9)
    info = DeviceMatrix.q1(E, None, "mass").storage_info()
    assert (info["storage"], info["ndiag"], info["runs"]) == ("diagonal", 27, 9)
    n = (E - 1) ** 3
    info = DeviceMatrix.q1(E, None, "mass", _node_row("reversed", n)).storage_info()
    assert (info["storage"], info["ndiag"]) == ("diagonal", 27)
    info = DeviceMatrix.q1(E, None, "mass", _node_row("random", n)).storage_info()
    assert info["storage"] == "rows"


def test_get_csr_returns_the_given_arrays():
    """MatGetCSR of a host-created matrix gives back its CSR arrays, also
    with rows whose columns are not sorted"""
    K, _ = problems.q1_fem_3d(6, problems.lognormal_kappa(6, 2))
    perm = np.random.default_rng(3).permutation(K.nrows)
    Ku = problems.symmetric_permutation(K, perm, sort_rows=False)
    for A in (K, Ku):
        got = DeviceMatrix(A).to_csr()
        assert np.array_equal(got.rowptr, A.rowptr)
        assert np.array_equal(got.colidx, A.colidx)
        assert _same_bits(got.vals, A.vals)


def test_device_assembly_solves_the_pencil():
    """kappa = 1 pencil from device-assembled K and M: the smallest
    eigenvalues match the closed form, and the solve equals the one on the
    host-assembled matrices bit for bit"""
    import gcge_old_amd as g

    E, nev = 10, 10
    flags = ["-gcge_conv_type", "RB", "-gcge_ev_tol", "1e-10", "-gcge_ev_max_it", "100"]
    K = DeviceMatrix.q1(E, None, "stiffness")
    M = DeviceMatrix.q1(E, None, "mass")
    out = g.solve(K, M, nev, flags, seed=5)
    assert out["num_conv"] == nev
    np.testing.assert_allclose(out["eval"], problems.q1_eigs(E, nev), rtol=1e-10)
    Kh, Mh = problems.q1_fem_3d(E)
    ref = g.solve(Kh, Mh, nev, flags, seed=5)
    assert out["num_iter"] == ref["num_iter"]
    assert _same_bits(out["eval"], ref["eval"])
    assert _same_bits(out["evec"], ref["evec"])


def test_device_assembly_rejects_bad_arguments():
    with pytest.raises(ValueError):
        DeviceMatrix.q1(4, None, "stiffness", np.zeros(27, dtype=np.int32))
    with pytest.raises(ValueError):
        DeviceMatrix.q1(4, np.ones(5), "stiffness")
    with pytest.raises(ValueError):
        DeviceMatrix.q1(4, None, "damping")


# ---- diagonal products in plane order ------------------------------------------------------


@pytest.mark.parametrize("E,which", [(93, "stiffness"), (10, "stiffness"), (12, "mass")])
def test_plane_order_products_equal_consecutive_order(E, which):
    """The two row-block orders of the diagonal products compute every row
    with the same arithmetic: equal bit for bit, for every column count
    (E = 93: plane stride 92^2 = 8464 rows, the order chosen by default)"""
    n = (E - 1) ** 3
    kap = problems.lognormal_kappa(E, 2) if which == "stiffness" else None
    dA = DeviceMatrix.q1(E, kap, which)
    S = (E - 1) ** 2
    if S >= 8192:
        assert dA.storage_info()["plane_rows"] == S
    for k in (1, 3, 10, 37):
        X = np.random.default_rng(k).standard_normal((n, k))
        dA.set_storage("diagonal-planes")
        assert dA.storage_info()["plane_rows"] == S
        y_planes = _spmm(dA, X)
        dA.set_storage("diagonal-consecutive")
        assert dA.storage_info()["plane_rows"] == 0
        y_cons = _spmm(dA, X)
        assert _same_bits(y_planes, y_cons)
    dA.set_storage("sliced")
    assert dA.storage_info()["plane_rows"] == 0


@pytest.mark.parametrize("E,which", [(93, "stiffness"), (10, "stiffness"), (12, "mass"),
                                     (130, "mass")])
def test_plane_tiles_equal_consecutive_order(E, which):
    """The LDS plane tiles add each row's products in the order of the
    row-block kernel: equal bit for bit, column counts that are not a
    multiple of the staged chunk and a partial last plane group included
    (E = 130: line stride 129, three staged positions per thread)"""
    n = (E - 1) ** 3
    kap = problems.lognormal_kappa(E, 5) if which == "stiffness" else None
    dA = DeviceMatrix.q1(E, kap, which)
    ks = (1, 3, 10, 37) if E < 100 else (2, 9)
    for k in ks:
        X = np.random.default_rng(10 + k).standard_normal((n, k))
        dA.set_storage("diagonal-tiles")
        y_tiles = _spmm(dA, X)
        dA.set_storage("diagonal-consecutive")
        y_cons = _spmm(dA, X)
        assert _same_bits(y_tiles, y_cons), (E, which, k)
    dA.set_storage("diagonal")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_plane_order_solve(precision):
    """solves through the block CG (products with the p.w epilogue, float32
    in the mixed form) agree in both orders"""
    import gcge_old_amd as g

    E, nev = 12, 10
    flags = ["-gcge_conv_type", "RB", "-gcge_ev_tol", "1e-10", "-gcge_ev_max_it", "100"]
    out = {}
    for order in ("diagonal-planes", "diagonal-consecutive", "diagonal-tiles"):
        K = DeviceMatrix.q1(E, problems.lognormal_kappa(E, 4), "stiffness")
        M = DeviceMatrix.q1(E, None, "mass")
        K.set_storage(order)
        M.set_storage(order)
        out[order] = g.solve(K, M, nev, flags, seed=3, precision=precision)
        assert out[order]["num_conv"] == nev
    a, b = out["diagonal-planes"]["eval"], out["diagonal-consecutive"]["eval"]
    np.testing.assert_allclose(a, b, rtol=1e-10)
    np.testing.assert_allclose(out["diagonal-tiles"]["eval"], b, rtol=1e-10)
