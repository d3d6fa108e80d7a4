"""The benchmark configurations at full size on the device (GPU).

C1 (2D FD 128^2, B = I) is test_hip_solver.py::test_fd2d_128_closed_form_rb_criterion.
C2  Q1 64^3, kappa = 1, nev 50: the closed-form spectrum (Kronecker sum),
    ||Kx - lambda Mx|| / (|lambda| ||Mx||) <= 1e-8, M-orthonormal vectors, in
    the benchmark's solver mode and in the reference's (soft locking,
    "multi" orthonormalisation, its own residual R), whose first three
    iterations follow the oracle run from the same block.
C3  Q1 128^3, lognormal kappa, nev 100, block 200 requested: the mixed and
    the float64 device solves agree eigenvalue by eigenvalue (1e-9), their
    converged subspaces by the largest principal angle (< 1e-6), and both
    meet the residual test.
C4  Q1 64^3, lognormal kappa, nev 20: dL/dkappa of L = sum w_k lambda_k on
    all 262144 elements -- Euler homogeneity sum_e kappa_e g_e = L (K is
    linear in kappa, M independent of it), fourth-order central differences
    on sampled elements -- and the eigenvector adjoint's shifted system.
C5  Q1 192^3, lognormal kappa, rows and columns randomly permuted, nev 200,
    block 400 requested: the products against the oracle on a row sample,
    a bounded mixed-precision solve (M-orthonormal block), and the
    eigenvalue adjoint on all 7077888 elements through the node -> row map
    by the identity sum_e kappa_e g_e = sum_c w_c x_c^T K x_c.

Matrices are assembled on the device (GCGE_HIP_MatCreateQ1, equal bit for bit
to the host assembly: test_hip_assembly.py).  Products and Gram blocks of the
checks run on the device; norms and small dense algebra on the host.
"""
import numpy as np
import pytest

import gcge_old_amd as gcge
import oracle
from gcge_old_amd import DeviceMatrix, DeviceMultiVec, _native, adjoint, problems

pytestmark = pytest.mark.gpu

BENCH_FLAGS = ["-gcge_conv_type", "RB", "-gcge_ev_tol", "1e-8", "-gcge_full_rr", "1",
               "-gcge_w_orth_type", "svqb", "-gcge_max_reorth_time", "2"]


def _products(dA, X):
    """A X on the device for a host block X."""
    n, k = X.shape
    dX = DeviceMultiVec.from_host(X)
    dY = DeviceMultiVec(n, k)
    _native.hip().GCGE_HIP_SpMM(dA.handle, dX.handle, 0, dY.handle, 0, k)
    return dY.get()


def _rb_residuals(dK, dM, lam, X):
    """||K x - lambda M x|| / (|lambda| ||M x||) per column."""
    KX, MX = _products(dK, X), _products(dM, X)
    R = KX - MX * lam
    return np.linalg.norm(R, axis=0) / (np.abs(lam) * np.linalg.norm(MX, axis=0))


def _gram(X, Y):
    """X^T Y on the device."""
    k1, k2 = X.shape[1], Y.shape[1]
    dX, dY = DeviceMultiVec.from_host(X), DeviceMultiVec.from_host(Y)
    G = np.zeros((k1, k2), order="F")
    _native.hip().GCGE_HIP_Gram(dX.handle, 0, k1, dY.handle, 0, k2, _native.as_dptr(G), k1)
    return G


def _largest_angle_m(dM, X, Y):
    """Largest principal angle between span(X) and span(Y), both
    M-orthonormal: arccos of the smallest singular value of X^T M Y."""
    s = np.linalg.svd(_gram(X, _products(dM, Y)), compute_uv=False)
    return float(np.arccos(np.clip(s.min(), -1.0, 1.0)))


# ---- C2 --------------------------------------------------------------------------


def test_c2_q1_64_benchmark_mode_closed_form():
    E, nev = 64, 50
    dK, dM = DeviceMatrix.q1(E, None, "stiffness"), DeviceMatrix.q1(E, None, "mass")
    print(f"C2 bench mode E={E} nev={nev}", flush=True)
    out = gcge.solve(dK, dM, nev, BENCH_FLAGS + ["-gcge_ev_max_it", "300"], seed=1)
    assert out["num_conv"] == nev
    np.testing.assert_allclose(np.sort(out["eval"]), problems.q1_eigs(E, nev), rtol=1e-10)
    X = out["evec"]
    assert np.all(_rb_residuals(dK, dM, out["eval"], X) <= 1.05e-8)
    G = _gram(X, _products(dM, X))
    assert np.abs(G - np.eye(nev)).max() < 1e-10


def test_c2_q1_64_reference_mode():
    """The reference's own solver settings (soft locking, "multi" W
    orthonormalisation, block size 10, residual ||r|| / (||x|| max(1, |lambda|))
    <= 1e-8 as src/gcge_eigsol.c:704-721 defines it): converged to the
    closed form to the accuracy that residual allows, and the first three
    iterations equal the oracle run from the same block (1e-9)."""
    E, nev = 64, 50
    K, M = problems.q1_fem_3d(E)
    dK, dM = DeviceMatrix(K), DeviceMatrix(M)
    x0 = oracle.random_fill(5, 0, 0, nev, K.nrows)
    print("C2 reference mode, 3 iterations against the oracle", flush=True)
    f3 = ["-gcge_ev_max_it", "3"]
    dev3 = gcge.solve(dK, dM, nev, f3, x0=x0, history_cap=8)
    cpu3 = oracle.solve(K, M, nev, f3, x0=x0, history_cap=8)
    assert dev3["history"].shape == cpu3["history"].shape
    np.testing.assert_array_equal(dev3["history"][:, :2], cpu3["history"][:, :2])
    np.testing.assert_allclose(dev3["history"][:, 2:], cpu3["history"][:, 2:], rtol=1e-9)
    print("C2 reference mode, full solve", flush=True)
    out = gcge.solve(dK, dM, nev, ["-gcge_ev_max_it", "200"], x0=x0)
    assert out["num_conv"] == nev
    # R = 1e-8 on ||x|| = 1 M-normalised vectors is a loose test for this
    # pencil (||Mx|| ~ h^3 ||x||): the reference reached 1.2e-6 here
    np.testing.assert_allclose(np.sort(out["eval"]), problems.q1_eigs(E, nev), rtol=1e-5)


# ---- C3 --------------------------------------------------------------------------


def test_c3_q1_128_lognormal_mixed_against_fp64():
    E, nev = 128, 100
    dK = DeviceMatrix.q1(E, problems.lognormal_kappa(E, 1), "stiffness")
    dM = DeviceMatrix.q1(E, None, "mass")
    flags = BENCH_FLAGS + ["-gcge_ev_max_it", "500", "-gcge_block_size", str(2 * nev)]
    out = {}
    for prec in ("mixed", "fp64"):
        print(f"C3 {prec} solve", flush=True)
        out[prec] = gcge.solve(dK, dM, nev, flags, seed=1, precision=prec)
        assert out[prec]["num_conv"] == nev, prec
        assert np.all(_rb_residuals(dK, dM, out[prec]["eval"], out[prec]["evec"]) <= 1.05e-8), prec
    m, f = out["mixed"], out["fp64"]
    np.testing.assert_allclose(m["eval"], f["eval"], rtol=1e-9)
    assert _largest_angle_m(dM, m["evec"], f["evec"]) < 1e-6
    dK.set_precision(False)
    dM.set_precision(False)


# ---- C4 --------------------------------------------------------------------------


def test_c4_q1_64_lognormal_adjoint():
    E, nev = 64, 20
    kappa = problems.lognormal_kappa(E, 1)
    w = np.random.default_rng(2).uniform(size=nev)
    flags = ["-gcge_conv_type", "RB", "-gcge_ev_tol", "1e-9", "-gcge_ev_max_it", "800",
             "-gcge_full_rr", "1", "-gcge_w_orth_type", "svqb", "-gcge_max_reorth_time", "2"]

    def solve(kv):
        dK = DeviceMatrix.q1(E, kv, "stiffness")
        dM = DeviceMatrix.q1(E, None, "mass")
        o = gcge.solve(dK, dM, nev, flags, seed=1)
        assert o["num_conv"] == nev
        return o, dK, dM

    def loss(kv):
        # the eigenvalues as Rayleigh quotients of the computed vectors
        # (error quadratic in the vectors' error), summed on the host
        o, dK, dM = solve(kv)
        X = o["evec"]
        rq = np.einsum("ij,ij->j", X, _products(dK, X)) / np.einsum("ij,ij->j", X, _products(dM, X))
        return float(w @ rq)

    print("C4 solve + eigenvalue adjoint", flush=True)
    out, dK, dM = solve(kappa)
    g = adjoint.q1_kappa_gradient(E, out["evec"], w)
    L = float(w @ out["eval"])
    assert abs(float(kappa @ g) - L) <= 1e-10 * abs(L)
    # fourth-order central differences (step 1e-2 kappa_e, truncation ~1e-8)
    # on elements drawn among the upper half of |kappa_e g_e| (smaller ones
    # sink below the ~1e-14 rounding of the loss at this step)
    rng = np.random.default_rng(3)
    big = np.flatnonzero(np.abs(kappa * g) > np.median(np.abs(kappa * g)))
    for e in rng.choice(big, 3, replace=False):
        h = 1e-2 * kappa[e]
        f = {}
        for j in (-2, -1, 1, 2):
            kv = kappa.copy()
            kv[e] += j * h
            f[j] = loss(kv)
        fd = (f[-2] - 8 * f[-1] + 8 * f[1] - f[2]) / (12 * h)
        print(f"C4 element {e}: adjoint {g[e]:.12e} fd {fd:.12e}", flush=True)
        assert abs(g[e] - fd) <= 1e-5 * abs(fd)
    # eigenvector adjoint of ||P x_1||^2: the shifted system it solves
    x1, lam1 = out["evec"][:, 0], float(out["eval"][0])
    mask = np.random.default_rng(4).uniform(size=x1.size) < 0.5
    gx = 2.0 * mask * x1
    y = adjoint.eigvec_adjoint_solve(dK, dM, x1, lam1, gx)
    Mx1 = _products(dM, x1[:, None])[:, 0]
    rhs = gx - Mx1 * (x1 @ gx)
    res = _products(dK, y[:, None])[:, 0] - lam1 * _products(dM, y[:, None])[:, 0] - rhs
    # the component along M x_1 is the projected-away part of the system
    res -= Mx1 * (x1 @ res)
    assert np.linalg.norm(res) <= 1e-6 * np.linalg.norm(rhs)
    assert abs(float(Mx1 @ y)) <= 1e-10 * np.linalg.norm(y) * np.linalg.norm(Mx1)
    with pytest.raises(ValueError):
        adjoint.eigvec_adjoint_solve(dK, dM, out["evec"][:, 1], float(out["eval"][1]), gx, index=1)


# ---- C5 --------------------------------------------------------------------------


def test_c5_q1_192_permuted_products_solve_adjoint():
    E, nev = 192, 200
    n = (E - 1) ** 3
    kappa = problems.lognormal_kappa(E, 1)
    perm = np.random.Generator(np.random.PCG64(1001)).permutation(n)  # row i holds node perm[i]
    node_row = np.empty(n, dtype=np.int32)
    node_row[perm] = np.arange(n, dtype=np.int32)
    print("C5 device assembly", flush=True)
    dK = DeviceMatrix.q1(E, kappa, "stiffness", node_row)
    dM = DeviceMatrix.q1(E, None, "mass", node_row)
    assert dK.storage_info()["storage"] == "rows"
    # products against the oracle on 2000 sampled rows
    print("C5 products against the oracle", flush=True)
    Kh = dK.to_csr()
    rows = np.sort(np.random.default_rng(6).choice(n, 2000, replace=False))
    lens = np.diff(Kh.rowptr)[rows]
    sub_ptr = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    idx = np.concatenate([np.arange(Kh.rowptr[r], Kh.rowptr[r + 1]) for r in rows])
    sub = problems.CSR(len(rows), n, sub_ptr, Kh.colidx[idx], Kh.vals[idx])
    del Kh
    for k in (4, 40):  # element gathers below 16 columns, row gathers from 16
        X = np.asfortranarray(np.random.default_rng(7 + k).standard_normal((n, k)))
        Y = _products(dK, X)[rows]
        Yo = oracle.spmm(oracle.Matrix(sub), X)
        bound = np.abs(sub.to_scipy()) @ np.abs(X)
        assert np.all(np.abs(Y - Yo) <= 64 * 2.2e-16 * bound), k
    del X
    # a bounded mixed-precision solve: the block comes back M-orthonormal
    print("C5 bounded mixed-precision solve", flush=True)
    flags = ["-gcge_conv_type", "RB", "-gcge_ev_tol", "1e-8", "-gcge_ev_max_it", "1",
             "-gcge_block_size", str(2 * nev), "-gcge_full_rr", "1", "-gcge_w_orth_type", "svqb",
             "-gcge_max_reorth_time", "2"]
    out = gcge.solve(dK, dM, nev, flags, seed=1, precision="mixed")
    dK.set_precision(False)
    dM.set_precision(False)
    assert out["block_size"] == nev
    Xd = DeviceMultiVec.from_host(out["evec"])
    del out
    lib = _native.hip()
    MX = DeviceMultiVec(n, nev)
    lib.GCGE_HIP_SpMM(dM.handle, Xd.handle, 0, MX.handle, 0, nev)
    G = np.zeros((nev, nev), order="F")
    lib.GCGE_HIP_Gram(Xd.handle, 0, nev, MX.handle, 0, nev, _native.as_dptr(G), nev)
    assert np.abs(G - np.eye(nev)).max() < 1e-10
    # eigenvalue adjoint on all elements through the node -> row map
    print("C5 adjoint identity", flush=True)
    w = np.random.default_rng(8).uniform(size=nev)
    g = adjoint.q1_kappa_gradient(E, Xd, w, node_row=node_row)
    lib.GCGE_HIP_SpMM(dK.handle, Xd.handle, 0, MX.handle, 0, nev)
    lib.GCGE_HIP_Gram(Xd.handle, 0, nev, MX.handle, 0, nev, _native.as_dptr(G), nev)
    xkx = float(w @ np.diag(G))
    assert abs(float(kappa @ g) - xkx) <= 1e-10 * abs(xkx)
