"""Eigenpair adjoints: the numpy restatement (oracle/adjoint.py) pinned by
Euler homogeneity and central finite differences of dense eigensolves (CPU),
and the device kernels against it (GPU).

Tolerances: homogeneity and device-vs-numpy are rounding-level (1e-10 /
1e-12 relative); finite differences with a relative step of 1e-6 are
accurate to about 1e-6 relative, so they are held to 1e-5.
"""
import numpy as np
import pytest

import oracle.adjoint as oadj
from gcge_old_amd import problems


def _kappa(E, seed):
    return problems.lognormal_kappa(E, seed)


def test_homogeneity_of_the_kappa_gradient():
    """K is linear in kappa and M does not depend on it, so
    sum_e kappa_e dlambda_c/dkappa_e = lambda_c for B-normalised x_c."""
    E = 5
    kappa = _kappa(E, 2)
    K, M = problems.q1_fem_3d(E, kappa)
    lam, X = oadj.dense_eigs(K, M, 4)
    for c in range(4):
        w = np.zeros(4)
        w[c] = 1.0
        g = oadj.q1_element_gradient(E, X, w)
        assert abs(kappa @ g - lam[c]) <= 1e-10 * lam[c]


def test_kappa_gradient_finite_differences():
    """dlambda_c/dkappa_e for every element of the 4^3 mesh against central
    differences (step 1e-6 kappa_e) of the dense generalised eigensolve."""
    E = 4
    kappa = _kappa(E, 3)
    K, M = problems.q1_fem_3d(E, kappa)
    lam, X = oadj.dense_eigs(K, M, 3)
    w = np.array([1.0, 0.5, 0.25])
    g = oadj.q1_element_gradient(E, X, w)
    fd = np.zeros(E ** 3)
    for e in range(E ** 3):
        d = 1e-6 * kappa[e]
        kp, km = kappa.copy(), kappa.copy()
        kp[e] += d
        km[e] -= d
        lp = oadj.dense_eigs(*problems.q1_fem_3d(E, kp), 3)[0]
        lm = oadj.dense_eigs(*problems.q1_fem_3d(E, km), 3)[0]
        fd[e] = w @ (lp - lm) / (2 * d)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-8 * np.abs(fd).max())


def test_pattern_gradient_finite_differences():
    """dlambda/dA on the pattern: perturbing a symmetric pair (i, j), (j, i)
    of A changes lambda by (g_ij + g_ji) delta; the dB gradient uses
    -lambda as weight."""
    E = 4
    K, M = problems.q1_fem_3d(E, _kappa(E, 4))
    lam, X = oadj.dense_eigs(K, M, 1)
    gA = oadj.pattern_gradient(K, X, [1.0])
    gB = oadj.pattern_gradient(M, X, [-lam[0]])
    rng = np.random.default_rng(0)
    for mat, grad, which in ((K, gA, "A"), (M, gB, "B")):
        rows_m = np.repeat(np.arange(mat.nrows), np.diff(mat.rowptr))
        for p in rng.choice(mat.nnz, 6, replace=False):
            i, j = rows_m[p], mat.colidx[p]
            q = np.flatnonzero((rows_m == j) & (mat.colidx == i))[0]
            d = 1e-6 * abs(mat.vals[p])
            out = []
            for sgn in (1, -1):
                v = mat.vals.copy()
                v[p] += sgn * d
                if q != p:
                    v[q] += sgn * d
                pert = problems.CSR(mat.nrows, mat.ncols, mat.rowptr, mat.colidx, v)
                pair = (pert, M) if which == "A" else (K, pert)
                out.append(oadj.dense_eigs(*pair, 1)[0][0])
            fd = (out[0] - out[1]) / (2 * d)
            want = grad[p] + (grad[q] if q != p else 0.0)
            assert abs(want - fd) <= 1e-5 * abs(fd) + 1e-7 * np.abs(grad).max()


def test_eigvec_adjoint_finite_differences():
    """L = ||P x_1||^2 for a fixed 0/1 mask P (sign-invariant): dL/dkappa_e =
    -y|e^T (h Khat) x_1|e with (K - lambda_1 M) y = (I - M x x^T) 2 P x_1."""
    E = 4
    kappa = _kappa(E, 5)
    K, M = problems.q1_fem_3d(E, kappa)
    rng = np.random.default_rng(1)
    mask = (rng.uniform(size=K.nrows) < 0.5).astype(float)

    def loss(kv):
        x = oadj.dense_eigs(*problems.q1_fem_3d(E, kv), 1)[1][:, 0]
        return np.sum((mask * x) ** 2)

    lam, X = oadj.dense_eigs(K, M, 1)
    x = X[:, 0]
    y = oadj.eigvec_adjoint(K, M, x, lam[0], 2 * mask * x)
    g = oadj.q1_element_gradient(E, x[:, None], [-1.0], Y=y[:, None])
    for e in rng.choice(E ** 3, 12, replace=False):
        d = 1e-6 * kappa[e]
        kp, km = kappa.copy(), kappa.copy()
        kp[e] += d
        km[e] -= d
        fd = (loss(kp) - loss(km)) / (2 * d)
        assert abs(g[e] - fd) <= 1e-5 * np.abs(g).max()


# ---- device kernels ------------------------------------------------------------

@pytest.mark.gpu
def test_device_kappa_and_pattern_gradients_match_numpy():
    """GCGE_HIP_Q1ElementGradient and GCGE_HIP_PatternGradient on device
    eigenvectors against the numpy restatement on the same vectors."""
    import gcge_old_amd as gcge
    from gcge_old_amd import adjoint

    E, nev = 10, 12  # nev >= 10: see the X-block width note in the solver tests
    kappa = _kappa(E, 6)
    K, M = problems.q1_fem_3d(E, kappa)
    out = gcge.solve(K, M, nev, ["-gcge_conv_type", "RB", "-gcge_ev_tol", "1e-10",
                                 "-gcge_ev_max_it", "200", "-gcge_full_rr", "1"], seed=2)
    assert out["num_conv"] == nev
    X, lam = out["evec"], out["eval"]
    w = np.linspace(1.0, 0.2, nev)
    g_dev = adjoint.q1_kappa_gradient(E, X, w)
    g_np = oadj.q1_element_gradient(E, X, w)
    np.testing.assert_allclose(g_dev, g_np, rtol=1e-12, atol=1e-14 * np.abs(g_np).max())
    # homogeneity on device eigenpairs
    assert abs(kappa @ g_dev - w @ lam) <= 1e-8 * abs(w @ lam)
    dK, dM = gcge.DeviceMatrix(K), gcge.DeviceMatrix(M)
    gA, gB = adjoint.pattern_gradients(dK, dM, X, lam, w)
    np.testing.assert_allclose(gA, oadj.pattern_gradient(K, X, w), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(gB, oadj.pattern_gradient(M, X, -w * lam), rtol=1e-12, atol=1e-15)


@pytest.mark.gpu
def test_device_kappa_gradient_finite_differences():
    """Eigenvalue loss sum_c w_c lambda_c on the 6^3 mesh: device gradient
    against central differences of device solves (8 elements)."""
    import gcge_old_amd as gcge
    from gcge_old_amd import adjoint

    E, nev = 6, 10
    kappa = _kappa(E, 7)
    # eigenvalue error ~ residual^2: 1e-10 leaves the differences' own error
    # (~1e-10 relative from rounding over a 1e-6 step) dominant
    flags = ["-gcge_conv_type", "RB", "-gcge_ev_tol", "1e-10", "-gcge_ev_max_it", "500",
             "-gcge_full_rr", "1"]
    w = np.linspace(1.0, 0.1, nev)

    def loss(kv):
        out = gcge.solve(*problems.q1_fem_3d(E, kv), nev, flags, seed=3)
        assert out["num_conv"] == nev
        return w @ out["eval"], out

    _, out = loss(kappa)
    g = adjoint.q1_kappa_gradient(E, out["evec"], w)
    rng = np.random.default_rng(2)
    for e in rng.choice(E ** 3, 8, replace=False):
        d = 1e-6 * kappa[e]
        kp, km = kappa.copy(), kappa.copy()
        kp[e] += d
        km[e] -= d
        fd = (loss(kp)[0] - loss(km)[0]) / (2 * d)
        assert abs(g[e] - fd) <= 1e-5 * abs(fd) + 1e-7 * np.abs(g).max()


@pytest.mark.gpu
def test_device_eigvec_adjoint_matches_dense():
    """Eigenvector loss ||P x_1||^2 on the 6^3 mesh: the device block-CG
    adjoint solve and element gradient against the dense restatement."""
    import gcge_old_amd as gcge
    from gcge_old_amd import adjoint

    E = 6
    kappa = _kappa(E, 8)
    K, M = problems.q1_fem_3d(E, kappa)
    lam, X = oadj.dense_eigs(K, M, 1)
    x = X[:, 0]
    mask = (np.random.default_rng(3).uniform(size=K.nrows) < 0.5).astype(float)
    gx = 2 * mask * x
    dK, dM = gcge.DeviceMatrix(K), gcge.DeviceMatrix(M)
    g_dev, y_dev = adjoint.q1_eigvec_kappa_gradient(E, dK, dM, x, lam[0], gx)
    y_np = oadj.eigvec_adjoint(K, M, x, lam[0], gx)
    np.testing.assert_allclose(y_dev, y_np, rtol=0, atol=1e-8 * np.abs(y_np).max())
    g_np = oadj.q1_element_gradient(E, x[:, None], [-1.0], Y=y_np[:, None])
    np.testing.assert_allclose(g_dev, g_np, rtol=0, atol=1e-8 * np.abs(g_np).max())


@pytest.mark.gpu
@pytest.mark.timeout(600)  # the first torch import on a fresh machine can take minutes
def test_torch_autograd_wrapper():
    """Q1Eigenvalues as a torch.autograd.Function: backward of sum w lambda
    equals the element gradient."""
    import torch

    from gcge_old_amd import adjoint

    E, nev = 6, 10
    kappa = torch.tensor(_kappa(E, 9), dtype=torch.float64, requires_grad=True)
    flags = ["-gcge_conv_type", "RB", "-gcge_ev_tol", "1e-10", "-gcge_ev_max_it", "300",
             "-gcge_full_rr", "1"]
    lam = adjoint.Q1Eigenvalues.apply(kappa, E, nev, flags, 4)
    w = torch.linspace(1.0, 3.0, nev, dtype=torch.float64)
    (w * lam).sum().backward()
    assert abs(float(kappa.detach() @ kappa.grad) - float(w @ lam.detach())) <= 1e-8 * float(w @ lam.detach())


def test_symmetric_permutation_matches_scipy():
    """P A P^T from the row gather equals the sparse triple product, with and
    without sorting the renumbered columns of each row."""
    import scipy.sparse as sp

    K, _ = problems.q1_fem_3d(6, _kappa(6, 4))
    perm = np.random.default_rng(8).permutation(K.nrows)
    m = K.to_scipy()
    P = sp.csr_matrix((np.ones(K.nrows), (np.arange(K.nrows), perm)), shape=m.shape)
    want = (P @ m @ P.T).toarray()
    for sort_rows in (True, False):
        Kp = problems.symmetric_permutation(K, perm, sort_rows=sort_rows)
        np.testing.assert_array_equal(Kp.to_scipy().toarray(), want)
        if sort_rows:
            for r in range(0, Kp.nrows, 37):
                assert np.all(np.diff(Kp.colidx[Kp.rowptr[r]:Kp.rowptr[r + 1]]) > 0)


@pytest.mark.gpu
def test_device_kappa_gradient_of_permuted_blocks():
    """Blocks of the permuted pencil (row i holds node perm[i]) with the node ->
    row map give the same element gradient as the unpermuted blocks."""
    from gcge_old_amd import adjoint

    E = 7
    rng = np.random.default_rng(9)
    n = (E - 1) ** 3
    X = rng.standard_normal((n, 5))
    w = rng.uniform(size=5)
    perm = rng.permutation(n)
    inv = np.empty(n, dtype=np.int64)
    inv[perm] = np.arange(n)
    print("q1 element gradient, permuted rows", flush=True)
    g = adjoint.q1_kappa_gradient(E, X, w)
    gp = adjoint.q1_kappa_gradient(E, X[perm], w, node_row=inv)
    np.testing.assert_allclose(gp, g, rtol=1e-14, atol=1e-15 * np.abs(g).max())
    np.testing.assert_allclose(g, oadj.q1_element_gradient(E, X, w), rtol=1e-12,
                               atol=1e-14 * np.abs(g).max())
