"""End-to-end GCG solves on the device backend (GPU).

Each solve starts from the same initial block on both sides, so the device
run and the oracle run (the same GCG core on the CPU restatement of the
reference CSR operations) follow the same iteration up to rounding.  Checks:
* eigenvalues agree with the oracle to 1e-10 relative and, where a closed
  form exists, with the exact spectrum;
* the same number of outer iterations (clusters of equal eigenvalues can
  shift a convergence decision by one iteration, hence +-2);
* every returned pair satisfies the residual test it was converged with;
* the converged subspace agrees with the oracle's (largest principal angle).
"""
import os

import numpy as np
import pytest
import scipy.linalg as sla

import gcge_old_amd as gcge
import oracle
from gcge_old_amd import _native, problems

pytestmark = pytest.mark.gpu


def _inv_sqrt(G):
    w, Q = np.linalg.eigh((G + G.T) / 2)
    return (Q / np.sqrt(w)) @ Q.T


def _largest_angle(X, Y, B=None):
    """Largest principal angle between span(X) and span(Y) in the B inner
    product (B may be only semidefinite, as the shipped mass matrices with
    Dirichlet rows are; it must be definite on the two spans)."""
    Bm = (lambda Z: Z) if B is None else (lambda Z: B.to_scipy() @ Z)
    Xo = X @ _inv_sqrt(X.T @ Bm(X))
    Yo = Y @ _inv_sqrt(Y.T @ Bm(Y))
    s = np.linalg.svd(Xo.T @ Bm(Yo), compute_uv=False)
    return float(np.arccos(np.clip(s.min(), -1.0, 1.0)))


def _residuals(A, B, ev, X):
    As = A.to_scipy()
    BX = X if B is None else B.to_scipy() @ X
    R = As @ X - BX * ev
    return np.linalg.norm(R, axis=0) / (np.abs(ev) * np.linalg.norm(BX, axis=0))


def _fixture_matrices(golden_dir, name):
    d = np.load(f"{golden_dir}/data/{name}.npz")
    return problems.CSR(int(d["n"]), int(d["n"]), d["rowptr"], d["colidx"], d["vals"])


CASES = {
    # name: (A, B, nev, flags, exact eigenvalues or None)
    "fd2d_32_nev10": (lambda g: (problems.fd_laplacian_2d(32), None), 10,
                      ["-gcge_block_size", "10"], lambda: problems.fd_laplacian_2d_eigs(32, 10)),
    "q1_e12_nev20": (lambda g: problems.q1_fem_3d(12), 20,
                     ["-gcge_ev_max_it", "60", "-gcge_ev_tol", "1e-10"], lambda: problems.q1_eigs(12, 20)),
    "A5M5_nev20": (lambda g: (_fixture_matrices(g, "A_5"), _fixture_matrices(g, "M_5")), 20,
                   ["-gcge_block_size", "10", "-gcge_ev_max_it", "40"], None),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_device_solve_matches_oracle(case, golden_dir):
    make, nev, flags, exact = CASES[case]
    A, B = make(golden_dir)
    x0 = oracle.random_fill(11, 0, 0, nev, A.nrows)
    dev = gcge.solve(A, B, nev, flags, x0=x0)
    cpu = oracle.solve(A, B, nev, flags, x0=x0)
    assert dev["num_conv"] == nev and cpu["num_conv"] == nev
    assert abs(dev["num_iter"] - cpu["num_iter"]) <= 2
    np.testing.assert_allclose(dev["eval"], cpu["eval"], rtol=1e-10)
    if exact is not None:
        np.testing.assert_allclose(np.sort(dev["eval"]), exact(), rtol=1e-10)
    tol = float(flags[flags.index("-gcge_ev_tol") + 1]) if "-gcge_ev_tol" in flags else 1e-8
    # reference criterion ||r|| / (||x|| max(1, |lambda|)) as reported by the solver
    assert np.all(dev["res"] < tol)
    # eigenvector subspace of the whole converged block
    assert _largest_angle(dev["evec"], cpu["evec"], B) < 1e-6


def test_fd2d_128_closed_form_rb_criterion():
    """2D FD Laplacian 128^2, B = I, nev = 10 at the relative residual
    ||Ax - lambda Bx|| / (|lambda| ||Bx||) <= 1e-8."""
    A = problems.fd_laplacian_2d(128)
    out = gcge.solve(A, None, 10, ["-gcge_conv_type", "RB", "-gcge_ev_tol", "1e-8",
                                   "-gcge_block_size", "20", "-gcge_ev_max_it", "100"], seed=3)
    assert out["num_conv"] == 10
    np.testing.assert_allclose(out["eval"], problems.fd_laplacian_2d_eigs(128, 10), rtol=1e-10)
    assert np.all(_residuals(A, None, out["eval"], out["evec"]) <= 1.05e-8)


def test_q1_generalized_closed_form():
    """Q1 FEM 20^3 (n = 6859), kappa = 1, nev = 30 against the Kronecker-sum
    spectrum, B-orthonormal eigenvectors."""
    K, M = problems.q1_fem_3d(20)
    out = gcge.solve(K, M, 30, ["-gcge_conv_type", "RB", "-gcge_ev_tol", "1e-9",
                                "-gcge_ev_max_it", "100"], seed=5)
    assert out["num_conv"] == 30
    np.testing.assert_allclose(np.sort(out["eval"]), problems.q1_eigs(20, 30), rtol=1e-10)
    X = out["evec"]
    G = X.T @ (M.to_scipy() @ X)
    assert np.abs(G - np.eye(30)).max() < 1e-10
    assert np.all(_residuals(K, M, out["eval"], X) <= 1.05e-9)


def test_mixed_precision_q1_closed_form():
    """Mixed-precision form (float32 inner block CG) on Q1 20^3, kappa = 1:
    the same eigenvalues as the closed form, the float64 residual test met."""
    K, M = problems.q1_fem_3d(20)
    flags = ["-gcge_conv_type", "RB", "-gcge_ev_tol", "1e-9", "-gcge_ev_max_it", "100"]
    out = gcge.solve(K, M, 30, flags, seed=5, precision="mixed")
    assert out["num_conv"] == 30
    np.testing.assert_allclose(np.sort(out["eval"]), problems.q1_eigs(20, 30), rtol=1e-10)
    assert np.all(_residuals(K, M, out["eval"], out["evec"]) <= 1.05e-9)


def test_mixed_precision_lognormal_against_fp64():
    """Q1 16^3 with lognormal kappa (no closed form): every eigenvalue of the
    mixed-precision run within 1e-9 relative of the float64 oracle run, and
    the converged subspaces agree."""
    E = 16
    K, M = problems.q1_fem_3d(E, problems.lognormal_kappa(E, seed=3))
    nev = 20
    flags = ["-gcge_conv_type", "RB", "-gcge_ev_tol", "1e-8", "-gcge_ev_max_it", "200"]
    x0 = oracle.random_fill(13, 0, 0, nev, K.nrows)
    mixed = gcge.solve(K, M, nev, flags, x0=x0, precision="mixed")
    ref = oracle.solve(K, M, nev, flags, x0=x0)
    assert mixed["num_conv"] == nev and ref["num_conv"] == nev
    np.testing.assert_allclose(mixed["eval"], ref["eval"], rtol=1e-9)
    assert np.all(_residuals(K, M, mixed["eval"], mixed["evec"]) <= 1.05e-8)
    assert _largest_angle(mixed["evec"], ref["evec"], M) < 1e-5


def test_full_rr_on_device():
    """-gcge_full_rr 1 through the device backend: the Q1 14^3, nev = 30
    start on which soft locking stalls converges to the closed form, with
    the float64 and the mixed-precision inner solves."""
    E, nev = 14, 30
    K, M = problems.q1_fem_3d(E)
    x0 = oracle.random_fill(2, 0, 0, nev, K.nrows)
    flags = ["-gcge_conv_type", "RB", "-gcge_ev_tol", "1e-8", "-gcge_ev_max_it", "80",
             "-gcge_full_rr", "1"]
    for precision in ("fp64", "mixed"):
        out = gcge.solve(K, M, nev, flags, x0=x0, precision=precision)
        assert out["num_conv"] == nev and out["num_iter"] < 60, (precision, out["num_iter"])
        np.testing.assert_allclose(np.sort(out["eval"]), problems.q1_eigs(E, nev), rtol=1e-10)
        assert np.all(_residuals(K, M, out["eval"], out["evec"]) <= 1.05e-8)


def test_large_rayleigh_ritz_on_device_eigensolver():
    """nev = 150 makes the projected matrix (nev + 2 x block) larger than 400,
    so the Rayleigh-Ritz eigensolve runs on rocSOLVER through the backend's
    dense eigensolver extension (smaller orders run on the host threads); the
    spectrum must still be the closed form."""
    E, nev = 12, 150
    K, M = problems.q1_fem_3d(E)
    out = gcge.solve(K, M, nev, ["-gcge_conv_type", "RB", "-gcge_ev_tol", "1e-8",
                                 "-gcge_ev_max_it", "200", "-gcge_full_rr", "1"], seed=7)
    assert out["num_conv"] == nev
    np.testing.assert_allclose(np.sort(out["eval"]), problems.q1_eigs(E, nev), rtol=1e-10)
    assert np.all(_residuals(K, M, out["eval"], out["evec"]) <= 1.05e-8)


def test_svqb_full_rr_device_matches_oracle():
    """The benchmark's solver options (-gcge_full_rr 1, -gcge_w_orth_type svqb)
    on the device against the same core on the oracle operations."""
    E, nev = 16, 40
    K, M = problems.q1_fem_3d(E)
    x0 = oracle.random_fill(1, 0, 0, nev, K.nrows)
    flags = ["-gcge_conv_type", "RB", "-gcge_ev_tol", "1e-8", "-gcge_ev_max_it", "150",
             "-gcge_full_rr", "1", "-gcge_w_orth_type", "svqb"]
    dev = gcge.solve(K, M, nev, flags, x0=x0)
    cpu = oracle.solve(K, M, nev, flags, x0=x0)
    assert dev["num_conv"] == nev and cpu["num_conv"] == nev
    assert abs(dev["num_iter"] - cpu["num_iter"]) <= 3
    np.testing.assert_allclose(dev["eval"], cpu["eval"], rtol=1e-10)
    np.testing.assert_allclose(np.sort(dev["eval"]), problems.q1_eigs(E, nev), rtol=1e-10)
    assert np.all(_residuals(K, M, dev["eval"], dev["evec"]) <= 1.05e-8)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_svqb_correction_block_cg_host_bookkeeping(precision, monkeypatch):
    """The SVQB path asks the device block CG for the correction to its
    starting values (BlockCGCorrection): with the bookkeeping on the device
    and, under GCGE_HIP_BCG_HOST=1, in the host loop, both in float64 and
    with the float32 inner solve, the solve reaches the closed-form spectrum
    and the two bookkeeping forms agree."""
    E, nev = 14, 24
    K, M = problems.q1_fem_3d(E)
    x0 = oracle.random_fill(5, 0, 0, nev, K.nrows)
    flags = ["-gcge_conv_type", "RB", "-gcge_ev_tol", "1e-8", "-gcge_ev_max_it", "150",
             "-gcge_full_rr", "1", "-gcge_w_orth_type", "svqb", "-gcge_max_reorth_time", "2"]
    dev = gcge.solve(K, M, nev, flags, x0=x0, precision=precision)
    monkeypatch.setenv("GCGE_HIP_BCG_HOST", "1")
    host = gcge.solve(K, M, nev, flags, x0=x0, precision=precision)
    for out in (dev, host):
        assert out["num_conv"] == nev
        np.testing.assert_allclose(np.sort(out["eval"]), problems.q1_eigs(E, nev), rtol=1e-10)
        assert np.all(_residuals(K, M, out["eval"], out["evec"]) <= 1.05e-8)
    np.testing.assert_allclose(host["eval"], dev["eval"], rtol=1e-10)


@pytest.mark.parametrize("eta", ["0", "1"])
def test_svqb_projection_pass_rules_on_device(eta):
    """-gcge_svqb_dgks_eta 0 (W as the reference forms it, both projection
    passes) and 1 (W as a correction, the second pass forced on every
    block) on the device: the closed-form spectrum with B-orthonormal
    eigenvectors, as with the default (the other tests)."""
    E, nev = 14, 24
    K, M = problems.q1_fem_3d(E)
    x0 = oracle.random_fill(6, 0, 0, nev, K.nrows)
    flags = ["-gcge_conv_type", "RB", "-gcge_ev_tol", "1e-8", "-gcge_ev_max_it", "150",
             "-gcge_full_rr", "1", "-gcge_w_orth_type", "svqb", "-gcge_max_reorth_time", "2",
             "-gcge_svqb_dgks_eta", eta]
    out = gcge.solve(K, M, nev, flags, x0=x0, precision="mixed")
    assert out["num_conv"] == nev
    np.testing.assert_allclose(np.sort(out["eval"]), problems.q1_eigs(E, nev), rtol=1e-10)
    assert np.all(_residuals(K, M, out["eval"], out["evec"]) <= 1.05e-8)
    X = out["evec"]
    assert np.abs(X.T @ (M.to_scipy() @ X) - np.eye(nev)).max() < 1e-10


def test_c_program_through_solver_init():
    """A C program on the public headers only (gcge_old_amd/csrc/examples/
    q1_solve.c): device-assembled Q1 pencil, GCGE_HIP_Solver_Init (the
    counterpart of the reference's GCGE_CSR_Solver_Init), GCGE_SOLVER_Solve;
    its eigenvalues against the closed form."""
    import subprocess

    exe = _native.lib_path("q1_solve")
    assert os.path.exists(exe), "q1_solve is not built"
    E, nev = 10, 10
    p = subprocess.run([exe, str(E), str(nev), "-gcge_conv_type", "RB", "-gcge_ev_tol", "1e-10",
                        "-gcge_ev_max_it", "100"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr[-2000:]
    ev = [float(line.split()[2]) for line in p.stdout.splitlines() if line.startswith("eval ")]
    assert len(ev) == nev
    np.testing.assert_allclose(ev, problems.q1_eigs(E, nev), rtol=1e-10)
    last = [line for line in p.stdout.splitlines() if line.startswith("num_iter")][-1].split()
    assert int(last[3]) >= nev


@pytest.mark.parametrize("nev,flags,seed",
                         [(20, ["-gcge_block_size", "10", "-gcge_ev_max_it", "40"], 7),
                          (30, ["-gcge_block_size", "10", "-gcge_ev_max_it", "60"], 8)])
def test_dirichlet_boundary_on_device(nev, flags, seed, golden_dir):
    """The reference's shipped pencil A_5/M_5 with dirichlet_boundary = 1 (the
    setting of its canonical example, example/csr_solver.c:51): the randomly
    drawn initial block gets its Dirichlet rows zeroed by the device
    SetDirichletBoundary (app/csr/gcge_app_csr.c:130-183) before the
    orthonormalisation.  Against the oracle run from the same random block
    (same counter-based generator, its own CPU zeroing) and against the
    reference binary started from that block with the rows zeroed
    (ref_runs.npz *_dirichlet): eigenvalues, iteration and per-iteration
    convergence counts; the Dirichlet rows of the eigenvectors stay zero."""
    A, B = _fixture_matrices(golden_dir, "A_5"), _fixture_matrices(golden_dir, "M_5")
    fl = flags + ["-gcge_dirichlet_boundary", "1"]
    dev = gcge.solve(A, B, nev, fl, seed=seed, history_cap=200)
    cpu = oracle.solve(A, B, nev, fl, seed=seed, history_cap=200)
    ref = np.load(os.path.join(golden_dir, "ref_runs.npz"))
    key = f"A5M5_nev{nev}_dirichlet"
    assert dev["num_iter"] == cpu["num_iter"] == int(ref[key + "/iters"][0])
    assert dev["num_conv"] == cpu["num_conv"] == int(ref[key + "/iters"][1])
    np.testing.assert_array_equal(dev["history"][:, :2], ref[key + "/conv_history"])
    scale = np.abs(ref[key + "/eval"]).max()
    np.testing.assert_allclose(dev["eval"], ref[key + "/eval"], rtol=1e-9, atol=1e-9 * scale)
    np.testing.assert_allclose(dev["eval"], cpu["eval"], rtol=1e-10, atol=1e-10 * scale)
    # the zeroed rows: B has no entry there, A exactly one (an eigenvector of
    # the pencil with lambda = A_ii / 0 is excluded, so they stay zero)
    Bs, As = B.to_scipy(), A.to_scipy()
    rows = [i for i in range(A.nrows) if not np.any(np.abs(Bs[i].data) > 1e-20)
            and np.count_nonzero(np.abs(As[i].data) > 1e-20) == 1]
    assert rows, "the fixture has Dirichlet rows"
    assert np.abs(dev["evec"][rows]).max() < 1e-12
