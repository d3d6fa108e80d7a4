"""Pins of the oracle (CPU).

The oracle restates the reference CSR backend and column-loop multivec
semantics; the reference's own tests hold no golden values (they print), so
each function is pinned by something a plausible mistake would violate:
brute-force dense products on non-symmetric inputs (a transposed operand or
a wrong index fails), closed-form spectra, statistical properties of the
generator, and runs of the reference solver itself (built from its sources;
the recorded outputs live in tests/golden/ref_runs.npz).
"""
import os

import numpy as np
import pytest

import oracle
from gcge_old_amd import problems

EPS = np.finfo(np.float64).eps


def _fixture(golden_dir, name):
    d = np.load(os.path.join(golden_dir, "data", name + ".npz"))
    return problems.CSR(int(d["n"]), int(d["n"]), d["rowptr"], d["colidx"], d["vals"])


def _nonsym(n, density, seed):
    import scipy.sparse as sp

    m = sp.random(n, n, density=density, random_state=np.random.default_rng(seed), format="csr")
    m = m + sp.diags(np.arange(1.0, n + 1.0))
    return problems.CSR.from_scipy(m)


@pytest.mark.parametrize("name", ["testA", "A_3", "M_5", "nonsym"])
def test_spmm_brute_force(name, golden_dir):
    A = _nonsym(300, 0.02, 1) if name == "nonsym" else _fixture(golden_dir, name)
    dense = A.to_scipy().toarray()
    X = np.random.default_rng(0).standard_normal((A.ncols, 5))
    got = oracle.spmm(oracle.Matrix(A), X)
    want = dense @ X
    bound = np.abs(dense) @ np.abs(X)
    assert np.all(np.abs(got - want) <= 8 * A.ncols * EPS * bound + 1e-300)


def test_testA_exact(golden_dir):
    """testA is the reference's 3x3 fixture; its product is exact in binary."""
    A = _fixture(golden_dir, "testA")
    X = np.eye(3)
    np.testing.assert_array_equal(oracle.spmm(oracle.Matrix(A), X), A.to_scipy().toarray())


def test_gram_and_linear_comb_brute_force():
    rng = np.random.default_rng(1)
    V = rng.standard_normal((500, 7))
    W = rng.standard_normal((500, 3))
    np.testing.assert_allclose(oracle.gram(V, W), V.T @ W, rtol=0, atol=1e-12)
    C = rng.standard_normal((7, 4))
    Y = rng.standard_normal((500, 4))
    np.testing.assert_allclose(oracle.linear_comb(V, C, 0.0, Y), V @ C, rtol=0, atol=1e-12)
    np.testing.assert_allclose(oracle.linear_comb(V, C, -2.0, Y), V @ C - 2.0 * Y, rtol=0, atol=1e-12)


def test_sym_eig_closed_form_and_numpy():
    n = 40
    T = 2 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    w, z = oracle.sym_eig(T)
    exact = 2 - 2 * np.cos(np.arange(1, n + 1) * np.pi / (n + 1))
    np.testing.assert_allclose(w, exact, rtol=0, atol=1e-13)
    np.testing.assert_allclose(z.T @ z, np.eye(n), atol=1e-13)
    np.testing.assert_allclose(T @ z, z * w, atol=1e-12)
    rng = np.random.default_rng(2)
    S = rng.standard_normal((30, 30))
    S = S + S.T
    upper_only = np.triu(S)  # only the upper triangle is read
    w2, _ = oracle.sym_eig(upper_only)
    np.testing.assert_allclose(w2, np.linalg.eigvalsh(S), atol=1e-12)


def test_random_generator_properties():
    X = oracle.random_fill(42, 0, 0, 4, 200000)
    assert X.min() >= -1.0 and X.max() < 1.0
    np.testing.assert_allclose(X.mean(axis=0), 0.0, atol=0.01)
    np.testing.assert_allclose(X.var(axis=0), 1.0 / 3.0, atol=0.01)
    np.testing.assert_array_equal(X, oracle.random_fill(42, 0, 0, 4, 200000))
    assert not np.array_equal(X[:, 0], oracle.random_fill(42, 1, 0, 1, 200000)[:, 0])
    assert not np.array_equal(X[:, 0], oracle.random_fill(43, 0, 0, 1, 200000)[:, 0])
    assert abs(np.corrcoef(X[:, 0], X[:, 1])[0, 1]) < 0.01


def test_csr_text_round_trip(tmp_path, golden_dir):
    A = _fixture(golden_dir, "A_5")
    p = str(tmp_path / "A.txt")
    A.write_text(p)
    h = oracle.lib().ORACLE_CSR_ReadFile(p.encode())
    assert h
    n, nnz = oracle.lib().ORACLE_CSR_Nrows(h), oracle.lib().ORACLE_CSR_Nnz(h)
    rp = np.zeros(n + 1, np.int32)
    ci = np.zeros(nnz, np.int32)
    v = np.zeros(nnz)
    oracle.lib().ORACLE_CSR_Export(h, oracle._i(rp), oracle._i(ci), oracle._d(v))
    oracle.lib().ORACLE_CSR_Free(h)
    np.testing.assert_array_equal(rp, A.rowptr)
    np.testing.assert_array_equal(ci, A.colidx)
    np.testing.assert_array_equal(v, A.vals)
    assert oracle.lib().ORACLE_CSR_ReadFile(str(tmp_path / "missing").encode()) is None


@pytest.mark.parametrize("name", ["A_5", "M_5"])
def test_petsc_binary_fixture_matches_text_fixture(name, golden_dir, tmp_path):
    """The reference ships A_5 / M_5 both as CSR text and as PETSc binary
    (the SLEPc driver's MatLoad input): the binary reader gives the same
    matrix bit for bit, and writing it back reproduces the file."""
    path = os.path.join(golden_dir, "data", name + ".petsc.bin")
    A = problems.CSR.read_petsc(path)
    T = _fixture(golden_dir, name)
    np.testing.assert_array_equal(A.rowptr, T.rowptr)
    np.testing.assert_array_equal(A.colidx, T.colidx)
    np.testing.assert_array_equal(A.vals, T.vals)
    out = str(tmp_path / "copy.bin")
    A.write_petsc(out)
    with open(out, "rb") as f1, open(path, "rb") as f2:
        assert f1.read() == f2.read()


def test_petsc_binary_rejects_other_files(tmp_path):
    bad = tmp_path / "bad.bin"
    bad.write_bytes(np.array([1211214, 2, 2, 0], ">i4").tobytes())  # a vector class id
    with pytest.raises(ValueError):
        problems.CSR.read_petsc(str(bad))
    short = tmp_path / "short.bin"
    short.write_bytes(np.array([1211216, 2, 2, 3, 1, 1], ">i4").tobytes())
    with pytest.raises(ValueError):
        problems.CSR.read_petsc(str(short))


def test_problem_generators_closed_form():
    """The generators themselves: dense eigenvalues of small instances equal
    the closed forms the end-to-end checks rely on."""
    import scipy.linalg as sla

    A = problems.fd_laplacian_2d(7).to_scipy().toarray()
    np.testing.assert_allclose(np.linalg.eigvalsh(A)[:10], problems.fd_laplacian_2d_eigs(7, 10),
                               rtol=1e-12)
    K, M = problems.q1_fem_3d(5)
    w = sla.eigh(K.to_scipy().toarray(), M.to_scipy().toarray(), eigvals_only=True)
    np.testing.assert_allclose(w[:20], problems.q1_eigs(5, 20), rtol=1e-11)
    # kappa scales K linearly and leaves M alone
    K2, M2 = problems.q1_fem_3d(5, kappa=np.full(125, 3.0))
    np.testing.assert_allclose(K2.vals, 3.0 * K.vals, rtol=1e-15)
    np.testing.assert_array_equal(M2.vals, M.vals)
    # face-neighbour entries of K are dropped, mass keeps the 27-point stencil
    E = 64
    N = E - 1
    assert K.nnz < M.nnz
    assert (3 * 4 - 2) ** 3 == M.nnz


def test_q1_assembly_against_quadrature_element_loop():
    """q1_fem_3d with random kappa against an element-by-element dense
    assembly whose element matrices come from 2-point Gauss quadrature of the
    trilinear shape functions (exact for these integrands), Dirichlet nodes
    removed afterwards: pins the element numerators, the kappa indexing
    ex + E ey + E^2 ez and the node numbering, which the kappa = 1 spectrum
    cannot."""
    import itertools

    E = 4
    h = 1.0 / E
    kappa = problems.lognormal_kappa(E, 7)
    g = np.array([0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0)])
    corners = list(itertools.product((0, 1), repeat=3))  # (ax, ay, az)

    def phi(a, xi):
        return np.prod([xi[d] if a[d] else 1.0 - xi[d] for d in range(3)])

    def dphi(a, xi):
        out = np.empty(3)
        for d in range(3):
            f = 1.0 if a[d] else -1.0
            out[d] = f * np.prod([xi[e] if a[e] else 1.0 - xi[e] for e in range(3) if e != d])
        return out

    Ke = np.zeros((8, 8))
    Me = np.zeros((8, 8))
    for xi in itertools.product(g, repeat=3):
        for p, a in enumerate(corners):
            for q, b in enumerate(corners):
                Ke[p, q] += 0.125 * dphi(a, xi) @ dphi(b, xi)  # weights 1/2 per axis
                Me[p, q] += 0.125 * phi(a, xi) * phi(b, xi)
    nodes = E + 1
    Kd = np.zeros((nodes ** 3, nodes ** 3))
    Md = np.zeros_like(Kd)
    for ez, ey, ex in itertools.product(range(E), repeat=3):
        ids = [(ex + a[0]) + nodes * (ey + a[1]) + nodes * nodes * (ez + a[2]) for a in corners]
        ke = kappa[ex + E * ey + E * E * ez]
        Kd[np.ix_(ids, ids)] += ke * h * Ke  # stiffness scales by h in 3D
        Md[np.ix_(ids, ids)] += h ** 3 * Me
    r = np.arange(1, E)
    inner = [x + nodes * y + nodes * nodes * z for z in r for y in r for x in r]
    K, M = problems.q1_fem_3d(E, kappa)
    np.testing.assert_allclose(K.to_scipy().toarray(), Kd[np.ix_(inner, inner)], rtol=0,
                               atol=1e-14 * np.abs(Kd).max())
    np.testing.assert_allclose(M.to_scipy().toarray(), Md[np.ix_(inner, inner)], rtol=0,
                               atol=1e-14 * np.abs(Md).max())


RUN_CASES = {
    "A5M5_nev20": (("A_5", "M_5"), 20, ["-gcge_block_size", "10", "-gcge_ev_max_it", "40"], 7, True),
    "A5M5_nev40": (("A_5", "M_5"), 40, ["-gcge_ev_max_it", "60"], 7, True),
    "A3M3_nev8": (("A_3", "M_3"), 8, ["-gcge_ev_max_it", "60"], 5, False),
    "fd2d_24_nev12": ("fd24", 12, ["-gcge_ev_max_it", "60"], 9, False),
    "q1_e10_nev15": ("q1e10", 15, ["-gcge_ev_max_it", "60", "-gcge_ev_tol", "1e-10"], 3, False),
}


def _build(spec, golden_dir):
    if isinstance(spec, tuple):
        return _fixture(golden_dir, spec[0]), _fixture(golden_dir, spec[1])
    if spec == "fd24":
        return problems.fd_laplacian_2d(24), None
    return problems.q1_fem_3d(10)


@pytest.mark.parametrize("case", sorted(RUN_CASES))
def test_core_on_oracle_matches_reference_runs(case, golden_dir):
    """The GCG core on the oracle operations against the recorded runs of the
    reference solver from the same initial block: same iteration count, same
    converged count and, where the spectrum has no repeated eigenvalues, the
    same convergence count at every iteration and the same eigenvalues at
    every iteration (1e-9 relative)."""
    g = np.load(os.path.join(golden_dir, "ref_runs.npz"))
    spec, nev, flags, seed, simple = RUN_CASES[case]
    A, B = _build(spec, golden_dir)
    x0 = oracle.random_fill(seed, 0, 0, nev, A.nrows)
    out = oracle.solve(A, B, nev, flags, x0=x0, history_cap=200)
    iters, nconv = g[case + "/iters"]
    assert out["num_iter"] == iters and out["num_conv"] == nconv
    ref_ev = g[case + "/eval"]
    if nconv == nev:
        np.testing.assert_allclose(out["eval"], ref_ev, rtol=1e-10)
    if simple:
        np.testing.assert_array_equal(out["history"][:, :2], g[case + "/conv_history"])
        np.testing.assert_allclose(out["history"][:, 2:], g[case + "/eval_history"], rtol=1e-9)


@pytest.mark.parametrize("nev,flags,seed",
                         [(20, ["-gcge_block_size", "10", "-gcge_ev_max_it", "40"], 7),
                          (30, ["-gcge_block_size", "10", "-gcge_ev_max_it", "60"], 8)])
def test_dirichlet_boundary_matches_reference_runs(nev, flags, seed, golden_dir):
    """dirichlet_boundary = 1 on A_5/M_5 (the setting of the reference's
    canonical example, example/csr_solver.c:51) with a randomly drawn
    initial block: the core zeroes the Dirichlet rows through the oracle's
    SetDirichletBoundary (app/csr/gcge_app_csr.c:130-183), then runs as the
    reference binary did from that block with the rows zeroed by the
    fixture script's own reading of the rule (ref_runs.npz *_dirichlet).  A
    missing or wrong zeroing changes the first orthonormalisation and every
    iterate after it."""
    g = np.load(os.path.join(golden_dir, "ref_runs.npz"))
    A, B = _fixture(golden_dir, "A_5"), _fixture(golden_dir, "M_5")
    out = oracle.solve(A, B, nev, flags + ["-gcge_dirichlet_boundary", "1"], seed=seed,
                       history_cap=200)
    key = f"A5M5_nev{nev}_dirichlet"
    assert [out["num_iter"], out["num_conv"]] == list(g[key + "/iters"])
    np.testing.assert_array_equal(out["history"][:, :2], g[key + "/conv_history"])
    np.testing.assert_allclose(out["eval"], g[key + "/eval"], rtol=1e-10)
    # without the zeroing the same random block follows another trajectory
    plain = oracle.solve(A, B, nev, flags, seed=seed, history_cap=200)
    assert not np.allclose(plain["history"][:, 2:min(plain['history'].shape[1], 6)][:3],
                           g[key + "/eval_history"][:3, :4], rtol=1e-6)


def test_spurious_pairs_of_shipped_example(golden_dir):
    """The shipped example (A_5/M_5, nev = 9): block_size is clamped to 9, so
    after the first loop iteration the X block (max_dim_x = nev + 20 = 29
    columns, src/gcge_eigsol.c:561) is wider than the 27-dimensional subspace;
    coefficient columns 27 and 28 lie outside the projected matrix
    (src/gcge_xpw.c:341-371), two basis vectors become zero, the projected
    matrix gets two null directions, and two pairs at lambda ~ 0 never
    converge, so the prefix count stays 0 for all 30 iterations -- in the
    reference run recorded from the same block and in the core.  Beyond that
    the two runs are not comparable: with a zero Ritz vector the block CG
    starts with columns already converged and then reads scalars from
    positions of its work array that this call never wrote
    (src/gcge_cg.c:411 with :465), i.e. whatever earlier steps left there."""
    A, B = _fixture(golden_dir, "A_5"), _fixture(golden_dir, "M_5")
    g = np.load(os.path.join(golden_dir, "ref_runs.npz"))
    out = oracle.solve(A, B, 9, [], x0=oracle.random_fill(7, 0, 0, 9, A.nrows))
    np.testing.assert_array_equal([out["num_iter"], out["num_conv"]], g["A5M5_nev9/iters"])
    assert out["num_iter"] == 30 and out["num_conv"] == 0
    assert np.all(np.abs(out["eval"][:2]) < 1e-9)
    assert np.all(np.abs(g["A5M5_nev9/eval"][:2]) < 1e-9)
    assert not np.any(out["res"][:2] < 1e-8)  # never converged (large or undefined)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build absent")
def test_live_reference_fd2d():
    A = problems.fd_laplacian_2d(20)
    x0 = oracle.random_fill(5, 0, 0, 12, A.nrows)
    r = oracle.ref_solve(A, None, 12, [], x0=x0)
    o = oracle.solve(A, None, 12, [], x0=x0)
    assert r["num_iter"] == o["num_iter"] and o["num_conv"] == 12
    np.testing.assert_allclose(o["eval"], r["eval"], rtol=1e-10)
    np.testing.assert_allclose(np.sort(o["eval"]), problems.fd_laplacian_2d_eigs(20, 12), rtol=1e-10)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build absent")
def test_live_reference_small_block_spurious_zeros():
    """nev = 8 clamps block_size to 8 < 10: the X block outgrows the subspace
    and both the reference and the core return spurious lambda ~ 0 pairs (the
    reference's own behaviour, reproduced)."""
    A = problems.fd_laplacian_2d(20)
    x0 = oracle.random_fill(5, 0, 0, 8, A.nrows)
    r = oracle.ref_solve(A, None, 8, [], x0=x0)
    o = oracle.solve(A, None, 8, [], x0=x0)
    assert r["num_iter"] == o["num_iter"] == 30
    assert r["num_conv"] == o["num_conv"] == 0
    assert np.sum(np.abs(r["eval"]) < 1e-9) == np.sum(np.abs(o["eval"]) < 1e-9) == 4


# ---- block CG ------------------------------------------------------------------

BCG_FIXTURES = ["fd12_rate", "fd12_zero_col", "fd12_fixed_iters", "q1e6_shift_B", "q1e6_shift_I"]


def bcg_matrices(spec):
    """The matrices named in bcg_runs.npz."""
    if spec == "fd12":
        return problems.fd_laplacian_2d(12), None
    if spec == "q1e6":
        return problems.q1_fem_3d(6)
    if spec == "q1e6A":
        return problems.q1_fem_3d(6)[0], None
    raise ValueError(spec)


@pytest.mark.parametrize("case", BCG_FIXTURES)
def test_block_cg_matches_reference_fixture(case, golden_dir):
    """The restated block CG against the reference GCGE_BCG outputs, bit for
    bit, NaN columns included (a column that starts converged shifts the
    position-indexed scalars of the reference bookkeeping)."""
    g = np.load(os.path.join(golden_dir, "bcg_runs.npz"))
    max_it, rate, if_shift, shift = g[case + "/params"]
    A, B = bcg_matrices(str(g[case + "/spec"]))
    x, r = oracle.block_cg(A, B, g[case + "/rhs"], g[case + "/x0"], int(max_it), float(rate),
                           int(if_shift), float(shift))
    np.testing.assert_array_equal(x, g[case + "/x"])
    np.testing.assert_array_equal(r, g[case + "/r"])
    if case == "fd12_zero_col":
        assert np.isnan(x).any()


def test_block_cg_exact_on_small_spd():
    """CG on an n x n SPD system reaches the solution in n steps; the
    returned residual block is b - (A + shift B) x."""
    import scipy.sparse as sp

    rng = np.random.default_rng(4)
    n = 6
    Q = rng.standard_normal((n, n))
    S = Q @ Q.T + n * np.eye(n)
    A = problems.CSR.from_scipy(sp.csr_matrix(S))
    Bd = np.diag(rng.uniform(1, 2, n))
    B = problems.CSR.from_scipy(sp.csr_matrix(Bd))
    b = rng.standard_normal((n, 3))
    for shift_args, Mop in (((0, 0.0), S), ((1, 0.7), S + 0.7 * Bd)):
        x, r = oracle.block_cg(A, B, b, np.zeros((n, 3)), max_it=n + 1, rate=1e-30,
                               if_shift=shift_args[0], shift=shift_args[1])
        np.testing.assert_allclose(x, np.linalg.solve(Mop, b), rtol=1e-10, atol=1e-12)
        np.testing.assert_allclose(r, b - Mop @ x, atol=1e-10)
    # identity shift when B is absent
    x, _ = oracle.block_cg(A, None, b, np.zeros((n, 3)), max_it=n + 1, rate=1e-30, if_shift=1,
                           shift=1.5)
    np.testing.assert_allclose(x, np.linalg.solve(S + 1.5 * np.eye(n), b), rtol=1e-10)


def test_block_cg_stopping_rule():
    """A column stops once ||r|| < rate ||r0||; the first step from x0 = 0
    is the steepest-descent step.  (With several columns, those left after
    the first column stops read another column's scalars -- the reference
    bookkeeping -- and may stagnate; the fixtures above pin that case.)"""
    A = problems.fd_laplacian_2d(16)
    rng = np.random.default_rng(5)
    b = rng.standard_normal((A.nrows, 4))
    for c in range(4):
        bc = b[:, c:c + 1]
        x, r = oracle.block_cg(A, None, bc, np.zeros_like(bc), max_it=200, rate=1e-3)
        ratio = np.linalg.norm(r) / np.linalg.norm(bc)
        assert ratio < 1e-3
        np.testing.assert_allclose(r, bc - A.to_scipy() @ x, atol=1e-10)
        # the step before the last did not meet the rate
        _, r_prev = oracle.block_cg(A, None, bc, np.zeros_like(bc), max_it=_steps(A, bc), rate=1e-3)
        assert np.linalg.norm(r_prev) / np.linalg.norm(bc) >= 1e-3
    _, r1 = oracle.block_cg(A, None, b, np.zeros_like(b), max_it=2, rate=1e-3)
    Am = A.to_scipy()
    alpha = (b * b).sum(0) / (b * (Am @ b)).sum(0)
    np.testing.assert_allclose(r1, b - alpha * (Am @ b), rtol=1e-12, atol=1e-12)


def _steps(A, b):
    """max_it that allows one step fewer than the single-column CG needs to
    reach rate 1e-3 (plain textbook CG count)."""
    Am = A.to_scipy()
    x = np.zeros_like(b)
    r = b.copy()
    p = r.copy()
    rr = float((r * r).sum())
    r0 = np.sqrt(rr)
    steps = 0
    while np.sqrt(rr) >= 1e-3 * r0:
        w = Am @ p
        a = rr / float((p * w).sum())
        x += a * p
        r -= a * w
        rr_new = float((r * r).sum())
        p = r + rr_new / rr * p
        rr = rr_new
        steps += 1
    return steps  # max_it = steps allows steps - 1 iterations


@pytest.mark.skipif(not os.path.exists(oracle.REF_BCG_BIN), reason="reference build absent")
def test_live_reference_block_cg():
    """Restatement against the reference GCGE_BCG on a larger shifted
    generalized system with mixed convergence speeds."""
    K, M = problems.q1_fem_3d(9)
    rng = np.random.default_rng(6)
    b = rng.standard_normal((K.nrows, 6)) * np.logspace(0, -3, 6)
    x0 = 1e-3 * rng.standard_normal((K.nrows, 6))
    for args in ((20, 1e-2, 0, 0.0), (30, 1e-4, 1, 12.0)):
        x, r = oracle.block_cg(K, M, b, x0, *args)
        xr, rr = oracle.ref_block_cg(K, M, b, x0, *args)
        np.testing.assert_array_equal(x, xr)
        np.testing.assert_array_equal(r, rr)


# ---- soft-locking stall and the full Rayleigh-Ritz option -------------------------

def test_full_rr_converges_where_soft_locking_stalls():
    """Q1 14^3, nev = 30, RB 1e-8 from the seed-2 block: zeroing the coupling
    of the converged prefix in the Rayleigh-Ritz step (the reference's soft
    locking) caps later vectors of a partly converged cluster above the
    tolerance; with -gcge_full_rr 1 every pair converges, to the closed-form
    spectrum."""
    E, nev = 14, 30
    K, M = problems.q1_fem_3d(E)
    x0 = oracle.random_fill(2, 0, 0, nev, K.nrows)
    flags = ["-gcge_conv_type", "RB", "-gcge_ev_tol", "1e-8", "-gcge_ev_max_it", "80"]
    full = oracle.solve(K, M, nev, flags + ["-gcge_full_rr", "1"], x0=x0)
    assert full["num_conv"] == nev and full["num_iter"] < 60
    np.testing.assert_allclose(np.sort(full["eval"]), problems.q1_eigs(E, nev), rtol=1e-10)
    X, lam = full["evec"], full["eval"]
    R = K.to_scipy() @ X - (M.to_scipy() @ X) * lam
    rb = np.linalg.norm(R, axis=0) / (np.abs(lam) * np.linalg.norm(M.to_scipy() @ X, axis=0))
    assert np.all(rb <= 1.05e-8)
    G = X.T @ (M.to_scipy() @ X)
    assert np.abs(G - np.eye(nev)).max() < 1e-10


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build absent")
def test_live_reference_soft_locking_stall():
    """The reference itself stalls on the same start at a strict tolerance
    of its own criterion: 150 iterations, fewer than nev pairs."""
    E, nev = 14, 30
    K, M = problems.q1_fem_3d(E)
    x0 = oracle.random_fill(2, 0, 0, nev, K.nrows)
    r = oracle.ref_solve(K, M, nev, ["-gcge_ev_tol", "3e-12", "-gcge_ev_max_it", "150"], x0=x0)
    assert r["num_iter"] == 150 and r["num_conv"] < nev


def test_svqb_w_orth_same_spectrum_and_iteration_count():
    """-gcge_w_orth_type svqb (SVQB passes inside W after the same
    projection against [X, P]) spans the same space as the recursive "multi"
    orthonormalisation, so the iteration converges to the same pairs in about
    as many steps; the eigenvectors come out B-orthonormal."""
    E, nev = 14, 30
    K, M = problems.q1_fem_3d(E)
    x0 = oracle.random_fill(3, 0, 0, nev, K.nrows)
    base = ["-gcge_conv_type", "RB", "-gcge_ev_tol", "1e-8", "-gcge_ev_max_it", "100",
            "-gcge_full_rr", "1"]
    multi = oracle.solve(K, M, nev, base, x0=x0)
    svqb = oracle.solve(K, M, nev, base + ["-gcge_w_orth_type", "svqb"], x0=x0)
    assert multi["num_conv"] == nev and svqb["num_conv"] == nev
    assert abs(multi["num_iter"] - svqb["num_iter"]) <= 3
    np.testing.assert_allclose(np.sort(svqb["eval"]), problems.q1_eigs(E, nev), rtol=1e-10)
    X = svqb["evec"]
    assert np.abs(X.T @ (M.to_scipy() @ X) - np.eye(nev)).max() < 1e-10


def test_svqb_one_pass_matches_two_passes():
    """-gcge_svqb_one_pass_cond (default 1e4) skips the second SVQB pass when
    the first kept every column of a well-conditioned block; 0 always runs
    two.  Both reach the closed-form spectrum in about as many iterations
    with B-orthonormal eigenvectors, and a threshold of 1e30 (one pass
    whenever no column is dropped) still gives B-orthonormal vectors on this
    pencil."""
    E, nev = 14, 30
    K, M = problems.q1_fem_3d(E)
    Ms = M.to_scipy()
    x0 = oracle.random_fill(5, 0, 0, nev, K.nrows)
    base = ["-gcge_conv_type", "RB", "-gcge_ev_tol", "1e-8", "-gcge_ev_max_it", "100",
            "-gcge_full_rr", "1", "-gcge_w_orth_type", "svqb"]
    runs = {c: oracle.solve(K, M, nev, base + ["-gcge_svqb_one_pass_cond", c], x0=x0)
            for c in ("0", "1e4", "1e30")}
    default = oracle.solve(K, M, nev, base, x0=x0)
    np.testing.assert_array_equal(default["eval"], runs["1e4"]["eval"])
    assert default["num_iter"] == runs["1e4"]["num_iter"]
    want = problems.q1_eigs(E, nev)
    for c, out in runs.items():
        assert out["num_conv"] == nev, c
        assert abs(out["num_iter"] - runs["0"]["num_iter"]) <= 3, c
        np.testing.assert_allclose(np.sort(out["eval"]), want, rtol=1e-10)
        X = out["evec"]
        assert np.abs(X.T @ (Ms @ X) - np.eye(nev)).max() < 1e-10, c


def test_two_projection_passes_with_svqb():
    """-gcge_max_reorth_time 2 (the benchmark's setting with svqb): the
    projection of W against [X, P] stops after two classical Gram-Schmidt
    passes, where the reference default runs a third whose coefficients are
    of relative size 1e-16 and below.  Same closed-form spectrum, about as
    many iterations, B-orthonormal eigenvectors; on a lognormal pencil the
    eigenvalues agree with the three-pass solve to 1e-12.  (W as the
    reference forms it: -gcge_svqb_dgks_eta 0.)"""
    E, nev = 14, 30
    K, M = problems.q1_fem_3d(E)
    x0 = oracle.random_fill(4, 0, 0, nev, K.nrows)
    base = ["-gcge_conv_type", "RB", "-gcge_ev_tol", "1e-8", "-gcge_ev_max_it", "100",
            "-gcge_full_rr", "1", "-gcge_w_orth_type", "svqb", "-gcge_svqb_dgks_eta", "0"]
    three = oracle.solve(K, M, nev, base, x0=x0)
    two = oracle.solve(K, M, nev, base + ["-gcge_max_reorth_time", "2"], x0=x0)
    assert three["num_conv"] == nev and two["num_conv"] == nev
    assert abs(three["num_iter"] - two["num_iter"]) <= 2
    np.testing.assert_allclose(np.sort(two["eval"]), problems.q1_eigs(E, nev), rtol=1e-10)
    X = two["evec"]
    assert np.abs(X.T @ (M.to_scipy() @ X) - np.eye(nev)).max() < 1e-10
    Kl, Ml = problems.q1_fem_3d(12, problems.lognormal_kappa(12, 3))
    x0 = oracle.random_fill(6, 0, 0, 20, Kl.nrows)
    a = oracle.solve(Kl, Ml, 20, base, x0=x0)
    b = oracle.solve(Kl, Ml, 20, base + ["-gcge_max_reorth_time", "2"], x0=x0)
    assert a["num_conv"] == 20 and b["num_conv"] == 20
    np.testing.assert_allclose(b["eval"], a["eval"], rtol=1e-12)


def test_svqb_w_as_correction_with_one_projection_pass():
    """-gcge_svqb_dgks_eta (default 1/sqrt 2, svqb only): W leaves the inner
    solve as the correction to its starting vector and the projection makes
    a second pass only for columns that kept less than eta of their B-norm
    in the first.  Against eta = 0 (W as the reference forms it, every
    pass): the same closed-form spectrum in about as many iterations, and
    on a lognormal pencil eigenvalues within 1e-12 and B-orthonormal
    eigenvectors; eta = 1 forces the second pass on every column."""
    E, nev = 14, 30
    K, M = problems.q1_fem_3d(E)
    x0 = oracle.random_fill(4, 0, 0, nev, K.nrows)
    base = ["-gcge_conv_type", "RB", "-gcge_ev_tol", "1e-8", "-gcge_ev_max_it", "100",
            "-gcge_full_rr", "1", "-gcge_w_orth_type", "svqb", "-gcge_max_reorth_time", "2"]
    runs = {eta: oracle.solve(K, M, nev, base + ["-gcge_svqb_dgks_eta", eta], x0=x0)
            for eta in ("0", "0.7071067811865476", "1")}
    default = oracle.solve(K, M, nev, base, x0=x0)
    np.testing.assert_array_equal(default["eval"], runs["0.7071067811865476"]["eval"])
    for eta, out in runs.items():
        assert out["num_conv"] == nev, eta
        assert abs(out["num_iter"] - runs["0"]["num_iter"]) <= 2, eta
        np.testing.assert_allclose(np.sort(out["eval"]), problems.q1_eigs(E, nev), rtol=1e-10)
        X = out["evec"]
        assert np.abs(X.T @ (M.to_scipy() @ X) - np.eye(nev)).max() < 1e-10, eta
    Kl, Ml = problems.q1_fem_3d(12, problems.lognormal_kappa(12, 3))
    x0 = oracle.random_fill(6, 0, 0, 20, Kl.nrows)
    a = oracle.solve(Kl, Ml, 20, base + ["-gcge_svqb_dgks_eta", "0"], x0=x0)
    b = oracle.solve(Kl, Ml, 20, base, x0=x0)
    assert a["num_conv"] == 20 and b["num_conv"] == 20
    np.testing.assert_allclose(b["eval"], a["eval"], rtol=1e-12)
    X = b["evec"]
    assert np.abs(X.T @ (Ml.to_scipy() @ X) - np.eye(20)).max() < 1e-10
