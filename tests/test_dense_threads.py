"""The host symmetric eigensolver (GCGE_Dense_SymEig, the Rayleigh-Ritz
eigensolve below the device threshold) gives the same bits with any number of
worker threads, and the right eigenpairs (CPU)."""
import os
import subprocess
import sys

import numpy as np
import pytest

from gcge_old_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# matrices: random of several orders (around the thread threshold, the
# parallel-reduction cut-off of 48 trailing rows and the Rayleigh-Ritz
# orders of the benchmark pencils), clustered spectra, a tridiagonal
# matrix, a diagonal one and zero
_RUNNER = r"""
import ctypes as C, sys
import numpy as np
lib = C.CDLL(sys.argv[1])
rng = np.random.default_rng(11)
mats = []
for n in (1, 2, 3, 47, 48, 49, 64, 90, 131, 257):
    a = rng.standard_normal((n, n))
    mats.append(a + a.T)
for n in (70, 200):
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.repeat(rng.standard_normal(n // 5 + 1), 5)[:n]
    mats.append((q * lam) @ q.T)
n = 130
mats.append(np.diag(rng.standard_normal(n)) + np.diag(np.ones(n - 1), 1) + np.diag(np.ones(n - 1), -1))
mats.append(np.diag(rng.standard_normal(100)))
mats.append(np.zeros((80, 80)))
out = {}
P = C.c_void_p
for i, A in enumerate(mats):
    n = A.shape[0]
    a = np.asfortranarray(A.copy())
    ev = np.zeros(n)
    V = np.zeros((n, n), order="F")
    W = np.zeros(n * n + 5 * n)
    lib.GCGE_Dense_SymEig(P(a.ctypes.data), n, n, P(ev.ctypes.data), P(V.ctypes.data), n,
                          P(W.ctypes.data))
    out[f"a{i}"], out[f"e{i}"], out[f"v{i}"] = A, ev, V
np.savez(sys.argv[2], **out)
"""


def _run(tmp_path, threads, min_n):
    env = dict(os.environ, GCGE_DENSE_THREADS=str(threads), GCGE_DENSE_MIN_N=str(min_n))
    out = tmp_path / f"eig_t{threads}_m{min_n}.npz"
    subprocess.run([sys.executable, "-c", _RUNNER, _native.lib_path("libgcge_amd.so"), str(out)],
                   check=True, env=env, timeout=300)
    return np.load(out)


def test_thread_count_does_not_change_a_bit(tmp_path):
    base = _run(tmp_path, 1, 2)
    for threads in (2, 3, 8):
        other = _run(tmp_path, threads, 2)
        for key in base.files:
            assert np.array_equal(base[key], other[key]), (threads, key)


def test_threaded_eigenpairs_are_correct(tmp_path):
    res = _run(tmp_path, 4, 2)
    count = len([k for k in res.files if k.startswith("a")])
    for i in range(count):
        A, ev, V = res[f"a{i}"], res[f"e{i}"], res[f"v{i}"]
        n = A.shape[0]
        scale = max(1.0, np.abs(A).max()) * n
        assert np.all(np.diff(ev) >= 0)
        assert np.max(np.abs(ev - np.linalg.eigvalsh(A))) < 1e-13 * scale
        assert np.max(np.abs(V.T @ V - np.eye(n))) < 1e-13 * n
        assert np.max(np.abs(A @ V - V * ev)) < 1e-13 * scale


@pytest.mark.parametrize("threads", [1, 8])
def test_pool_survives_repeated_and_mixed_orders(tmp_path, threads):
    """Orders below and above the threshold in turn, in one process (the
    pool and its sweep log are reused and grown)."""
    res = _run(tmp_path, threads, 60)
    ref = _run(tmp_path, 1, 100000)
    for key in ref.files:
        assert np.array_equal(ref[key], res[key]), key


_PROD_RUNNER = r"""
import ctypes as C, sys
import numpy as np
lib = C.CDLL(sys.argv[1])
P = C.c_void_p
ci = lambda v: C.byref(C.c_int(v))
cd = lambda v: C.byref(C.c_double(v))
rng = np.random.default_rng(3)
out = {}
i = 0
for (M, N, K) in [(300, 100, 300), (100, 300, 100), (7, 5, 3), (1, 9, 4), (200, 2, 50), (301, 99, 17)]:
    for ta in "NT":
        for tb in "NT":
            for al, be in [(1.0, 0.0), (-0.5, 2.0), (0.0, 3.0)]:
                A = np.asfortranarray(rng.standard_normal((M, K) if ta == "N" else (K, M)))
                B = np.asfortranarray(rng.standard_normal((K, N) if tb == "N" else (N, K)))
                C0 = np.asfortranarray(rng.standard_normal((M + 2, N)))
                Cm = C0.copy(order="F")
                lib.GCGE_Default_DenseMatDotDenseMat(C.c_char_p(ta.encode()), C.c_char_p(tb.encode()),
                    ci(M), ci(N), ci(K), cd(al), P(A.ctypes.data), ci(A.shape[0]), P(B.ctypes.data),
                    ci(B.shape[0]), cd(be), P(Cm.ctypes.data), ci(M + 2))
                want = al * ((A if ta == "N" else A.T) @ (B if tb == "N" else B.T)) + be * C0[:M]
                out[f"g{i}"], out[f"w{i}"], out[f"r{i}"] = Cm, want, C0
                i += 1
for (M, N) in [(300, 100), (100, 300), (5, 7)]:
    for side in "LR":
        for uplo in "UL":
            S = M if side == "L" else N
            A = np.asfortranarray(rng.standard_normal((S, S)))
            B = np.asfortranarray(rng.standard_normal((M, N)))
            C0 = np.asfortranarray(rng.standard_normal((M, N)))
            Cm = C0.copy(order="F")
            lib.GCGE_Default_DenseSymMatDotDenseMat(C.c_char_p(side.encode()), C.c_char_p(uplo.encode()),
                ci(M), ci(N), cd(0.7), P(A.ctypes.data), ci(S), P(B.ctypes.data), ci(M), cd(0.25),
                P(Cm.ctypes.data), ci(M))
            Sy = np.triu(A) + np.triu(A, 1).T if uplo == "U" else np.tril(A) + np.tril(A, -1).T
            want = 0.7 * (Sy @ B if side == "L" else B @ Sy) + 0.25 * C0
            out[f"g{i}"], out[f"w{i}"], out[f"r{i}"] = Cm, want, C0
            i += 1
np.savez(sys.argv[2], **out)
"""


def _run_prod(tmp_path, threads):
    env = dict(os.environ, GCGE_DENSE_THREADS=str(threads), GCGE_DENSE_MIN_FLOPS="1")
    out = tmp_path / f"prod_t{threads}.npz"
    subprocess.run([sys.executable, "-c", _PROD_RUNNER, _native.lib_path("libgcge_amd.so"), str(out)],
                   check=True, env=env, timeout=300)
    return np.load(out)


def test_dense_products_same_bits_on_threads(tmp_path):
    """GCGE_Default_DenseMatDotDenseMat (every transpose case, beta 0 / 1 /
    other, leading dimension above M) and DenseSymMatDotDenseMat (both sides
    and triangles): the threaded form deals out whole columns, so every
    thread count gives the single-threaded bits, and the results are the
    products numpy forms."""
    base = _run_prod(tmp_path, 1)
    for threads in (3, 8):
        other = _run_prod(tmp_path, threads)
        for key in base.files:
            if key.startswith("g"):
                assert np.array_equal(base[key], other[key]), (threads, key)
    for key in base.files:
        if key.startswith("g"):
            i = key[1:]
            got, want = base[key], base["w" + i]
            M = want.shape[0]
            assert np.allclose(got[:M], want, rtol=1e-12, atol=1e-12 * max(1.0, np.abs(want).max()))
            assert np.array_equal(got[M:], base["r" + i][M:])  # rows beyond M untouched
