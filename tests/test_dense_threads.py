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
