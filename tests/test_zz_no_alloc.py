"""GCGE_HIP_GramNoAlloc, the Gram entry point that allocates nothing (GPU).

Two calls in a row through the C ABI, the second with a smaller input,
each against GCGE_HIP_Gram on the same blocks: bit for bit while the partial
sums fit the fixed device array (same kernels, same row ranges), to 1e-13
of the largest entry when the row ranges are capped; an oversized m * k is
refused with an error and no device work.
"""
import numpy as np
import pytest

from gcge_old_amd import DeviceMultiVec, _native

pytestmark = pytest.mark.gpu


def _both(n, m, k, seed):
    lib = _native.hip()
    rng = np.random.default_rng(seed)
    V = DeviceMultiVec.from_host(rng.standard_normal((n, m)))
    W = DeviceMultiVec.from_host(rng.standard_normal((n, k)))
    g_ref = np.zeros((m, k), order="F")
    lib.GCGE_HIP_Gram(V.handle, 0, m, W.handle, 0, k, _native.as_dptr(g_ref), m)
    g = np.full((m + 3, k), np.nan, order="F")  # leading dimension m + 3
    rc = lib.GCGE_HIP_GramNoAlloc(V.handle, 0, m, W.handle, 0, k, _native.as_dptr(g), m + 3)
    assert rc == 0, lib.GCGE_HIP_LastError().decode()
    assert np.all(np.isnan(g[m:]))  # rows beyond m untouched
    return g[:m], g_ref


@pytest.mark.parametrize("first,second", [((100000, 40, 30), (5000, 7, 5)),
                                          ((250047, 90, 10), (3000, 17, 33))])
def test_no_alloc_gram_equals_gram_two_calls(first, second):
    """the second call smaller than the first: the array keeps the first
    call's contents, which the second must not read"""
    for i, (n, m, k) in enumerate((first, second)):
        print(f"GramNoAlloc n={n} m={m} k={k}", flush=True)
        g, g_ref = _both(n, m, k, 10 + i)
        assert np.array_equal(g.view(np.int64), g_ref.view(np.int64)), (n, m, k)


def test_no_alloc_gram_capped_ranges_then_smaller():
    """300 x 100 over 2048383 rows needs more partial sums than the array
    holds: fewer row ranges, the same result to rounding; then a small call"""
    g, g_ref = _both(2048383, 300, 100, 3)
    assert np.abs(g - g_ref).max() <= 1e-13 * np.abs(g_ref).max()
    g, g_ref = _both(1000, 3, 2, 4)
    assert np.array_equal(g.view(np.int64), g_ref.view(np.int64))


def test_no_alloc_gram_refuses_oversized_output():
    lib = _native.hip()
    V = DeviceMultiVec(64, 800)
    G = np.zeros((800, 800), order="F")
    rc = lib.GCGE_HIP_GramNoAlloc(V.handle, 0, 800, V.handle, 0, 800, _native.as_dptr(G), 800)
    assert rc == 1
    assert "exceeds" in lib.GCGE_HIP_LastError().decode()
    assert not G.any()
