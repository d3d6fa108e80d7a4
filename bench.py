"""Benchmark: seconds to nev eigenpairs of the Q1 FEM pencil on one MI355X.

Default workload (one "step" = one complete GCG solve): trilinear Q1
hexahedral elements on the uniform 128^3 mesh of the unit cube, Dirichlet
nodes eliminated (n = 127^3 = 2048383), stiffness K with a piecewise-constant
coefficient log kappa_e ~ N(0, 1) i.i.d. from the seed and consistent mass M,
nev = 100 smallest eigenpairs of K x = lambda M x (block size 200 requested,
which the reference's rule clamps to nev), converged to
||Kx - lambda Mx|| / (|lambda| ||Mx||) <= 1e-8.  K and M are assembled on the
device from kappa before timing; the initial block is drawn on the device
inside each solve from the seed.

value is the mixed-precision form: the inner block CG (directions, products,
residuals and its correction of the Ritz vectors) in float32, float64 for the
B-Gram blocks, their orthonormalisation, the Rayleigh-Ritz matrices and dense
eigensolve, and the residual check.  The plain float64 solve of the same
seeds is timed beside it (fp64), with the largest relative eigenvalue
difference, the largest principal angle between the two converged subspaces
and the largest final residual of each.

--workload q1-64 times the 64^3 kappa = 1 pencil (nev 50) with the float64
solve as value (eigenvalues against the closed form), mixed beside it.

The solver runs with -gcge_full_rr 1: the converged prefix stays in the
Rayleigh-Ritz step with its computed coupling.  The reference's soft locking
zeroes that coupling, which at this tolerance stalls some starts short of nev
pairs (the reference build does the same); --reference-locking restores it.
The W block is B-orthonormalised with -gcge_w_orth_type svqb (the reference's
projection against [X, P], then one or two SVQB passes instead of the
recursive single-column leaves: a few host round trips instead of about 40,
same span); W leaves the inner solve as the correction to its starting
vector, and the projection makes a second pass only where a column kept
less than 1/sqrt(2) of its B-norm in the first (-gcge_svqb_dgks_eta, the
Daniel-Gragg-Kaufman-Stewart criterion), at most two passes
(-gcge_max_reorth_time 2; the reference's 3rd pass removes components of
relative size 1e-16 and below); --reference-orth restores "multi", W as the
reference forms it and three passes.

The JSON line carries:
  value          mean seconds per solve over the timed steps (lower is better)
  roofline       the block-CG products (the dominant kernel; float32 in the
                 mixed form): algorithmic bytes over their HIP-event time
                 summed over the timed region (events recorded by the
                 dispatch packets, hipExtLaunchKernelGGL; nothing else is
                 bracketed during the timed steps), against 8 TB/s; traffic
                 from the committed rocprofv3 counter summary in profiles/
  kernels_one_solve  per-kernel event time of one extra, untimed solve with
                 every launch bracketed
  cpu_baseline   the reference's CSR path for the same work on one host
                 core: the solve's operations counted as the single-vector
                 CSR calls the reference's column loops make (products,
                 inner products, axpbys, dense eigensolves), each priced by
                 the oracle port's own call timed at the workload's size
                 (the reference's vendored LAPACK for the dense eigensolve)
  fp64 / mixed   the other precision on the same seeds
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "s to nev eigpairs @resid 1e-8: FP32-inner vs FP64 on GPU; SpMM HBM GB/s"
HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=2)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--workload", choices=("q1-128-lognormal", "q1-64", "q1-64-adjoint",
                                          "q1-192-permuted"),
                   default="q1-128-lognormal",
                   help="q1-128-lognormal: Q1 128^3, lognormal kappa, nev 100, block 200 "
                        "requested, value = mixed-precision solve (default); "
                        "q1-64: Q1 64^3, kappa = 1, nev 50, value = float64 solve; "
                        "q1-64-adjoint: Q1 64^3 lognormal kappa, nev 20, gradient of "
                        "sum w_k lambda_k + ||P x_1||^2 with respect to all kappa_e, "
                        "checked by central differences on --fd elements at each of --fd-steps; "
                        "q1-192-permuted: Q1 192^3 lognormal kappa, rows and columns "
                        "symmetrically permuted by a seeded random permutation, nev 200, block "
                        "400 requested, a bounded number of GCG iterations (--max-it, default "
                        "3 here) in float64 and mixed precision, then the eigenvalue adjoint on "
                        "all elements")
    p.add_argument("--fd", type=int, default=64, help="elements checked by central differences "
                                                      "(q1-64-adjoint)")
    p.add_argument("--fd-steps", type=float, nargs="+", default=[1e-6, 1e-4, 1e-3, 1e-2],
                   help="relative steps of the central differences (q1-64-adjoint)")
    p.add_argument("--reorder", choices=("none", "rcm"), default="none",
                   help="q1-192-permuted: reorder the permuted pencil by reverse Cuthill-McKee "
                        "(computed from the matrix graph, a preprocessing step) before the solve")
    p.add_argument("--assembly", choices=("device", "host"), default="device",
                   help="assemble K and M on the device from kappa (default), or on the host "
                        "(numpy) and copy the CSR arrays")
    p.add_argument("--mesh", type=int, default=None, help="elements per axis (overrides the workload)")
    p.add_argument("--nev", type=int, default=None)
    p.add_argument("--kappa-seed", type=int, default=1)
    p.add_argument("--tol", type=float, default=1e-8)
    p.add_argument("--max-it", type=int, default=None,
                   help="GCG iteration cap (a bounded run for profiling; solves stopped by it "
                        "report all_converged false)")
    p.add_argument("--other-steps", type=int, default=-1,
                   help="solves of the other precision (default: min(--steps, 2); 0: skip)")
    p.add_argument("--cpu", choices=("model", "none"), default="model",
                   help="CPU baseline: op-count model priced by the port's timed calls")
    p.add_argument("--cpu-reference-iters", type=int, default=0,
                   help="also time the reference binary (oracle/_ref) on 1 and 1 + this many "
                        "GCG iterations of the same pencil and scale the difference to the "
                        "device iteration count (reference mode: its soft locking, multi "
                        "orthonormalisation and R residual)")
    p.add_argument("--reference-locking", action="store_true",
                   help="soft locking as in the reference (no -gcge_full_rr)")
    p.add_argument("--reference-orth", action="store_true",
                   help="recursive \"multi\" W orthonormalisation as in the reference "
                        "(default here: -gcge_w_orth_type svqb)")
    p.add_argument("--max-reorth-time", type=int, default=None,
                   help="-gcge_max_reorth_time: projection passes of W against [X, P] "
                        "(default 2 with the svqb W orthonormalisation, the reference's 3 "
                        "with --reference-orth)")
    p.add_argument("--no-profile", action="store_true",
                   help="time without per-kernel events (no roofline numbers)")
    p.add_argument("--no-breakdown", action="store_true",
                   help="skip the extra untimed solve with every launch bracketed")
    p.add_argument("--mfma-probe", action="store_true",
                   help="isolated Gram / linear-combination launches for matrix-core counters")
    p.add_argument("--pmc-probe", action="store_true",
                   help="only the workload's product launches (the block-CG float32 product, "
                        "float64 products) and plain copies of known size, for rocprofv3 --pmc "
                        "counter passes; prints one line")
    p.add_argument("--kernels", action="store_true",
                   help="also time SpMM, Gram, linear combination and a plain copy in isolation "
                        "(printed on stderr; the copy is the calibration for HBM counters)")
    return p.parse_args()


def traffic_file(E, lognormal):
    return os.path.join(ROOT, "profiles", f"spmm_q1_{E}{'_lognormal' if lognormal else ''}_hbm_traffic.md")


def read_traffic(E, lognormal, key):
    """HBM bytes per launch of the product `key` from the committed counter
    summary: the line '<key>: <bytes>'."""
    try:
        with open(traffic_file(E, lognormal)) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def pmc_probe(dK, k_cg):
    """Isolated launches for the counter passes, three of each: the block-CG
    float32 product on K with k_cg columns (through the block CG with a
    fixed step count: one direction update, product and x/r update per
    step), the float64 products with 10 (the float64 block-CG shape at 64^3),
    k_cg and 2 k_cg columns, and a plain copy of n x 64 doubles (the
    calibration)."""
    import gcge_old_amd as gcge
    from gcge_old_amd import _native

    lib = _native.hip()
    n = dK.nrows
    X = gcge.DeviceMultiVec(n, 2 * k_cg)
    Y = gcge.DeviceMultiVec(n, 2 * k_cg)
    lib.GCGE_HIP_RandomFill(X.handle, 0, 2 * k_cg)
    lib.GCGE_HIP_RandomFill(Y.handle, 0, 2 * k_cg)
    # float32 block CG: 4 iterations = 3 products of k_cg columns (rate 0:
    # no column stops)
    dK.set_precision(True)
    lib.GCGE_HIP_SetLinearSolverParams(4, 0.0, 0, 0.0, None)
    lib.GCGE_HIP_MultiLinearSolve(dK.handle, Y.handle, 0, X.handle, 0, k_cg)
    lib.GCGE_HIP_SetLinearSolverParams(20, 1e-2, 0, 0.0, None)
    dK.set_precision(False)
    widths = sorted({10, k_cg, 2 * k_cg})
    for k in widths:
        for _ in range(3):
            lib.GCGE_HIP_SpMM(dK.handle, X.handle, 0, Y.handle, 0, k)
    for _ in range(3):
        lib.GCGE_HIP_Axpby(1.0, X.handle, 0, 0.0, Y.handle, 0, 64)
    lib.GCGE_HIP_Synchronize()
    fixed64 = dK.nnz * 12 + 4 * (n + 1)
    fixed32 = dK.nnz * 8 + 4 * (n + 1)
    print(json.dumps({"pmc_probe": True, "n": n, "nnz": dK.nnz,
                      "spmm32_cg_algorithmic_bytes": fixed32 + 8 * k_cg * n,
                      "spmm64_algorithmic_bytes": {str(k): fixed64 + 16 * k * n for k in widths},
                      "copy_bytes_read": 8 * n * 64, "copy_bytes_written": 8 * n * 64}),
          flush=True)


def mfma_probe():
    """Isolated Gram and linear-combination launches at the solver's shapes
    (n, m, k) = (2048383, 300, 100) (128^3 projection), (250047, 60, 10) and
    (250047, 70, 50) (64^3), three of each, for the matrix-core counter
    passes; prints their event times."""
    import ctypes as C

    import gcge_old_amd as gcge
    from gcge_old_amd import _native

    lib = _native.hip()
    for n, m, k in ((2048383, 300, 100), (250047, 60, 10), (250047, 70, 50)):
        V, W = gcge.DeviceMultiVec(n, m), gcge.DeviceMultiVec(n, k)
        lib.GCGE_HIP_RandomFill(V.handle, 0, m)
        lib.GCGE_HIP_RandomFill(W.handle, 0, k)
        G = np.zeros((m, k), order="F")
        Cm = np.asfortranarray(np.random.default_rng(1).standard_normal((m, k)))
        lib.GCGE_HIP_ProfileReset()
        lib.GCGE_HIP_ProfileEnable(1)
        for _ in range(3):
            lib.GCGE_HIP_Gram(V.handle, 0, m, W.handle, 0, k, _native.as_dptr(G), m)
        for _ in range(3):
            lib.GCGE_HIP_LinearComb(V.handle, 0, m, _native.as_dptr(Cm), m, 0.0, W.handle, 0, k)
        lib.GCGE_HIP_ProfileEnable(0)
        res = {"n": n, "m": m, "k": k, "gflop": round(2.0 * n * m * k / 1e9, 2)}
        for name in ("gram", "lincomb"):
            ms, cnt, by = C.c_double(), C.c_int(), C.c_double()
            lib.GCGE_HIP_KernelStats(name.encode(), C.byref(ms), C.byref(cnt), C.byref(by))
            us = ms.value * 1e3 / max(cnt.value, 1)
            res[name + "_us"] = round(us, 1)
            res[name + "_tflops"] = round(2.0 * n * m * k / (us * 1e-6) / 1e12, 1) if us else None
        print(json.dumps(res), flush=True)
        del V, W


def kernel_sweep(dK, dM):
    """Isolated kernel timings on the workload's matrices (stderr)."""
    import ctypes as C

    import gcge_old_amd as gcge
    from gcge_old_amd import _native

    lib = _native.hip()
    n = dM.nrows
    X = gcge.DeviceMultiVec(n, 300)
    Y = gcge.DeviceMultiVec(n, 200)
    lib.GCGE_HIP_RandomFill(X.handle, 0, 300)

    def timed(name, fn, reps=10, extra=()):
        """us per launch and GB/s of kernel `name`; `extra` kernels of the
        same calls (a second reduction pass) are added to the time"""
        lib.GCGE_HIP_ProfileReset()
        lib.GCGE_HIP_ProfileEnable(1)
        for _ in range(reps):
            fn()
        lib.GCGE_HIP_ProfileEnable(0)
        ms, cnt, byt = C.c_double(0.0), C.c_int(0), C.c_double(0.0)
        lib.GCGE_HIP_KernelStats(name.encode(), C.byref(ms), C.byref(cnt), C.byref(byt))
        total_ms = ms.value
        for e in extra:
            ms2, cnt2 = C.c_double(0.0), C.c_int(0)
            lib.GCGE_HIP_KernelStats(e.encode(), C.byref(ms2), C.byref(cnt2), None)
            total_ms += ms2.value
        us = total_ms * 1e3 / max(cnt.value, 1)
        gbs = byt.value / (total_ms * 1e-3) / 1e9 if total_ms > 0 else 0.0
        return us, gbs

    for mat, label in ((dK, "K"), (dM, "M")):
        for k in (1, 10, 50, 100, 200):
            us, gbs = timed("spmm", lambda: lib.GCGE_HIP_SpMM(mat.handle, X.handle, 0, Y.handle, 0, k))
            print(f"# kernel spmm {label} k={k}: {us:.1f} us/launch, {gbs:.0f} GB/s", file=sys.stderr)
    G = np.zeros((300, 200), order="F")
    for m, k in ((70, 10), (60, 60), (200, 100), (300, 100), (1, 1)):
        us, gbs = timed("gram", lambda: lib.GCGE_HIP_Gram(X.handle, 0, m, Y.handle, 0, k,
                                                          _native.as_dptr(G), 300),
                        extra=("gram_reduce",))
        print(f"# kernel gram m={m} k={k} (with its reduction pass): {us:.1f} us/call, {gbs:.0f} GB/s, "
              f"{2.0 * n * m * k / (us * 1e-6) / 1e12:.1f} TFLOP/s", file=sys.stderr)
    Cm = np.asfortranarray(np.random.default_rng(0).standard_normal((300, 200)))
    for m, k in ((90, 50), (200, 100), (300, 100), (70, 10)):
        us, gbs = timed("lincomb", lambda: lib.GCGE_HIP_LinearComb(X.handle, 0, m, _native.as_dptr(Cm),
                                                                   300, 0.0, Y.handle, 0, k))
        print(f"# kernel lincomb m={m} k={k}: {us:.1f} us/launch, {gbs:.0f} GB/s, "
              f"{2.0 * n * m * k / (us * 1e-6) / 1e12:.1f} TFLOP/s", file=sys.stderr)
    # plain copy of 64 columns (n x 64 doubles read and written): the counter
    # calibration dispatch
    us, _ = timed("axpby", lambda: lib.GCGE_HIP_Axpby(1.0, X.handle, 0, 0.0, Y.handle, 0, 64))
    copy_bytes = 2 * 8.0 * n * 64
    print(f"# kernel copy {copy_bytes / 2**20:.1f} MiB moved: {us:.1f} us/launch, "
          f"{copy_bytes / (us * 1e-6) / 1e9:.0f} GB/s", file=sys.stderr)


KERNELS = ("spmm", "spmm32", "gram", "gram_reduce", "lincomb", "axpby", "axpby_list", "swap",
           "coldots", "bcg_p", "bcg_xr", "axpy_cols", "residual", "random", "convert")
WORK_NAMES = ("spmv", "spmv_nnz", "spmv_rows", "dot", "axpby", "dense_eig", "dense_eig_n3",
              "vec_len")


def kernel_stats(lib):
    """Per-kernel HIP-event totals since the last profile reset."""
    import ctypes as C

    kernels = {}
    for name in KERNELS:
        ms, cnt, byt = C.c_double(0.0), C.c_int(0), C.c_double(0.0)
        lib.GCGE_HIP_KernelStats(name.encode(), C.byref(ms), C.byref(cnt), C.byref(byt))
        if cnt.value:
            kernels[name] = {"ms": ms.value, "launches": cnt.value, "bytes": byt.value}
    return kernels


def work_counts(lib, reset):
    from gcge_old_amd import _native

    w = np.zeros(len(WORK_NAMES))
    lib.GCGE_HIP_WorkCounts(_native.as_dptr(w), len(w), 1 if reset else 0)
    return dict(zip(WORK_NAMES, w.tolist()))


def cpu_model(K, M, work, reps=60):
    """Seconds the reference's CSR path needs for `work` (counts per solve,
    work_counts) on one host core: every count priced by the oracle port's
    own single-vector call (app/csr/gcge_app_csr.c:297-378 restated, serial
    loops), timed here at the workload's size; the dense eigensolves by the
    reference's vendored LAPACK dsyev (oracle/_ref) at the solve's mean order,
    or the core's host QL solver when that library is absent."""
    import ctypes as C

    import oracle

    L = oracle.lib()
    dp = C.POINTER(C.c_double)
    n = K.nrows
    rng = np.random.default_rng(0)
    x = rng.standard_normal(n)
    y = rng.standard_normal(n)
    r = np.zeros(n)
    t_start = time.perf_counter()
    spmv = {}
    for name, A in (("K", K), ("M", M)):
        Am = oracle.Matrix(A)
        L.ORACLE_SpMM(Am.handle, 1, x.ctypes.data_as(dp), n, r.ctypes.data_as(dp), n)  # warm
        t0 = time.perf_counter()
        for _ in range(reps):
            L.ORACLE_SpMM(Am.handle, 1, x.ctypes.data_as(dp), n, r.ctypes.data_as(dp), n)
        spmv[name] = (time.perf_counter() - t0) / reps
        del Am
    per_nnz = (spmv["K"] + spmv["M"]) / (K.nnz + M.nnz)
    g = np.zeros(1)
    one = np.ones(1)
    L.ORACLE_Gram(n, 1, 1, x.ctypes.data_as(dp), n, y.ctypes.data_as(dp), n, g.ctypes.data_as(dp), 1)
    t0 = time.perf_counter()
    for _ in range(10 * reps):
        L.ORACLE_Gram(n, 1, 1, x.ctypes.data_as(dp), n, y.ctypes.data_as(dp), n,
                      g.ctypes.data_as(dp), 1)
    t_dot = (time.perf_counter() - t0) / (10 * reps)
    t0 = time.perf_counter()
    for _ in range(10 * reps):
        L.ORACLE_LinearComb(n, 1, 1, x.ctypes.data_as(dp), n, one.ctypes.data_as(dp), 1, 1.0,
                            y.ctypes.data_as(dp), n)
    t_axpby = (time.perf_counter() - t0) / (10 * reps)
    # dense eigensolve at the mean order of the solve's calls
    n_eig = int(round((work["dense_eig_n3"] / work["dense_eig"]) ** (1.0 / 3.0))) if work["dense_eig"] else 0
    t_eig, eig_kind = 0.0, "none"
    if n_eig > 0:
        a = rng.standard_normal((n_eig, n_eig))
        a = np.asfortranarray(a + a.T)
        t_eig, eig_kind = _time_dense_eig(a, reps=10)
    secs = (work["spmv_nnz"] * per_nnz + work["dot"] * t_dot + work["axpby"] * t_axpby
            + work["dense_eig"] * t_eig)
    sample_s = time.perf_counter() - t_start
    return {
        "value": round(secs, 1),
        "unit": "s",
        "cores": 1,
        "kind": "port",
        "sample": (f"the solve's work counted as the reference's single-vector CSR calls "
                   f"({work['spmv']:.0f} SpMV, {work['dot']:.0f} inner products, "
                   f"{work['axpby']:.0f} axpbys of length {n}, {work['dense_eig']:.0f} dense "
                   f"eigensolves of mean order {n_eig} per solve), each priced by the port's own "
                   f"call timed on this host at this size: SpMV K {spmv['K'] * 1e3:.1f} ms, "
                   f"M {spmv['M'] * 1e3:.1f} ms, inner product {t_dot * 1e3:.2f} ms, axpby "
                   f"{t_axpby * 1e3:.2f} ms, dense eigensolve {t_eig * 1e3:.1f} ms ({eig_kind}); "
                   f"{sample_s:.1f} s of timed CPU work"),
        "unit_costs_s": {"spmv_K": spmv["K"], "spmv_M": spmv["M"], "dot": t_dot, "axpby": t_axpby,
                         "dense_eig": t_eig},
        "work_per_solve": {k: round(v, 1) for k, v in work.items()},
    }


def _time_dense_eig(a, reps=3):
    """Seconds per dsyev of the symmetric matrix a: the reference's vendored
    netlib LAPACK when built, else the core's host QL solver."""
    import ctypes as C

    dp = C.POINTER(C.c_double)
    n = a.shape[0]
    lapack = os.path.join(ROOT, "oracle", "_ref", "liblapack_ref.so")
    w = np.zeros(n)
    if os.path.exists(lapack):
        L = C.CDLL(lapack)
        info = C.c_int(0)
        work = np.zeros(1)
        b = a.copy(order="F")
        L.dsyev_(b"V", b"U", C.byref(C.c_int(n)), b.ctypes.data_as(dp), C.byref(C.c_int(n)),
                 w.ctypes.data_as(dp), work.ctypes.data_as(dp), C.byref(C.c_int(-1)), C.byref(info))
        lw = int(work[0])
        work = np.zeros(lw)
        t0 = time.perf_counter()
        for _ in range(reps):
            b = a.copy(order="F")
            L.dsyev_(b"V", b"U", C.byref(C.c_int(n)), b.ctypes.data_as(dp), C.byref(C.c_int(n)),
                     w.ctypes.data_as(dp), work.ctypes.data_as(dp), C.byref(C.c_int(lw)),
                     C.byref(info))
        return (time.perf_counter() - t0) / reps, "reference netlib dsyev"
    from gcge_old_amd import _native

    core = _native.core()
    f = core.GCGE_Default_DenseMatEigenSolver
    ip = C.POINTER(C.c_int)
    f.argtypes = [dp, C.c_int, C.c_int, dp, dp, C.c_int, C.c_int, C.c_int, ip, dp]
    z = np.zeros((n, n), order="F")
    iw = np.zeros(10 * n, dtype=np.int32)
    dw = np.zeros(30 * n * n)
    t0 = time.perf_counter()
    for _ in range(reps):
        b = a.copy(order="F")
        f(b.ctypes.data_as(dp), n, n, w.ctypes.data_as(dp), z.ctypes.data_as(dp), n, 1, n,
          iw.ctypes.data_as(ip), dw.ctypes.data_as(dp))
    return (time.perf_counter() - t0) / reps, "host QL of the core"


def cpu_reference_run(K, M, nev, tol, iters, dev_iters):
    """The reference binary (oracle/_ref/gcge_ref, built from its sources, one
    core) on 1 and 1 + iters GCG iterations of the same pencil; the
    difference per iteration scaled to the device run's iteration count."""
    import oracle

    if not oracle.ref_available():
        return None
    flags = ["-gcge_ev_tol", str(tol)]
    r1 = oracle.ref_solve(K, M, nev, flags + ["-gcge_ev_max_it", "1"])
    r2 = oracle.ref_solve(K, M, nev, flags + ["-gcge_ev_max_it", str(1 + iters)])
    d_it = max(r2["num_iter"] - r1["num_iter"], 1)
    per_iter = (r2["time"] - r1["time"]) / d_it
    return {"value": round(per_iter * dev_iters, 1), "unit": "s", "cores": 1, "kind": "reference",
            "sample": f"reference binary: {r1['num_iter']} and {r2['num_iter']} GCG iterations in "
                      f"{r1['time']:.1f} s and {r2['time']:.1f} s on one host core, "
                      f"{per_iter:.2f} s per steady-state iteration, scaled to the device run's "
                      f"{dev_iters} iterations"}


def leading_complete_clusters(ev, rel_gap=1e-6):
    """Number of leading eigenpairs before the last cluster of ev (ascending;
    a cluster = eigenvalues within rel_gap of their neighbour)."""
    ev = np.sort(np.asarray(ev))
    m = len(ev)
    while m > 1 and abs(ev[m - 1] - ev[m - 2]) <= rel_gap * abs(ev[m - 1]):
        m -= 1
    return m - 1 if m > 1 else len(ev)


def reference_residuals(gcge, lib, dK, dM, ev, X):
    """Largest residual of the last solve's pairs in the reference's own
    definition (conv type "R", src/gcge_eigsol.c:704-721):
    ||K x - lambda M x|| / ||x|| / max(1, |lambda|), on the device."""
    from gcge_old_amd import _native

    k = X.shape[1]
    Xd = gcge.DeviceMultiVec.from_host(X)
    KX, MX = gcge.DeviceMultiVec(X.shape[0], k), gcge.DeviceMultiVec(X.shape[0], k)
    lib.GCGE_HIP_SpMM(dK.handle, Xd.handle, 0, KX.handle, 0, k)
    lib.GCGE_HIP_SpMM(dM.handle, Xd.handle, 0, MX.handle, 0, k)
    for j in range(k):  # KX_j -= lambda_j MX_j
        lib.GCGE_HIP_Axpby(-float(ev[j]), MX.handle, j, 1.0, KX.handle, j, 1)
    G = np.zeros((k, k), order="F")
    lib.GCGE_HIP_Gram(KX.handle, 0, k, KX.handle, 0, k, _native.as_dptr(G), k)
    rr = np.sqrt(np.maximum(np.diag(G), 0.0))
    lib.GCGE_HIP_Gram(Xd.handle, 0, k, Xd.handle, 0, k, _native.as_dptr(G), k)
    xn = np.sqrt(np.diag(G))
    return float(np.max(rr / xn / np.maximum(1.0, np.abs(ev))))


def principal_angle(gcge, lib, dM, X1, X2):
    """Largest principal angle between span(X1) and span(X2), both
    M-orthonormal: the singular values of X1^T M X2 are the cosines."""
    from gcge_old_amd import _native

    k = X1.shape[1]
    A = gcge.DeviceMultiVec.from_host(X1)
    B = gcge.DeviceMultiVec.from_host(X2)
    MB = gcge.DeviceMultiVec(X2.shape[0], k)
    lib.GCGE_HIP_SpMM(dM.handle, B.handle, 0, MB.handle, 0, k)
    G = np.zeros((k, k), order="F")
    lib.GCGE_HIP_Gram(A.handle, 0, k, MB.handle, 0, k, _native.as_dptr(G), k)
    s = np.linalg.svd(G, compute_uv=False)
    smin = float(min(s.min(), 1.0))
    return float(np.sqrt(max(0.0, 1.0 - smin * smin)))


def solve_workload(args):
    """One complete GCG solve per step on the Q1 pencil, in the value's
    precision (mixed at 128^3 lognormal, float64 at 64^3), the other precision
    on the same seeds beside it."""
    import gcge_old_amd as gcge
    from gcge_old_amd import _native, problems

    lib = _native.hip()
    lognormal = args.workload == "q1-128-lognormal"
    E = args.mesh or (128 if lognormal else 64)
    nev = args.nev or (100 if lognormal else 50)
    max_it = args.max_it or 500
    value_prec, other_prec = ("mixed", "fp64") if lognormal else ("fp64", "mixed")
    t0 = time.perf_counter()
    kappa = problems.lognormal_kappa(E, args.kappa_seed) if lognormal else None
    if args.assembly == "device":
        dK = gcge.DeviceMatrix.q1(E, kappa, "stiffness")
        dM = gcge.DeviceMatrix.q1(E, None, "mass")
        lib.GCGE_HIP_Synchronize()
        setup_s = time.perf_counter() - t0
        K = M = None
    else:
        K, M = problems.q1_fem_3d(E, kappa)
        t1 = time.perf_counter()
        dK, dM = gcge.DeviceMatrix(K), gcge.DeviceMatrix(M)
        lib.GCGE_HIP_Synchronize()
        setup_s = time.perf_counter() - t1
    n = dK.nrows
    flags = ["-gcge_conv_type", "RB", "-gcge_ev_tol", str(args.tol), "-gcge_ev_max_it", str(max_it)]
    if lognormal:
        flags += ["-gcge_block_size", str(2 * nev)]
    if not args.reference_locking:
        flags += ["-gcge_full_rr", "1"]
    if not args.reference_orth:
        flags += ["-gcge_w_orth_type", "svqb"]
    reorth = args.max_reorth_time
    if reorth is None and not args.reference_orth:
        reorth = 2
    if reorth is not None:
        flags += ["-gcge_max_reorth_time", str(reorth)]
    print(f"# workload: Q1 {E}^3 {'lognormal kappa ' if lognormal else ''}n={n} nnz(K)={dK.nnz} "
          f"nnz(M)={dM.nnz} nev={nev} ({args.assembly} assembly {setup_s:.2f}s, storage "
          f"{dK.storage_info()})", file=sys.stderr, flush=True)
    if args.pmc_probe:
        pmc_probe(dK, nev)
        return
    if args.kernels:
        kernel_sweep(dK, dM)

    def run(prec, seed, mode, its=max_it):
        lib.GCGE_HIP_ProfileEnable(mode)
        f = flags if its == max_it else flags[:5] + [str(its)] + flags[6:]
        out = gcge.solve(dK, dM, nev, f, seed=seed, precision=prec)
        lib.GCGE_HIP_ProfileEnable(0)
        return out

    prof_mode = 0 if args.no_profile else 2
    warm_launches = 0
    # buffers of the other precision allocated before anything is timed
    run(other_prec, 99, 0, its=2)
    for w in range(args.warmup):
        lib.GCGE_HIP_ProfileReset()
        out = run(value_prec, 100 + w, prof_mode)
        k = kernel_stats(lib)
        warm_launches = max(warm_launches, sum(v["launches"] for n_, v in k.items()
                                               if n_ in ("spmm", "spmm32")))
        print(f"# warmup {w} ({value_prec}): {out['time']:.4f}s iters={out['num_iter']} "
              f"conv={out['num_conv']}", file=sys.stderr, flush=True)

    # timed region: events on the product launches only (recorded by the
    # dispatches themselves), so the solve time carries no profiling
    # packets; the events are created beforehand (two per launch, with a
    # margin)
    lib.GCGE_HIP_ProfileReset()
    if not args.no_profile:
        lib.GCGE_HIP_ProfileReserve(int(2 * 1.5 * max(warm_launches, 2000) * args.steps))
    work_counts(lib, reset=True)
    lib.GCGE_HIP_Synchronize()
    times, iters, outs = [], [], []
    t_start = time.perf_counter()
    for s in range(args.steps):
        out = run(value_prec, 1 + s, prof_mode)
        times.append(out["time"])
        iters.append(out["num_iter"])
        outs.append(out)
        print(f"# step {s} ({value_prec}): {out['time']:.4f}s iters={out['num_iter']} "
              f"conv={out['num_conv']}", file=sys.stderr, flush=True)
    lib.GCGE_HIP_Synchronize()
    wall = time.perf_counter() - t_start
    kernels = kernel_stats(lib)
    work = {k: v / max(args.steps, 1) for k, v in work_counts(lib, reset=True).items()}
    work["vec_len"] = n
    value = float(np.mean(times))

    # the other precision on the same seeds
    other = None
    other_steps = min(args.steps, 2) if args.other_steps < 0 else args.other_steps
    if other_steps > 0:
        lib.GCGE_HIP_ProfileReset()
        o_times, o_iters, diffs, angles, o_outs = [], [], [], [], []
        angle_cols = None
        for s in range(other_steps):
            o = run(other_prec, 1 + s, prof_mode)
            o_times.append(o["time"])
            o_iters.append(o["num_iter"])
            ref = outs[s % len(outs)]
            both = min(o["num_conv"], ref["num_conv"])
            d = np.abs(o["eval"][:both] - ref["eval"][:both]) / np.abs(ref["eval"][:both])
            diffs.append(float(d.max()) if both else None)
            if both == nev:
                # over the eigenvectors below the last eigenvalue cluster: that
                # cluster may continue beyond nev (the kappa = 1 cube has
                # eigenvalues of multiplicity 3 and 6), and then any basis of
                # its computed part is as valid as another
                m = leading_complete_clusters(ref["eval"])
                angles.append(principal_angle(gcge, lib, dM, ref["evec"][:, :m], o["evec"][:, :m]))
                angle_cols = m
            print(f"# other {s} ({other_prec}): {o['time']:.4f}s iters={o['num_iter']} "
                  f"conv={o['num_conv']} max rel eig diff {diffs[-1]} angle "
                  f"{angles[-1] if angles else None}", file=sys.stderr, flush=True)
            o_outs.append({"num_conv": o["num_conv"], "res": o["res"]})
            del o
        ko = kernel_stats(lib)
        o_value = float(np.mean(o_times))
        other = {"precision": other_prec, "value": round(o_value, 5), "unit": "s",
                 "steps": other_steps, "iterations": o_iters,
                 "all_converged": all(o["num_conv"] == nev for o in o_outs),
                 "max_rel_eig_diff_vs_value": max(d for d in diffs if d is not None) if any(
                     d is not None for d in diffs) else None,
                 "max_principal_angle_vs_value": max(angles) if angles else None,
                 "principal_angle_columns": angle_cols,
                 "max_final_residual": float(max(np.max(o["res"]) for o in o_outs))}
        fp64_s = o_value if other_prec == "fp64" else value
        mixed_s = value if other_prec == "fp64" else o_value
        other["mixed_speedup_vs_fp64"] = round(fp64_s / mixed_s, 3)
        for name in ("spmm", "spmm32"):
            if name in ko and ko[name]["ms"] > 0:
                other[name + "_gbs"] = round(ko[name]["bytes"] / (ko[name]["ms"] * 1e-3) / 1e9, 1)

    # per-kernel breakdown: one more solve, untimed, every launch bracketed
    breakdown = {}
    if not args.no_profile and not args.no_breakdown:
        lib.GCGE_HIP_ProfileReset()
        run(value_prec, 1, 1)
        breakdown = kernel_stats(lib)

    dom = "spmm32" if value_prec == "mixed" else "spmm"
    roofline = None
    if dom in kernels and kernels[dom]["ms"] > 0:
        sp = kernels[dom]
        achieved = sp["bytes"] / (sp["ms"] * 1e-3) / 1e9
        key = "spmm32_cg" if dom == "spmm32" else "spmm64_k10"
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": read_traffic(E, lognormal, key + "_traffic_bytes"),
                    "traffic_algorithmic_bytes": read_traffic(E, lognormal, key + "_algorithmic_bytes"),
                    "traffic_note": "HBM bytes (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated) "
                                    "of one isolated launch of this product at "
                                    + ("the block-CG shape (float32, nev columns)" if dom == "spmm32"
                                       else "10 columns (the block-CG shape)")
                                    + ", next to that launch's algorithmic bytes",
                    "kernel": ("float32 block-CG products (direction block times K with p.w; "
                               "k_dia_tile LDS plane tiles from 8 columns, k_dia_spmm below)"
                               if dom == "spmm32" else
                               "float64 products (k_dia_tile LDS plane tiles from 8 columns, "
                               "k_dia_spmm below)"),
                    "launches": sp["launches"],
                    "avg_launch_us": round(sp["ms"] * 1e3 / sp["launches"], 2),
                    "algorithmic_bytes_per_launch": round(sp["bytes"] / sp["launches"]),
                    "bytes_formula": "nnz*(s_val+4) + 4(n+1) + s_x*k*(n_cols+n_rows), k = columns "
                                     "of the launch (the block CG records them on the device)"}
    spmm_rates = {name: {"gbs": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1),
                         "launches": v["launches"],
                         "avg_launch_us": round(v["ms"] * 1e3 / v["launches"], 2)}
                  for name, v in kernels.items() if name in ("spmm", "spmm32") and v["ms"] > 0}

    last = outs[-1]
    eig_err = None
    if not lognormal:
        exact = problems.q1_eigs(E, nev)
        eig_err = float(np.max(np.abs(np.sort(last["eval"]) - exact) / exact))

    cpu = None
    if args.cpu == "model":
        try:
            Kh = K if K is not None else dK.to_csr()
            Mh = M if M is not None else dM.to_csr()
            cpu = cpu_model(Kh, Mh, work)
            cpu["value"] = round(cpu["value"], 1)
            if args.cpu_reference_iters > 0:
                cpu["reference_run"] = cpu_reference_run(Kh, Mh, nev, args.tol, args.cpu_reference_iters,
                                                         int(np.mean(iters)))
            del Kh, Mh
        except Exception as e:  # the baseline is reported, never required
            print(f"# cpu baseline skipped: {e}", file=sys.stderr)

    X = gcge.DeviceMultiVec(n, nev)
    lib.GCGE_HIP_Synchronize()
    t = time.perf_counter()
    X.get()
    down = time.perf_counter() - t
    del X
    ref_res = reference_residuals(gcge, lib, dK, dM, last["eval"], last["evec"])
    line = {
        "metric": METRIC,
        "value": round(value, 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(value * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f32 inner block CG (directions, products, residuals, correction), f64 Gram/"
                  "orthonormalisation/Rayleigh-Ritz/residual check" if value_prec == "mixed" else "f64"),
        "data": "synthetic",
        "config": {"workload": f"Q1 FEM {E}^3 unit cube, "
                               f"{'lognormal kappa (seed %d)' % args.kappa_seed if lognormal else 'kappa=1'}, "
                               f"K x = lambda M x, nev={nev}"
                               f"{', block 200 requested' if lognormal else ''}, "
                               f"{'mixed precision' if value_prec == 'mixed' else 'FP64'}, "
                               f"||r||/(|lambda| ||Mx||) <= {args.tol:g}",
                   "precision": value_prec,
                   "rayleigh_ritz": "reference soft locking" if args.reference_locking else "full (-gcge_full_rr 1)",
                   "w_orth": "multi (reference)" if args.reference_orth else
                             "svqb (-gcge_w_orth_type svqb; one pass below Gram condition 1e4, else two)",
                   "projection_passes": ("1, a 2nd where a column kept < 1/sqrt(2) of its B-norm "
                                         "(W as a correction, -gcge_svqb_dgks_eta)"
                                         if not args.reference_orth and reorth == 2 else
                                         reorth if reorth is not None else 3),
                   "n": n, "nnz_K": dK.nnz, "nnz_M": dM.nnz, "nev": nev,
                   "block_size": last["block_size"], "assembly": args.assembly,
                   "parallelism": "1 GPU"},
        "iterations": iters,
        "all_converged": all(o["num_conv"] == nev for o in outs),
        "max_final_residual": float(max(np.max(o["res"]) for o in outs)),
        "max_final_residual_reference_R": ref_res,
        "max_rel_eig_err_vs_closed_form": eig_err,
        "wall_s_timed_region": round(wall, 4),
        "roofline": roofline,
        "spmm_rates": spmm_rates,
        "kernels_one_solve": {k: {"ms": round(v["ms"], 3), "launches": v["launches"]}
                              for k, v in breakdown.items()},
        "cpu_baseline": cpu,
        "host_transfer": {"setup_s": round(setup_s, 4), "assembly": args.assembly,
                          "evec_download_s": round(down, 5),
                          "s_per_solve_incl_transfers": round(value + down, 4),
                          "note": "setup = device assembly from kappa (or host CSR upload with "
                                  "--assembly host), once; value excludes it and the download "
                                  "of the eigenvectors (inputs resident)"},
        "fp64" if other_prec == "fp64" else "mixed": other,
    }
    print(json.dumps(line), flush=True)


def adjoint_workload(args):
    """Gradient of L = sum_k w_k lambda_k + ||P x_1||^2 with respect to every
    kappa_e of the lognormal Q1 64^3 problem (nev 20): one solve, the
    eigenvalue adjoint (element kernel), the eigenvector adjoint (shifted
    block CG + element kernel); then central differences of both terms on
    --fd random elements at each relative step of --fd-steps (default the
    specified 1e-6, and 1e-4, 1e-3, 1e-2), each difference quotient from two
    solves of the perturbed pencil (assembled on the device)."""
    import gcge_old_amd as gcge
    from gcge_old_amd import _native, adjoint, problems

    lib = _native.hip()
    E = args.mesh or 64
    nev = args.nev or 20
    n = (E - 1) ** 3
    kappa = problems.lognormal_kappa(E, args.kappa_seed)
    rng = np.random.default_rng(args.kappa_seed + 1)
    w = rng.uniform(size=nev)
    mask = (np.random.default_rng(args.kappa_seed + 2).uniform(size=n) < 0.5)
    # tolerance 1e-9: the lognormal pencil does not reach 1e-11 in 800
    # iterations
    flags = ["-gcge_conv_type", "RB", "-gcge_ev_tol", str(min(args.tol, 1e-9)),
             "-gcge_ev_max_it", "800", "-gcge_full_rr", "1", "-gcge_w_orth_type", "svqb",
             "-gcge_max_reorth_time", "2"]
    dM = gcge.DeviceMatrix.q1(E, None, "mass")

    def products(dA, X):
        dX = gcge.DeviceMultiVec.from_host(X)
        dY = gcge.DeviceMultiVec(X.shape[0], X.shape[1])
        lib.GCGE_HIP_SpMM(dA.handle, dX.handle, 0, dY.handle, 0, X.shape[1])
        return dY.get()

    def evaluate(kv, grad):
        dK = gcge.DeviceMatrix.q1(E, kv, "stiffness")
        ts = time.perf_counter()
        out = gcge.solve(dK, dM, nev, flags, seed=1)
        if out["num_conv"] < nev:
            raise RuntimeError(f"solve converged {out['num_conv']} of {nev} pairs")
        x1 = out["evec"][:, 0]
        if not grad:
            # difference quotients need both terms beyond the rounding of the
            # solver's Ritz values (~1e-13 relative, from its Gram sums):
            # eigenvalues as Rayleigh quotients of the computed vectors (error
            # quadratic in the vector error), and x_1 refined by one Newton
            # (Jacobi-Davidson) step, (K - lam M) d = -r with d M-orthogonal
            # to x_1, through the same shifted block CG
            X = out["evec"]
            KX, MX = products(dK, X), products(dM, X)
            rq = np.einsum("ij,ij->j", X, KX) / np.einsum("ij,ij->j", X, MX)
            lam1 = float(rq[0])
            r = KX[:, 0] - lam1 * MX[:, 0]
            x1 = x1 + adjoint.eigvec_adjoint_solve(dK, dM, x1, lam1, -r, rate=1e-12)
            x1 = x1 / np.sqrt(float(x1 @ products(dM, x1[:, None])[:, 0]))
            return (float(w @ rq), float(np.sum(x1[mask] ** 2))), out, None
        terms = (float(w @ out["eval"]), float(np.sum(x1[mask] ** 2)))
        t0 = time.perf_counter()
        g = adjoint.q1_kappa_gradient(E, out["evec"], w)
        t1 = time.perf_counter()
        gv, _ = adjoint.q1_eigvec_kappa_gradient(E, dK, dM, x1, out["eval"][0], 2.0 * mask * x1)
        t2 = time.perf_counter()
        return terms, out, (g, gv, t1 - t0, t2 - t1, t2 - ts)

    # the timed span starts with the matrices resident on the device (the
    # device assembly is excluded) and ends with both gradients on the host
    _, out, (g_val, g_vec, t_eval, t_vec, total) = evaluate(kappa, True)
    print(f"# adjoint: solve {out['time']:.3f}s iters={out['num_iter']} conv={out['num_conv']}, "
          f"eigenvalue adjoint {t_eval*1e3:.1f} ms, eigenvector adjoint {t_vec:.3f}s",
          file=sys.stderr, flush=True)
    elements = rng.choice(E ** 3, args.fd, replace=False)
    checks = {}
    for step in args.fd_steps:
        ev, evv, gm_val, gm_vec = [], [], [], []
        for e in elements:
            d = step * kappa[e]
            t = {}
            for j in (-1, 1):
                kv = kappa.copy()
                kv[e] += j * d
                t[j] = evaluate(kv, False)[0]
            fds = [(t[1][i] - t[-1][i]) / (2 * d) for i in range(2)]
            ev.append(abs(g_val[e] - fds[0]) / max(abs(fds[0]), 1e-300))
            evv.append(abs(g_vec[e] - fds[1]) / max(abs(fds[1]), 1e-300))
            gm_val.append(abs(g_val[e] - fds[0]) / np.abs(g_val).max())
            gm_vec.append(abs(g_vec[e] - fds[1]) / np.abs(g_vec).max())
        ev, evv = np.array(ev), np.array(evv)
        checks[f"{step:g}"] = {
            "eigenvalue_term_rel_err": {"max": float(ev.max()), "median": float(np.median(ev)),
                                        "within_1e-5": int((ev <= 1e-5).sum())},
            "eigenvector_term_rel_err": {"max": float(evv.max()), "median": float(np.median(evv)),
                                         "within_1e-5": int((evv <= 1e-5).sum())},
            "max_err_over_max_gradient": [float(max(gm_val)), float(max(gm_vec))]}
        print(f"# central differences, relative step {step:g}: {checks[f'{step:g}']}",
              file=sys.stderr, flush=True)
    homog = float(kappa @ g_val - w @ out["eval"]) / float(w @ out["eval"])
    line = {"metric": "s for the eigenpair gradient (solve + adjoints), all kappa_e",
            "value": round(total, 4), "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 0,
            "ms_per_step": round(total * 1e3, 2), "higher_is_better": False, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"Q1 FEM {E}^3 lognormal kappa (seed {args.kappa_seed}), nev={nev}, "
                                   "L = sum w_k lambda_k + ||P x_1||^2, dL/dkappa_e for all elements",
                       "n": n, "elements": E ** 3, "nev": nev},
            "solve_s": round(out["time"], 4), "eigenvalue_adjoint_ms": round(t_eval * 1e3, 3),
            "eigenvector_adjoint_s": round(t_vec, 4),
            "fd_elements": args.fd, "fd_checks_by_relative_step": checks,
            "homogeneity_rel_err": homog}
    print(json.dumps(line), flush=True)


def permuted_workload(args):
    """Scale and irregular access: the lognormal Q1 pencil on the 192^3 mesh
    (n = 191^3), rows and columns permuted by a seeded random permutation, so
    every product gathers X at random rows.  GCG iterations are bounded
    (the per-iteration cost is the measured quantity; a full solve takes
    hundreds of iterations at this size): seconds per iteration in float64
    and in the mixed form (difference of a 1-iteration and a
    (1 + --max-it)-iteration run, so that the set-up cancels), the
    products' rate, then dL/dkappa_e of
    L = sum_c w_c lambda_c for all E^3 elements on the device through the
    node -> row map, checked by the identity sum_e kappa_e g_e =
    sum_c w_c x_c^T K x_c (exact for any vectors, K linear in kappa)."""
    import ctypes as C

    import gcge_old_amd as gcge
    from gcge_old_amd import _native, adjoint, problems

    lib = _native.hip()
    E = args.mesh or 192
    nev = args.nev or 200
    iters = args.max_it or 3
    t0 = time.time()
    kappa = problems.lognormal_kappa(E, args.kappa_seed)
    n = (E - 1) ** 3
    perm = np.random.Generator(np.random.PCG64(args.kappa_seed + 1000)).permutation(n)
    # row i of the pencil holds node node_of_row[i]
    node_of_row = perm

    def inverse(p):
        q = np.empty(n, dtype=np.int32)
        q[p] = np.arange(n, dtype=np.int32)
        return q

    if args.assembly == "device":
        gcge.init()
        dK = gcge.DeviceMatrix.q1(E, kappa, "stiffness", inverse(node_of_row))
        dM = gcge.DeviceMatrix.q1(E, kappa, "mass", inverse(node_of_row))
    else:
        K, M = problems.q1_fem_3d(E, kappa)
        Kp = problems.symmetric_permutation(K, perm, sort_rows=False)
        Mp = problems.symmetric_permutation(M, perm, sort_rows=False)
        del K, M
        dK, dM = gcge.DeviceMatrix(Kp), gcge.DeviceMatrix(Mp)
    lib.GCGE_HIP_Synchronize()
    t_prep = time.time() - t0
    t_rcm = None
    if args.reorder == "rcm":
        from scipy.sparse.csgraph import reverse_cuthill_mckee

        t1 = time.time()
        Mp = dM.to_csr() if args.assembly == "device" else Mp
        g = Mp.to_scipy()  # M has every stored position of K
        g.sort_indices()
        r = reverse_cuthill_mckee(g, symmetric_mode=True)
        del g
        node_of_row = perm[r]
        if args.assembly == "device":
            del dK, dM, Mp
            dK = gcge.DeviceMatrix.q1(E, kappa, "stiffness", inverse(node_of_row))
            dM = gcge.DeviceMatrix.q1(E, kappa, "mass", inverse(node_of_row))
        else:
            Kp = problems.symmetric_permutation(Kp, r)
            Mp = problems.symmetric_permutation(Mp, r)
            dK, dM = gcge.DeviceMatrix(Kp), gcge.DeviceMatrix(Mp)
        t_rcm = time.time() - t1
    if args.assembly == "host":
        del Kp, Mp
    print(f"# workload: Q1 {E}^3 lognormal kappa, randomly permuted, n={n} nnz(K)={dK.nnz} "
          f"nnz(M)={dM.nnz} nev={nev} (set-up: {args.assembly} assembly + permutation + device "
          f"matrices {t_prep:.2f}s, reordering {args.reorder} {t_rcm}, storage "
          f"{dK.storage_info()})", file=sys.stderr, flush=True)
    def flags_for(max_it):
        return ["-gcge_conv_type", "RB", "-gcge_ev_tol", str(args.tol), "-gcge_ev_max_it",
                str(max_it), "-gcge_block_size", str(2 * nev), "-gcge_full_rr", "1",
                "-gcge_w_orth_type", "svqb", "-gcge_max_reorth_time", "2"]

    # seconds per iteration as the difference of a 1-iteration and a
    # (1 + iters)-iteration run from the same start: the set-up (the initial
    # X orthonormalisation takes nev single-column products) cancels
    legs = {}
    out = None
    for prec in ("fp64", "mixed"):
        first = gcge.solve(dK, dM, nev, flags_for(1), seed=1, precision=prec)
        lib.GCGE_HIP_ProfileReset()
        lib.GCGE_HIP_ProfileEnable(2)
        out = gcge.solve(dK, dM, nev, flags_for(1 + iters), seed=1, precision=prec)
        lib.GCGE_HIP_ProfileEnable(0)
        k = kernel_stats(lib)
        d_it = max(out["num_iter"] - first["num_iter"], 1)
        legs[prec] = {"s_per_iteration": round((out["time"] - first["time"]) / d_it, 4),
                      "iterations": [first["num_iter"], out["num_iter"]],
                      "s_total": [round(first["time"], 3), round(out["time"], 3)],
                      "block_size": out["block_size"]}
        del first
        for name in ("spmm", "spmm32"):  # float64 products; float32 inner-CG products
            if name in k and k[name]["ms"] > 0:
                sp = k[name]
                legs[prec][name + "_gbs"] = round(sp["bytes"] / (sp["ms"] * 1e-3) / 1e9, 1)
                legs[prec][name + "_avg_launch_ms"] = round(sp["ms"] / sp["launches"], 3)
        print(f"# {prec}: {legs[prec]}", file=sys.stderr, flush=True)
        if prec == "fp64":
            X = out["evec"]
    dK.set_precision(False)
    dM.set_precision(False)
    # eigenvalue adjoint over all elements on the (float64 leg's) vectors
    w = np.random.default_rng(args.kappa_seed + 1).uniform(size=nev)
    Xd = gcge.DeviceMultiVec.from_host(X)
    del X
    inv = np.empty(n, dtype=np.int64)
    inv[node_of_row] = np.arange(n)
    lib.GCGE_HIP_Synchronize()
    t = time.perf_counter()
    g = adjoint.q1_kappa_gradient(E, Xd, w, node_row=inv)
    t_adj = time.perf_counter() - t
    Yd = gcge.DeviceMultiVec(n, nev)
    lib.GCGE_HIP_SpMM(dK.handle, Xd.handle, 0, Yd.handle, 0, nev)
    G = np.zeros((nev, nev), order="F")
    lib.GCGE_HIP_Gram(Xd.handle, 0, nev, Yd.handle, 0, nev, _native.as_dptr(G), nev)
    xkx = float(w @ np.diag(G))
    ident = abs(float(kappa @ g) - xkx) / abs(xkx)
    fp64, mixed = legs["fp64"]["s_per_iteration"], legs["mixed"]["s_per_iteration"]
    line = {"metric": "s per GCG iteration (difference of bounded runs) on the permuted 192^3 pencil: "
                      "FP64 vs FP32-inner",
            "value": fp64, "unit": "s", "n_gpus": 1, "steps": iters, "warmup": 0,
            "ms_per_step": round(fp64 * 1e3, 2), "higher_is_better": False, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"Q1 FEM {E}^3 lognormal kappa (seed {args.kappa_seed}), rows and "
                                   "columns randomly permuted, nev=200, block 400 requested"
                                   + (", then reordered by reverse Cuthill-McKee" if t_rcm is not None else ""),
                       "reorder": args.reorder, "reorder_s": t_rcm,
                       "assembly": args.assembly, "setup_s": round(t_prep, 3),
                       "n": n, "nnz_K": dK.nnz, "nnz_M": dM.nnz, "nev": nev,
                       "rayleigh_ritz": "full (-gcge_full_rr 1)", "w_orth": "svqb",
                       "projection_passes": "1, a 2nd where a column kept < 1/sqrt(2) of its B-norm"},
            "fp64": legs["fp64"], "mixed": legs["mixed"],
            "mixed_speedup_per_iteration": round(fp64 / mixed, 3),
            "eigenvalue_adjoint_all_elements_ms": round(t_adj * 1e3, 2),
            "adjoint_identity_rel_err": ident,
            "adjoint_identity_sides": [float(kappa @ g), xkx], "elements": E ** 3}
    print(json.dumps(line), flush=True)



def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU; --gpus {args.gpus} is not supported (the "
              "multi-GPU row partition is a design, not built)", file=sys.stderr)
        sys.exit(2)
    import gcge_old_amd as gcge

    gcge.init(0)
    if args.mfma_probe:
        mfma_probe()
        return
    if args.workload == "q1-64-adjoint":
        adjoint_workload(args)
        return
    if args.workload == "q1-192-permuted":
        permuted_workload(args)
        return
    solve_workload(args)


if __name__ == "__main__":
    main()
