"""Benchmark: seconds to nev eigenpairs of the Q1 FEM pencil on one MI355X.

Workload (one "step" = one complete GCG solve): trilinear Q1 hexahedral
elements on the uniform 64^3 mesh of the unit cube, Dirichlet nodes
eliminated (n = 63^3 = 250047), stiffness K and consistent mass M with
kappa = 1, nev = 50 smallest eigenpairs of K x = lambda M x, float64
throughout, converged to ||Kx - lambda Mx|| / (|lambda| ||Mx||) <= 1e-8.
Matrices are uploaded before timing; the initial block is drawn on the
device inside each solve from the seed.  The solver runs with
-gcge_full_rr 1: the converged prefix stays in the Rayleigh-Ritz step with
its computed coupling.  The reference's soft locking zeroes that coupling,
which at this tolerance stalls some starts short of nev pairs (the reference
build does the same); --reference-locking restores it.  The W block is
B-orthonormalised with -gcge_w_orth_type svqb (the reference's projection
against [X, P], then two SVQB passes instead of the recursive single-column
leaves: 5 host round trips instead of about 40, same span);
--reference-orth restores "multi".

The JSON line carries:
  value          mean seconds per solve over the timed steps (lower is better)
  roofline       the block SpMM (dominant kernel): algorithmic bytes over its
                 HIP-event time summed over the timed region, against 8 TB/s
                 (the events of the SpMM launches are recorded by the
                 dispatch packets, hipExtLaunchKernelGGL; nothing else is
                 bracketed during the timed steps); traffic from the
                 committed rocprofv3 counter summary in profiles/
  kernels_one_solve  per-kernel event time of one extra, untimed solve with
                 every launch bracketed
  cpu_baseline   the reference CSR path built from its own sources (or the
                 oracle restatement when that binary is absent), one core,
                 timed on a bounded number of GCG iterations of the same
                 workload and scaled to the device run's iteration count
  mixed          the mixed-precision form on the same workload and seeds
                 (float32 products and directions in the inner block CG,
                 float64 everywhere else): seconds per solve, speed-up over
                 the float64 solve and the largest relative eigenvalue
                 difference from it
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
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--workload", choices=("q1-64", "q1-128-lognormal", "q1-64-adjoint",
                                          "q1-192-permuted"),
                   default="q1-64",
                   help="q1-64: Q1 64^3, kappa = 1, nev 50 (default, float64 value); "
                        "q1-128-lognormal: Q1 128^3, lognormal kappa, nev 100, block 200 "
                        "requested (the mixed-precision comparison on one seed); "
                        "q1-64-adjoint: Q1 64^3 lognormal kappa, nev 20, gradient of "
                        "sum w_k lambda_k + ||P x_1||^2 with respect to all kappa_e, "
                        "checked by fourth-order central differences on --fd elements; "
                        "q1-192-permuted: Q1 192^3 lognormal kappa, rows and columns "
                        "symmetrically permuted by a seeded random permutation, nev 200, block "
                        "400 requested, a bounded number of GCG iterations (--max-it, default "
                        "3 here) in float64 and mixed precision, then the eigenvalue adjoint on "
                        "all elements")
    p.add_argument("--fd", type=int, default=8, help="elements checked by finite differences "
                                                     "(q1-64-adjoint)")
    p.add_argument("--reorder", choices=("none", "rcm"), default="none",
                   help="q1-192-permuted: reorder the permuted pencil by reverse Cuthill-McKee "
                        "(computed from the matrix graph, a preprocessing step) before the solve")
    p.add_argument("--assembly", choices=("device", "host"), default="device",
                   help="q1-192-permuted: assemble K and M on the device from kappa and the "
                        "node -> row map, or on the host (numpy) and copy the CSR arrays")
    p.add_argument("--mesh", type=int, default=None, help="elements per axis (overrides the workload)")
    p.add_argument("--nev", type=int, default=None)
    p.add_argument("--kappa-seed", type=int, default=1)
    p.add_argument("--tol", type=float, default=1e-8)
    p.add_argument("--max-it", type=int, default=None,
                   help="GCG iteration cap (a bounded run for profiling; solves stopped by it "
                        "report all_converged false)")
    p.add_argument("--cpu-iters", type=int, default=3,
                   help="GCG iterations timed for the CPU baseline (0: skip)")
    p.add_argument("--mixed-steps", type=int, default=-1,
                   help="solves of the mixed-precision form (default: --steps; 0: skip)")
    p.add_argument("--reference-locking", action="store_true",
                   help="soft locking as in the reference (no -gcge_full_rr)")
    p.add_argument("--reference-orth", action="store_true",
                   help="recursive \"multi\" W orthonormalisation as in the reference "
                        "(default here: -gcge_w_orth_type svqb)")
    p.add_argument("--no-profile", action="store_true",
                   help="time without per-kernel events (no roofline numbers)")
    p.add_argument("--mfma-probe", action="store_true",
                   help="isolated Gram / linear-combination launches for matrix-core counters")
    p.add_argument("--pmc-probe", action="store_true",
                   help="only run 5 SpMM launches (K, k=10) and 5 plain copies of n x 64 doubles, "
                        "for rocprofv3 --pmc counter passes; prints one line")
    p.add_argument("--kernels", action="store_true",
                   help="also time SpMM, Gram, linear combination and a plain copy in isolation "
                        "(printed on stderr; the copy is the calibration for HBM counters)")
    return p.parse_args()


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "spmm_q1_64_k10_hbm_traffic.md")


def read_traffic():
    """HBM bytes per SpMM launch (K, k = 10) from the committed counter
    summary: the line 'traffic_bytes_per_launch: <value>'."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def pmc_probe(K):
    """Isolated launches for the counter passes: SpMM on K with 10 columns
    (the block-CG shape of the workload) and a plain copy of known size."""
    import gcge_old_amd as gcge
    from gcge_old_amd import _native

    lib = _native.hip()
    n = K.nrows
    dK = gcge.DeviceMatrix(K)
    X = gcge.DeviceMultiVec(n, 64)
    Y = gcge.DeviceMultiVec(n, 64)
    lib.GCGE_HIP_RandomFill(X.handle, 0, 64)
    for _ in range(5):
        lib.GCGE_HIP_SpMM(dK.handle, X.handle, 0, Y.handle, 0, 10)
    for _ in range(5):
        lib.GCGE_HIP_Axpby(1.0, X.handle, 0, 0.0, Y.handle, 0, 64)
    lib.GCGE_HIP_Synchronize()
    spmm_bytes = K.nnz * 12 + 4 * (n + 1) + 8 * 10 * 2 * n
    print(json.dumps({"pmc_probe": True, "spmm_algorithmic_bytes": spmm_bytes,
                      "copy_bytes_read": 8 * n * 64, "copy_bytes_written": 8 * n * 64}), flush=True)


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
            res[name + "_us"] = round(ms.value * 1e3 / max(cnt.value, 1), 1)
        print(json.dumps(res), flush=True)
        del V, W


def kernel_sweep(K, M):
    """Isolated kernel timings on the workload's matrices (stderr)."""
    import ctypes as C

    import gcge_old_amd as gcge
    from gcge_old_amd import _native

    lib = _native.hip()
    n = M.nrows
    dK, dM = gcge.DeviceMatrix(K), gcge.DeviceMatrix(M)
    X = gcge.DeviceMultiVec(n, 96)
    Y = gcge.DeviceMultiVec(n, 96)
    lib.GCGE_HIP_RandomFill(X.handle, 0, 96)

    def timed(name, fn, reps=20, extra=()):
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
        for k in (1, 4, 10, 16, 50):
            us, gbs = timed("spmm", lambda: lib.GCGE_HIP_SpMM(mat.handle, X.handle, 0, Y.handle, 0, k))
            print(f"# kernel spmm {label} k={k}: {us:.1f} us/launch, {gbs:.0f} GB/s", file=sys.stderr)
    G = np.zeros((96, 96), order="F")
    for m, k in ((70, 10), (60, 60), (90, 10), (10, 10), (1, 1)):
        us, gbs = timed("gram", lambda: lib.GCGE_HIP_Gram(X.handle, 0, m, Y.handle, 0, k,
                                                          _native.as_dptr(G), 96),
                        extra=("gram_reduce",))
        print(f"# kernel gram m={m} k={k} (with its reduction pass): {us:.1f} us/call, {gbs:.0f} GB/s",
              file=sys.stderr)
    Cm = np.asfortranarray(np.random.default_rng(0).standard_normal((96, 96)))
    for m, k in ((90, 50), (80, 30), (70, 10)):
        us, gbs = timed("lincomb", lambda: lib.GCGE_HIP_LinearComb(X.handle, 0, m, _native.as_dptr(Cm),
                                                                   96, 0.0, Y.handle, 0, k))
        print(f"# kernel lincomb m={m} k={k}: {us:.1f} us/launch, {gbs:.0f} GB/s", file=sys.stderr)
    # plain copy of 64 columns (n x 64 doubles read and written): the counter
    # calibration dispatch
    us, _ = timed("axpby", lambda: lib.GCGE_HIP_Axpby(1.0, X.handle, 0, 0.0, Y.handle, 0, 64))
    copy_bytes = 2 * 8.0 * n * 64
    print(f"# kernel copy {copy_bytes / 2**20:.1f} MiB moved: {us:.1f} us/launch, "
          f"{copy_bytes / (us * 1e-6) / 1e9:.0f} GB/s", file=sys.stderr)


KERNELS = ("spmm", "spmm32", "gram", "gram_reduce", "lincomb", "axpby", "axpby_list", "swap",
           "coldots", "bcg_p", "bcg_xr", "axpy_cols", "residual", "random")


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


def host_transfer(gcge, lib, n, nev, upload_s, value):
    """The solve with its host transfers: the matrices' upload (host CSR to
    the device layouts, once) and the eigenvectors' download (n x nev
    doubles, once per solve), next to the resident-input value."""
    X = gcge.DeviceMultiVec(n, nev)
    lib.GCGE_HIP_Synchronize()
    t0 = time.perf_counter()
    X.get()
    down = time.perf_counter() - t0
    return {"matrix_upload_s": round(upload_s, 4), "evec_download_s": round(down, 5),
            "s_per_solve_incl_transfers": round(value + upload_s + down, 4),
            "note": "upload includes the host-side sliced-ELLPACK/diagonal conversion; "
                    "value excludes both transfers (inputs resident)"}


def mixed_leg(gcge, lib, dK, dM, nev, flags, steps, fp64_outs, fp64_value):
    """The mixed-precision form on the same seeds as the float64 steps."""
    lib.GCGE_HIP_ProfileReset()
    lib.GCGE_HIP_ProfileEnable(2)
    times, iters, errs = [], [], []
    for s in range(steps):
        out = gcge.solve(dK, dM, nev, flags, seed=1 + s, precision="mixed")
        times.append(out["time"])
        iters.append(out["num_iter"])
        ref = fp64_outs[s % len(fp64_outs)]["eval"]
        errs.append(float(np.max(np.abs(np.sort(out["eval"]) - np.sort(ref)) / np.abs(ref))))
        print(f"# mixed step {s}: {out['time']:.4f}s iters={out['num_iter']} "
              f"conv={out['num_conv']} max rel eig diff {errs[-1]:.2e}", file=sys.stderr, flush=True)
    lib.GCGE_HIP_ProfileEnable(0)
    k = kernel_stats(lib)
    dK.set_precision(False)
    dM.set_precision(False)
    value = float(np.mean(times))
    out = {"value": round(value, 5), "unit": "s", "speedup_vs_fp64": round(fp64_value / value, 3),
           "iterations": iters, "max_rel_eig_diff_vs_fp64": max(errs),
           "dtype": "f32 inner block CG products and directions, f64 elsewhere"}
    if "spmm32" in k and k["spmm32"]["ms"] > 0:
        sp = k["spmm32"]
        out["spmm32_gbs"] = round(sp["bytes"] / (sp["ms"] * 1e-3) / 1e9, 1)
    return out


def cpu_baseline(K, M, nev, tol, iters, dev_iters):
    """Time `iters` GCG iterations of the reference CSR path on one core and
    scale per-iteration cost to the device run's iteration count."""
    import oracle

    flags = ["-gcge_ev_max_it", str(iters), "-gcge_ev_tol", str(tol)]
    if oracle.ref_available():
        kind = "reference"
        r = oracle.ref_solve(K, M, nev, flags)
        secs, done = r["time"], r["num_iter"]
    else:
        kind = "port"
        r = oracle.solve(K, M, nev, flags + ["-gcge_conv_type", "RB"])
        secs, done = r["time"], r["num_iter"]
    per_iter = secs / max(done, 1)
    return {
        "value": per_iter * dev_iters,
        "unit": "s",
        "cores": 1,
        "kind": kind,
        "sample": f"{done} GCG iterations of the same workload in {secs:.1f} s on one host core "
                  f"({per_iter:.2f} s/iteration), scaled to the device run's {dev_iters} iterations",
    }


def adjoint_workload(args):
    """Gradient of L = sum_k w_k lambda_k + ||P x_1||^2 with respect to every
    kappa_e of the lognormal Q1 64^3 problem (nev 20): one solve, the
    eigenvalue adjoint (element kernel), the eigenvector adjoint (shifted
    block CG + element kernel); central differences on a few elements."""
    import gcge_old_amd as gcge
    from gcge_old_amd import adjoint, problems

    E = args.mesh or 64
    nev = args.nev or 20
    kappa = problems.lognormal_kappa(E, args.kappa_seed)
    rng = np.random.default_rng(args.kappa_seed + 1)
    w = rng.uniform(size=nev)
    # tolerance 1e-9: the lognormal pencil does not reach 1e-11 in 800
    # iterations
    flags = ["-gcge_conv_type", "RB", "-gcge_ev_tol", str(min(args.tol, 1e-9)),
             "-gcge_ev_max_it", "800", "-gcge_full_rr", "1", "-gcge_w_orth_type", "svqb"]

    def evaluate(kv, grad):
        K, M = problems.q1_fem_3d(E, kv)
        dK, dM = gcge.DeviceMatrix(K), gcge.DeviceMatrix(M)
        ts = time.perf_counter()
        out = gcge.solve(dK, dM, nev, flags, seed=1)
        if out["num_conv"] < nev:
            raise RuntimeError(f"solve converged {out['num_conv']} of {nev} pairs")
        x1 = out["evec"][:, 0]
        mask = (np.random.default_rng(args.kappa_seed + 2).uniform(size=K.nrows) < 0.5)
        if not grad:
            # difference quotients need both terms beyond the rounding of the
            # solver's Ritz values (~1e-13 relative, from its Gram sums):
            # eigenvalues as Rayleigh quotients of the computed vectors
            # (error quadratic in the vector error), and x_1 refined by one
            # Newton (Jacobi-Davidson) step, (K - lam M) d = -r with d
            # M-orthogonal to x_1, through the same shifted block CG
            Ks, Ms = K.to_scipy(), M.to_scipy()
            X = out["evec"]
            rq = np.einsum("ij,ij->j", X, Ks @ X) / np.einsum("ij,ij->j", X, Ms @ X)
            lam1 = float(rq[0])
            r = Ks @ x1 - lam1 * (Ms @ x1)
            x1 = x1 + adjoint.eigvec_adjoint_solve(dK, dM, x1, lam1, -r, rate=1e-12)
            x1 = x1 / np.sqrt(float(x1 @ (Ms @ x1)))
            return (float(w @ rq), float(np.sum(x1[mask] ** 2))), out, None
        terms = (float(w @ out["eval"]), float(np.sum(x1[mask] ** 2)))
        t0 = time.perf_counter()
        g = adjoint.q1_kappa_gradient(E, out["evec"], w)
        t1 = time.perf_counter()
        gv, _ = adjoint.q1_eigvec_kappa_gradient(E, dK, dM, x1, out["eval"][0], 2.0 * mask * x1)
        t2 = time.perf_counter()
        return terms, out, (g, gv, t1 - t0, t2 - t1, t2 - ts)

    # the timed span starts with the matrices resident on the device (host
    # assembly and upload excluded) and ends with both gradients on the host
    _, out, (g_val, g_vec, t_eval, t_vec, total) = evaluate(kappa, True)
    print(f"# adjoint: solve {out['time']:.3f}s iters={out['num_iter']} conv={out['num_conv']}, "
          f"eigenvalue adjoint {t_eval*1e3:.1f} ms, eigenvector adjoint {t_vec:.3f}s",
          file=sys.stderr, flush=True)
    # fourth-order central differences with the relative step 1e-2: one
    # element moves L by ~1e-9 of its size per unit step, and x_1 is only
    # determined to ~1e-11 in float64 (the residual K x - lam M x cannot be
    # evaluated more accurately), so smaller steps drown in that noise at
    # elements with small gradients; the truncation error of the
    # fourth-order formula at this step is ~1e-8
    errs_val, errs_vec, errs_val_max, errs_vec_max = [], [], [], []
    for e in rng.choice(E ** 3, args.fd, replace=False):
        d = 1e-2 * kappa[e]
        t = {}
        for j in (-2, -1, 1, 2):
            kv = kappa.copy()
            kv[e] += j * d
            t[j] = evaluate(kv, False)[0]
        fds = [(t[-2][i] - 8 * t[-1][i] + 8 * t[1][i] - t[2][i]) / (12 * d) for i in range(2)]
        errs_val.append(abs(g_val[e] - fds[0]) / max(abs(fds[0]), 1e-300))
        errs_vec.append(abs(g_vec[e] - fds[1]) / max(abs(fds[1]), 1e-300))
        errs_val_max.append(abs(g_val[e] - fds[0]) / np.abs(g_val).max())
        errs_vec_max.append(abs(g_vec[e] - fds[1]) / np.abs(g_vec).max())
        print(f"# element {e}: eigenvalue term adjoint {g_val[e]:.10e} fd {fds[0]:.10e} "
              f"rel {errs_val[-1]:.2e}; eigenvector term adjoint {g_vec[e]:.10e} fd {fds[1]:.10e} "
              f"rel {errs_vec[-1]:.2e}", file=sys.stderr, flush=True)
    homog = float(kappa @ g_val - w @ out["eval"]) / float(w @ out["eval"])
    line = {"metric": "s for the eigenpair gradient (solve + adjoints), all kappa_e",
            "value": round(total, 4), "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 0,
            "ms_per_step": round(total * 1e3, 2), "higher_is_better": False, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"Q1 FEM {E}^3 lognormal kappa (seed {args.kappa_seed}), nev={nev}, "
                                   "L = sum w_k lambda_k + ||P x_1||^2, dL/dkappa_e for all elements",
                       "n": (E - 1) ** 3, "elements": E ** 3, "nev": nev},
            "solve_s": round(out["time"], 4), "eigenvalue_adjoint_ms": round(t_eval * 1e3, 3),
            "eigenvector_adjoint_s": round(t_vec, 4),
            "fd_elements": args.fd,
            "max_rel_err_vs_fd_eigenvalue_term": max(errs_val) if errs_val else None,
            "max_rel_err_vs_fd_eigenvector_term": max(errs_vec) if errs_vec else None,
            "max_err_vs_fd_over_max_gradient": [max(errs_val_max), max(errs_vec_max)] if errs_val else None,
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
                "-gcge_w_orth_type", "svqb"]

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
                       "rayleigh_ritz": "full (-gcge_full_rr 1)", "w_orth": "svqb"},
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
    from gcge_old_amd import _native, problems

    gcge.init(0)
    if args.mfma_probe:
        mfma_probe()
        return
    lib = _native.hip()
    if args.workload == "q1-64-adjoint":
        adjoint_workload(args)
        return
    if args.workload == "q1-192-permuted":
        permuted_workload(args)
        return
    if args.max_it is None:
        args.max_it = 500
    lognormal = args.workload == "q1-128-lognormal"
    E = args.mesh or (128 if lognormal else 64)
    if args.nev is None:
        args.nev = 100 if lognormal else 50
    t0 = time.time()
    kappa = problems.lognormal_kappa(E, args.kappa_seed) if lognormal else None
    K, M = problems.q1_fem_3d(E, kappa)
    assemble_s = time.time() - t0
    # the host CSR arrays cross PCIe once (with the host-side layout conversion)
    t0 = time.perf_counter()
    dK, dM = gcge.DeviceMatrix(K), gcge.DeviceMatrix(M)
    lib.GCGE_HIP_Synchronize()
    upload_s = time.perf_counter() - t0
    flags = ["-gcge_conv_type", "RB", "-gcge_ev_tol", str(args.tol), "-gcge_ev_max_it", str(args.max_it)]
    if lognormal:
        flags += ["-gcge_block_size", str(2 * args.nev)]
    if not args.reference_locking:
        flags += ["-gcge_full_rr", "1"]
    if not args.reference_orth:
        flags += ["-gcge_w_orth_type", "svqb"]
    print(f"# workload: Q1 {E}^3 {'lognormal kappa ' if lognormal else ''}n={K.nrows} nnz(K)={K.nnz} "
          f"nnz(M)={M.nnz} nev={args.nev} (assembly {assemble_s:.1f}s)", file=sys.stderr, flush=True)
    if args.pmc_probe:
        pmc_probe(K)
        return
    if args.kernels:
        kernel_sweep(K, M)

    warm_launches = 0
    for w in range(args.warmup):
        lib.GCGE_HIP_ProfileReset()
        lib.GCGE_HIP_ProfileEnable(0 if args.no_profile else 2)
        out = gcge.solve(dK, dM, args.nev, flags, seed=100 + w)
        lib.GCGE_HIP_ProfileEnable(0)
        warm_launches = max(warm_launches, kernel_stats(lib).get("spmm", {}).get("launches", 0))
        print(f"# warmup {w}: {out['time']:.4f}s iters={out['num_iter']} conv={out['num_conv']}",
              file=sys.stderr, flush=True)

    # timed region: events on the SpMM launches only (recorded by the
    # dispatches themselves), so the solve time carries no profiling packets;
    # the events are created beforehand (two per launch, with a margin)
    lib.GCGE_HIP_ProfileReset()
    if not args.no_profile:
        lib.GCGE_HIP_ProfileReserve(int(2 * 1.5 * max(warm_launches, 2000) * args.steps))
    lib.GCGE_HIP_ProfileEnable(0 if args.no_profile else 2)
    lib.GCGE_HIP_Synchronize()
    times, iters, outs = [], [], []
    t_start = time.perf_counter()
    for s in range(args.steps):
        out = gcge.solve(dK, dM, args.nev, flags, seed=1 + s)
        times.append(out["time"])
        iters.append(out["num_iter"])
        outs.append(out)
        print(f"# step {s}: {out['time']:.4f}s iters={out['num_iter']} conv={out['num_conv']}",
              file=sys.stderr, flush=True)
    lib.GCGE_HIP_Synchronize()
    wall = time.perf_counter() - t_start
    lib.GCGE_HIP_ProfileEnable(0)
    kernels = kernel_stats(lib)

    # per-kernel breakdown: one more solve, untimed, every launch bracketed
    breakdown = {}
    if not args.no_profile:
        lib.GCGE_HIP_ProfileReset()
        lib.GCGE_HIP_ProfileEnable(1)
        gcge.solve(dK, dM, args.nev, flags, seed=1)
        lib.GCGE_HIP_ProfileEnable(0)
        breakdown = kernel_stats(lib)

    value = float(np.mean(times))
    roofline = None
    if "spmm" in kernels and kernels["spmm"]["ms"] > 0:
        sp = kernels["spmm"]
        achieved = sp["bytes"] / (sp["ms"] * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    # the counter summary covers the 64^3 K product only
                    "traffic": read_traffic() if (E == 64 and not lognormal) else None,
                    "traffic_note": "HBM bytes per SpMM launch on K with 10 columns (the block-CG "
                                    "shape), rocprofv3 FETCH_SIZE + WRITE_SIZE scaled by a copy "
                                    "calibration; algorithmic bytes of that launch (CSR formula): "
                                    f"{K.nnz * 12 + 4 * (K.nrows + 1) + 160 * K.nrows}; the "
                                    "diagonal-storage kernel reads no column indices",
                    "kernel": "spmm", "launches": sp["launches"],
                    "avg_launch_us": round(sp["ms"] * 1e3 / sp["launches"], 2),
                    "algorithmic_bytes_per_launch": round(sp["bytes"] / sp["launches"])}

    # accuracy of the last solve against the closed form (kappa = 1 only)
    last = outs[-1]
    eig_err = None
    if not lognormal:
        exact = problems.q1_eigs(E, args.nev)
        eig_err = float(np.max(np.abs(np.sort(last["eval"]) - exact) / exact))

    mixed = None
    mixed_steps = args.steps if args.mixed_steps < 0 else args.mixed_steps
    if mixed_steps > 0:
        mixed = mixed_leg(gcge, lib, dK, dM, args.nev, flags, mixed_steps, outs, value)

    cpu = None
    if args.cpu_iters > 0:
        try:
            cpu = cpu_baseline(K, M, args.nev, args.tol, args.cpu_iters, int(np.mean(iters)))
        except Exception as e:  # the baseline is reported, never required
            print(f"# cpu baseline skipped: {e}", file=sys.stderr)

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
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"Q1 FEM {E}^3 unit cube, "
                               f"{'lognormal kappa (seed %d)' % args.kappa_seed if lognormal else 'kappa=1'}, "
                               f"K x = lambda M x, nev={args.nev}, FP64, "
                               f"||r||/(|lambda| ||Mx||) <= {args.tol:g}",
                   "rayleigh_ritz": "reference soft locking" if args.reference_locking else "full (-gcge_full_rr 1)",
                   "w_orth": "multi (reference)" if args.reference_orth else "svqb (-gcge_w_orth_type svqb)",
                   "n": K.nrows, "nnz_K": K.nnz, "nnz_M": M.nnz, "nev": args.nev,
                   "block_size": last["block_size"], "parallelism": "1 GPU"},
        "iterations": iters,
        "all_converged": all(o["num_conv"] == args.nev for o in outs),
        "max_rel_eig_err_vs_closed_form": eig_err,
        "wall_s_timed_region": round(wall, 4),
        "roofline": roofline,
        "kernels_one_solve": {k: {"ms": round(v["ms"], 3), "launches": v["launches"]}
                              for k, v in breakdown.items()},
        "cpu_baseline": cpu,
        "host_transfer": host_transfer(gcge, lib, K.nrows, args.nev, upload_s, value),
        "mixed": mixed,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
