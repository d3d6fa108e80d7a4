"""ORACLE -- test infrastructure only.

Python access to the CPU restatement (liboracle_csr.so) and to the reference
built from its own sources (_ref/gcge_ref).  Only tests/, the smoke check and
the CPU-baseline leg of bench.py import this package; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_csr.so")
REF_BIN = os.path.join(HERE, "_ref", "gcge_ref")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built (make -C oracle)")
        _lib = C.CDLL(LIB)
        L = _lib
        L.ORACLE_CSR_Create.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip, _dp]
        L.ORACLE_CSR_Create.restype = C.c_void_p
        L.ORACLE_CSR_Free.argtypes = [C.c_void_p]
        L.ORACLE_CSR_ReadFile.argtypes = [C.c_char_p]
        L.ORACLE_CSR_ReadFile.restype = C.c_void_p
        L.ORACLE_CSR_WriteFile.argtypes = [C.c_void_p, C.c_char_p]
        L.ORACLE_CSR_Nrows.argtypes = [C.c_void_p]
        L.ORACLE_CSR_Nnz.argtypes = [C.c_void_p]
        L.ORACLE_CSR_Export.argtypes = [C.c_void_p, _ip, _ip, _dp]
        L.ORACLE_SetSeed.argtypes = [C.c_uint64]
        L.ORACLE_Backend.restype = C.c_void_p
        L.ORACLE_RandomValue.argtypes = [C.c_uint64] * 4
        L.ORACLE_RandomValue.restype = C.c_double
        L.ORACLE_SymEig.argtypes = [C.c_int, _dp, C.c_int, _dp, _dp, C.c_int]
        L.ORACLE_SpMM.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int, _dp, C.c_int]
        L.ORACLE_Gram.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp,
                                  C.c_int]
        L.ORACLE_LinearComb.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_int, _dp, C.c_int,
                                        C.c_double, _dp, C.c_int]
        L.ORACLE_RandomFill.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.c_int, _dp,
                                        C.c_int]
        L.ORACLE_BlockCG.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_void_p, C.c_int, _dp,
                                     C.c_int, _dp, C.c_int, C.c_int, C.c_double]
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


class Matrix:
    """An oracle CSR matrix (owns a native copy)."""

    def __init__(self, csr):
        self.csr = csr
        self.rowptr = np.ascontiguousarray(csr.rowptr, dtype=np.int32)
        self.colidx = np.ascontiguousarray(csr.colidx, dtype=np.int32)
        self.vals = np.ascontiguousarray(csr.vals, dtype=np.float64)
        self.handle = lib().ORACLE_CSR_Create(csr.nrows, csr.ncols, csr.nnz, _i(self.rowptr),
                                              _i(self.colidx), _d(self.vals))

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.ORACLE_CSR_Free(self.handle)
            self.handle = None


def spmm(A: Matrix, X: np.ndarray) -> np.ndarray:
    X = np.asfortranarray(X, dtype=np.float64)
    n, k = X.shape
    Y = np.zeros((A.csr.nrows, k), order="F")
    lib().ORACLE_SpMM(A.handle, k, _d(X), n, _d(Y), A.csr.nrows)
    return Y


def gram(V: np.ndarray, W: np.ndarray) -> np.ndarray:
    V = np.asfortranarray(V, dtype=np.float64)
    W = np.asfortranarray(W, dtype=np.float64)
    n, m = V.shape
    k = W.shape[1]
    G = np.zeros((m, k), order="F")
    lib().ORACLE_Gram(n, m, k, _d(V), n, _d(W), n, _d(G), m)
    return G


def linear_comb(V: np.ndarray, Cm: np.ndarray, beta: float = 0.0, Y: np.ndarray | None = None):
    V = np.asfortranarray(V, dtype=np.float64)
    Cm = np.asfortranarray(Cm, dtype=np.float64)
    n, m = V.shape
    k = Cm.shape[1]
    Y = np.zeros((n, k), order="F") if Y is None else np.asfortranarray(Y.copy(), dtype=np.float64)
    lib().ORACLE_LinearComb(n, m, k, _d(V), n, _d(Cm), m, beta, _d(Y), n)
    return Y


def sym_eig(a: np.ndarray):
    a = np.asfortranarray(a, dtype=np.float64)
    n = a.shape[0]
    w = np.zeros(n)
    z = np.zeros((n, n), order="F")
    lib().ORACLE_SymEig(n, _d(a), n, _d(w), _d(z), n)
    return w, z


def random_fill(seed: int, draw: int, col0: int, ncols: int, n: int) -> np.ndarray:
    X = np.zeros((n, ncols), order="F")
    lib().ORACLE_RandomFill(seed, draw, col0, ncols, n, _d(X), n)
    return X


def block_cg(A, B, rhs: np.ndarray, x0: np.ndarray, max_it: int = 20, rate: float = 1e-2,
             if_shift: int = 0, shift: float = 0.0):
    """The reference block CG (src/gcge_cg.c:302-501) restated in C: returns
    (x, r) -- the solutions and the final residual block that the
    reference leaves in its right-hand-side block.  A, B: problems.CSR."""
    Am = Matrix(A)
    Bm = Matrix(B) if B is not None else None
    r = np.asfortranarray(rhs, dtype=np.float64).copy(order="F")
    x = np.asfortranarray(x0, dtype=np.float64).copy(order="F")
    n, k = r.shape
    lib().ORACLE_BlockCG(Am.handle, if_shift, shift, None if Bm is None else Bm.handle, k, _d(r),
                         n, _d(x), n, max_it, rate)
    return x, r


def solve(A, B, nev: int, flags: list[str] | None = None, x0: np.ndarray | None = None,
          seed: int = 1, history_cap: int = 0):
    """Run the solver core on the oracle CPU operations (column loops, serial
    CSR).  Returns a dict like gcge_old_amd.solve."""
    from gcge_old_amd import _native

    Am = A if isinstance(A, Matrix) else Matrix(A)
    Bm = None if B is None else (B if isinstance(B, Matrix) else Matrix(B))
    n = Am.csr.nrows
    lib().ORACLE_SetSeed(seed)
    return _native_solve(_native, lib().ORACLE_Backend(), Am.handle,
                         None if Bm is None else Bm.handle, n, nev, flags or [], x0, history_cap,
                         keep=(Am, Bm))


def _native_solve(_native, backend, Ah, Bh, n, nev, flags, x0, history_cap, keep=None):
    core = _native.core()
    argv = _native.argv_array(["gcge"] + list(flags))
    ev = np.zeros(nev)
    vec = np.zeros((n, nev), order="F")
    res = np.zeros(nev)
    info = np.zeros(4)
    hist = np.zeros((max(history_cap, 1), nev + 2))
    x0p = None
    if x0 is not None:
        x0 = np.asfortranarray(x0, dtype=np.float64)
        if x0.shape != (n, nev):
            raise ValueError(f"x0 must have shape ({n}, {nev}), got {x0.shape}")
        x0p = _d(x0)
    rc = core.GCGE_SolveWithBackend(backend, Ah, Bh, n, nev, len(flags) + 1, argv, x0p, _d(ev),
                                    _d(vec), _d(res), _d(info), _d(hist) if history_cap else None,
                                    history_cap)
    if rc == 2:
        raise ValueError("a -gcge_nev flag must equal the nev argument")
    if rc != 0:
        raise RuntimeError("operations table incomplete")
    h = hist[: int(np.count_nonzero(hist[:, 0]))] if history_cap else None
    return dict(eval=ev, evec=vec, res=res, num_iter=int(info[0]), num_conv=int(info[1]),
                time=float(info[2]), block_size=int(info[3]), history=h)


REF_HIP_BIN = os.path.join(HERE, "_ref", "gcge_ref_hip")
REF_BCG_BIN = os.path.join(HERE, "_ref", "gcge_ref_bcg")


def ref_block_cg(A, B, rhs: np.ndarray, x0: np.ndarray, max_it: int = 20, rate: float = 1e-2,
                 if_shift: int = 0, shift: float = 0.0):
    """The reference GCGE_BCG built from its sources (oracle/_ref/gcge_ref_bcg)
    on its CSR operations; returns (x, r) like block_cg."""
    if not os.path.exists(REF_BCG_BIN):
        raise RuntimeError(f"{REF_BCG_BIN} is not built (make -C oracle ref)")
    n, k = rhs.shape
    with tempfile.TemporaryDirectory() as d:
        a_path = os.path.join(d, "A.bin")
        A.write_bin(a_path)
        b_path = "-"
        if B is not None:
            b_path = os.path.join(d, "B.bin")
            B.write_bin(b_path)
        r_path, x_path, out = (os.path.join(d, f) for f in ("rhs.bin", "x0.bin", "out.bin"))
        np.asfortranarray(rhs, dtype=np.float64).T.tofile(r_path)
        np.asfortranarray(x0, dtype=np.float64).T.tofile(x_path)
        args = [REF_BCG_BIN, a_path, b_path, str(k), str(if_shift), repr(float(shift)),
                str(max_it), repr(float(rate)), r_path, x_path, out]
        p = subprocess.run(args, capture_output=True, text=True, timeout=600)
        if p.returncode != 0:
            raise RuntimeError(f"reference block CG failed: {p.stderr[-2000:]}")
        data = np.fromfile(out, dtype=np.float64)
    x = data[: n * k].reshape(k, n).T.copy(order="F")
    r = data[n * k:].reshape(k, n).T.copy(order="F")
    return x, r


def ref_available() -> bool:
    return os.path.exists(REF_BIN)


def ref_hip_available() -> bool:
    return os.path.exists(REF_HIP_BIN)


def ref_solve(A, B, nev: int, flags: list[str] | None = None, x0: np.ndarray | None = None,
              print_eval: bool = False, timeout: float = 3600.0, on_device: bool = False):
    """Run the reference solver (built from its sources) in a subprocess, on
    its CSR backend or (on_device) on the device backend through
    GCGE_HIP_SetOps.  Returns a dict with eval, res, evec, num_iter,
    num_conv, time and, when print_eval, the per-check eigenvalue history
    parsed from its output."""
    binary = REF_HIP_BIN if on_device else REF_BIN
    if not os.path.exists(binary):
        raise RuntimeError(f"{binary} is not built (make -C oracle ref ref-hip)")
    with tempfile.TemporaryDirectory() as d:
        a_path = os.path.join(d, "A.bin")
        A.write_bin(a_path)
        b_path = "-"
        if B is not None:
            b_path = os.path.join(d, "B.bin")
            B.write_bin(b_path)
        x_path = "-"
        if x0 is not None:
            x_path = os.path.join(d, "x0.bin")
            np.asfortranarray(x0, dtype=np.float64).T.tofile(x_path)
        out = os.path.join(d, "out.bin")
        args = [binary, a_path, b_path, str(nev), x_path, out] + list(flags or [])
        args += ["-gcge_print_para", "0", "-gcge_print_final_part_time", "0",
                 "-gcge_print_eval", "1" if print_eval else "0"]
        p = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
        if p.returncode != 0:
            raise RuntimeError(f"reference run failed: {p.stderr[-2000:]}")
        with open(out, "rb") as f:
            hdr = np.fromfile(f, dtype=np.int32, count=4)
            nev_o, num_iter, num_conv, n = (int(v) for v in hdr)
            secs = float(np.fromfile(f, dtype=np.float64, count=1)[0])
            ev = np.fromfile(f, dtype=np.float64, count=nev_o)
            res = np.fromfile(f, dtype=np.float64, count=nev_o)
            vec = np.fromfile(f, dtype=np.float64, count=n * nev_o).reshape(nev_o, n).T
    hist = None
    if print_eval:
        rows: dict[int, dict[int, float]] = {}
        for m in re.finditer(r"eval\((\d+),\s*(\d+)\) = ([-+0-9.eE]+)", p.stdout):
            rows.setdefault(int(m.group(1)), {})[int(m.group(2)) - 1] = float(m.group(3))
        conv = {int(m.group(1)): int(m.group(2))
                for m in re.finditer(r"num_conv_continuous\(\s*(\d+)\) =\s*(\d+)", p.stdout)}
        hist = np.array([[it, conv.get(it, -1)] + [rows[it][i] for i in range(nev_o)]
                         for it in sorted(rows)])
    return dict(eval=ev, res=res, evec=np.asfortranarray(vec), num_iter=num_iter,
                num_conv=num_conv, time=secs, history=hist, stdout=p.stdout)
