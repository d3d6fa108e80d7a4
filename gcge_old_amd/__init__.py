"""gcge_old_amd -- a GCG eigensolver backend for AMD Instinct MI355X.

Python mirror of the solver API.  The work happens in two native libraries
(see ``_native``): the GCG core in C (``libgcge_amd.so``) and the HIP
backend (``libgcge_hip.so``).  This module only moves data in and out:

    import gcge_old_amd as gcge
    from gcge_old_amd import problems
    K, M = problems.q1_fem_3d(64)
    out = gcge.solve(K, M, nev=50, flags=["-gcge_ev_tol", "1e-8"])
    out["eval"], out["evec"], out["num_iter"], out["time"]

``flags`` are the reference "-gcge_*" command-line flags
(reference src/gcge_para.c:128-358), parsed by the native core.
"""
from __future__ import annotations

import numpy as np

from . import _native, problems  # noqa: F401

__all__ = ["init", "DeviceMatrix", "DeviceMultiVec", "solve", "problems"]

_initialised = False


def init(device: int = 0) -> None:
    """Select the GPU and create the backend stream.  Raises when no usable
    GPU is present: there is no CPU fallback for the device path."""
    global _initialised
    if _initialised:
        return
    lib = _native.hip()
    if lib.GCGE_HIP_Init(device) != 0:
        raise RuntimeError("GCGE_HIP_Init failed: " + lib.GCGE_HIP_LastError().decode())
    _initialised = True


class DeviceMatrix:
    """A CSR matrix resident in device memory (sliced ELLPACK layout, plus
    diagonal storage when its entries lie on few diagonals)."""

    def __init__(self, csr: problems.CSR):
        init()
        self.csr = csr
        self.nrows, self.ncols, self.nnz = csr.nrows, csr.ncols, csr.nnz
        rp = np.ascontiguousarray(csr.rowptr, dtype=np.int32)
        ci = np.ascontiguousarray(csr.colidx, dtype=np.int32)
        v = np.ascontiguousarray(csr.vals, dtype=np.float64)
        self.handle = _native.hip().GCGE_HIP_MatCreateCSR(csr.nrows, csr.ncols, csr.nnz,
                                                          _native.as_iptr(rp), _native.as_iptr(ci),
                                                          _native.as_dptr(v))
        if not self.handle:
            raise RuntimeError("GCGE_HIP_MatCreateCSR failed")
        self.fp32 = False

    def set_precision(self, fp32: bool) -> None:
        """fp32=True: keep a float32 copy of the values and run the inner
        block CG products in float32 (mixed-precision form)."""
        _native.hip().GCGE_HIP_MatSetPrecision(self.handle, 1 if fp32 else 0)
        self.fp32 = bool(fp32)

    STORAGES = ("sliced", "diagonal", "rows")

    def set_storage(self, storage) -> bool:
        """Products through "diagonal" storage (when the matrix has it),
        "sliced" ELLPACK with element gathers, or sliced ELLPACK with "rows"
        gathers from a row-major copy of the column block (from 16 columns);
        True/False mean "diagonal"/"sliced".  Returns whether diagonal storage
        is now used."""
        if storage is True or storage is False:
            storage = "diagonal" if storage else "sliced"
        used = _native.hip().GCGE_HIP_MatSetStorage(self.handle, self.STORAGES.index(storage))
        return used == 1

    def storage_info(self) -> dict:
        nd, ng = np.zeros(1, dtype=np.int32), np.zeros(1, dtype=np.int32)
        used = _native.hip().GCGE_HIP_MatStorageInfo(self.handle, _native.as_iptr(nd),
                                                     _native.as_iptr(ng))
        return {"diagonal": used == 1, "storage": self.STORAGES[used], "ndiag": int(nd[0]),
                "runs": int(ng[0])}

    def __del__(self):
        if getattr(self, "handle", None):
            _native.hip().GCGE_HIP_MatDestroy(self.handle)
            self.handle = None


class DeviceMultiVec:
    """n x ncols float64 column block in device memory."""

    def __init__(self, n: int, ncols: int):
        init()
        self.n, self.ncols = n, ncols
        self.handle = _native.hip().GCGE_HIP_MultiVecCreate(n, ncols)

    @classmethod
    def from_host(cls, X: np.ndarray) -> "DeviceMultiVec":
        X = np.asfortranarray(X, dtype=np.float64)
        mv = cls(X.shape[0], X.shape[1])
        mv.set(X)
        return mv

    def set(self, X: np.ndarray, start: int = 0) -> None:
        X = np.asfortranarray(X, dtype=np.float64)
        _native.hip().GCGE_HIP_MultiVecSetHost(self.handle, start, X.shape[1], _native.as_dptr(X),
                                               X.shape[0])

    def get(self, start: int = 0, ncols: int | None = None) -> np.ndarray:
        ncols = self.ncols - start if ncols is None else ncols
        X = np.zeros((self.n, ncols), order="F")
        _native.hip().GCGE_HIP_MultiVecGetHost(self.handle, start, ncols, _native.as_dptr(X),
                                               self.n)
        return X

    def __del__(self):
        if getattr(self, "handle", None):
            _native.hip().GCGE_HIP_MultiVecDestroy(self.handle)
            self.handle = None


def solve(A, B, nev: int, flags: list[str] | None = None, x0: np.ndarray | None = None,
          seed: int = 1, history_cap: int = 0, precision: str = "fp64") -> dict:
    """Smallest nev eigenpairs of A x = lambda B x on the GPU (B may be None).

    A and B are ``problems.CSR`` or ``DeviceMatrix``.  ``precision`` is
    "fp64" (float64 throughout) or "mixed" (the products inside the inner
    block CG in float32; Gram matrices, orthonormalisation, Rayleigh-Ritz and
    the residual check stay float64).  Returns a dict with eval, evec
    (n x nev), res (last residuals), num_iter, num_conv, time (device solve
    seconds, inputs resident), block_size and history."""
    if precision not in ("fp64", "mixed"):
        raise ValueError(f"precision must be 'fp64' or 'mixed', not {precision!r}")
    init()
    Ad = A if isinstance(A, DeviceMatrix) else DeviceMatrix(A)
    Bd = None if B is None else (B if isinstance(B, DeviceMatrix) else DeviceMatrix(B))
    for m in (Ad, Bd):
        if m is not None:
            m.set_precision(precision == "mixed")
    lib = _native.hip()
    lib.GCGE_HIP_SetSeed(seed)
    flags = list(flags or [])
    n = Ad.nrows
    core = _native.core()
    argv = _native.argv_array(["gcge"] + flags)
    ev = np.zeros(nev)
    vec = np.zeros((n, nev), order="F")
    res = np.zeros(nev)
    info = np.zeros(4)
    hist = np.zeros((max(history_cap, 1), nev + 2))
    x0p = None
    if x0 is not None:
        x0 = np.asfortranarray(x0, dtype=np.float64)
        x0p = _native.as_dptr(x0)
    rc = core.GCGE_SolveWithBackend(lib.GCGE_HIP_Backend(), Ad.handle,
                                    None if Bd is None else Bd.handle, n, nev, len(flags) + 1,
                                    argv, x0p, _native.as_dptr(ev), _native.as_dptr(vec),
                                    _native.as_dptr(res), _native.as_dptr(info),
                                    _native.as_dptr(hist) if history_cap else None, history_cap)
    if rc != 0:
        raise RuntimeError("operations table incomplete")
    h = hist[: int(np.count_nonzero(hist[:, 0]))] if history_cap else None
    return dict(eval=ev, evec=vec, res=res, num_iter=int(info[0]), num_conv=int(info[1]),
                time=float(info[2]), block_size=int(info[3]), history=h)
