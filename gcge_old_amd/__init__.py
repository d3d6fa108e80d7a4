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

    @classmethod
    def q1(cls, E: int, kappa=None, which: str = "stiffness", node_row=None) -> "DeviceMatrix":
        """Q1 stiffness (``which="stiffness"``, coefficients ``kappa``: E^3
        element values, None for 1) or mass (``"mass"``) matrix of the E^3
        mesh assembled on the device, the same matrix bit for bit as
        ``problems.q1_fem_3d`` (rows and columns renumbered by ``node_row``:
        row ``node_row[i]`` holds node i, rows sorted by column)."""
        init()
        if which not in ("stiffness", "mass"):
            raise ValueError("which must be 'stiffness' or 'mass'")
        w = 0 if which == "stiffness" else 1
        kp = None
        if kappa is not None and w == 0:
            kp = np.ascontiguousarray(kappa, dtype=np.float64).ravel()
            if kp.size != E ** 3:
                raise ValueError(f"kappa needs {E ** 3} element values, got {kp.size}")
        nr = None
        if node_row is not None:
            nr = np.ascontiguousarray(node_row, dtype=np.int32).ravel()
            n = (E - 1) ** 3
            if nr.size != n:
                raise ValueError(f"node_row needs {n} entries, got {nr.size}")
            if n and (nr.min() < 0 or nr.max() >= n or
                      not np.all(np.bincount(nr, minlength=n) == 1)):
                raise ValueError("node_row must be a permutation of 0..n-1")
        self = cls.__new__(cls)
        self.csr = None
        self.handle = _native.hip().GCGE_HIP_MatCreateQ1(
            E, _native.as_dptr(kp) if kp is not None else None, w,
            _native.as_iptr(nr) if nr is not None else None)
        if not self.handle:
            raise RuntimeError("GCGE_HIP_MatCreateQ1 failed")
        n = (E - 1) ** 3
        self.nrows = self.ncols = n
        self.nnz = int(_native.hip().GCGE_HIP_MatGetCSR(self.handle, None, None, None))
        self.fp32 = False
        return self

    def to_csr(self) -> problems.CSR:
        """The matrix copied back from the device in CSR form."""
        rp = np.zeros(self.nrows + 1, dtype=np.int32)
        ci = np.zeros(self.nnz, dtype=np.int32)
        v = np.zeros(self.nnz, dtype=np.float64)
        _native.hip().GCGE_HIP_MatGetCSR(self.handle, _native.as_iptr(rp), _native.as_iptr(ci),
                                        _native.as_dptr(v))
        return problems.CSR(self.nrows, self.ncols, rp, ci, v)

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
        True/False mean "diagonal"/"sliced".  "diagonal-consecutive" and
        "diagonal-planes" fix the order of the diagonal row blocks,
        "diagonal-tiles" uses the LDS plane tiles of 27-point plane layouts
        for every column count (same products bit for bit; "diagonal"
        chooses: plane tiles from 8 columns).  Returns whether diagonal
        storage is now used."""
        if storage is True or storage is False:
            storage = "diagonal" if storage else "sliced"
        code = {"diagonal-consecutive": 3, "diagonal-planes": 4, "diagonal-tiles": 5}.get(storage)
        if code is None:
            code = self.STORAGES.index(storage)
        used = _native.hip().GCGE_HIP_MatSetStorage(self.handle, code)
        return used == 1

    def storage_info(self) -> dict:
        nd, ng = np.zeros(1, dtype=np.int32), np.zeros(1, dtype=np.int32)
        used = _native.hip().GCGE_HIP_MatStorageInfo(self.handle, _native.as_iptr(nd),
                                                     _native.as_iptr(ng))
        return {"diagonal": used == 1, "storage": self.STORAGES[used], "ndiag": int(nd[0]),
                "runs": int(ng[0]),
                "plane_rows": int(_native.hip().GCGE_HIP_MatDiagonalPlane(self.handle))}

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
        if x0.shape != (n, nev):
            raise ValueError(f"x0 must have shape ({n}, {nev}), got {x0.shape}")
        x0p = _native.as_dptr(x0)
    rc = core.GCGE_SolveWithBackend(lib.GCGE_HIP_Backend(), Ad.handle,
                                    None if Bd is None else Bd.handle, n, nev, len(flags) + 1,
                                    argv, x0p, _native.as_dptr(ev), _native.as_dptr(vec),
                                    _native.as_dptr(res), _native.as_dptr(info),
                                    _native.as_dptr(hist) if history_cap else None, history_cap)
    if rc == 2:
        raise ValueError("a -gcge_nev flag must equal the nev argument")
    if rc != 0:
        raise RuntimeError("operations table incomplete")
    h = hist[: int(np.count_nonzero(hist[:, 0]))] if history_cap else None
    return dict(eval=ev, evec=vec, res=res, num_iter=int(info[0]), num_conv=int(info[1]),
                time=float(info[2]), block_size=int(info[3]), history=h)
