"""ctypes bindings of the native libraries.

``libgcge_amd.so``  the GCG core (gcge.h): solver facade, parameters and the
                    one-call driver GCGE_SolveWithBackend.
``libgcge_hip.so``  the MI355X backend (gcge_hip.h): device CSR matrices and
                    multivecs, block kernels, the operations table.

Both are built in-tree by ``build()`` in ``__graft_entry__`` (or
``python -m gcge_old_amd.build``).  Loading fails loudly when a library is
missing: there is no pure-Python or CPU stand-in for the device path.
"""
from __future__ import annotations

import ctypes as C
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")

_libs: dict[str, C.CDLL] = {}


def lib_path(name: str) -> str:
    return os.path.join(LIB_DIR, name)


def load(name: str) -> C.CDLL:
    if name not in _libs:
        path = lib_path(name)
        if not os.path.exists(path):
            raise RuntimeError(
                f"{path} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
                "from the repository root")
        # local binding: published globally, the ROCm libraries this library
        # links (/opt/rocm) would capture the symbols of the ROCm copies a
        # later `import torch` loads, and that import would crash
        _libs[name] = C.CDLL(path, mode=C.RTLD_LOCAL)
        _declare(name, _libs[name])
    return _libs[name]


def core() -> C.CDLL:
    return load("libgcge_amd.so")


def hip() -> C.CDLL:
    return load("libgcge_hip.so")


dptr = C.POINTER(C.c_double)
iptr = C.POINTER(C.c_int)


def _declare(name: str, lib: C.CDLL) -> None:
    if name == "libgcge_amd.so":
        lib.GCGE_SolveWithBackend.argtypes = [
            C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
            C.POINTER(C.c_char_p), dptr, dptr, dptr, dptr, dptr, dptr, C.c_int]
        lib.GCGE_SolveWithBackend.restype = C.c_int
    elif name == "libgcge_hip.so":
        lib.GCGE_HIP_Init.argtypes = [C.c_int]
        lib.GCGE_HIP_Init.restype = C.c_int
        lib.GCGE_HIP_Finalize.argtypes = []
        lib.GCGE_HIP_LastError.argtypes = []
        lib.GCGE_HIP_LastError.restype = C.c_char_p
        lib.GCGE_HIP_Backend.argtypes = []
        lib.GCGE_HIP_Backend.restype = C.c_void_p
        lib.GCGE_HIP_SetSeed.argtypes = [C.c_uint64]
        lib.GCGE_HIP_MatCreateCSR.argtypes = [C.c_int, C.c_int, C.c_int, iptr, iptr, dptr]
        lib.GCGE_HIP_MatCreateCSR.restype = C.c_void_p
        lib.GCGE_HIP_MatCreateQ1.argtypes = [C.c_int, dptr, C.c_int, iptr]
        lib.GCGE_HIP_MatCreateQ1.restype = C.c_void_p
        lib.GCGE_HIP_MatGetCSR.argtypes = [C.c_void_p, iptr, iptr, dptr]
        lib.GCGE_HIP_MatGetCSR.restype = C.c_int
        lib.GCGE_HIP_MatDestroy.argtypes = [C.c_void_p]
        lib.GCGE_HIP_MatSetPrecision.argtypes = [C.c_void_p, C.c_int]
        lib.GCGE_HIP_MatSetStorage.argtypes = [C.c_void_p, C.c_int]
        lib.GCGE_HIP_MatSetStorage.restype = C.c_int
        lib.GCGE_HIP_MatStorageInfo.argtypes = [C.c_void_p, iptr, iptr]
        lib.GCGE_HIP_MatStorageInfo.restype = C.c_int
        lib.GCGE_HIP_MatDiagonalPlane.argtypes = [C.c_void_p]
        lib.GCGE_HIP_MatDiagonalPlane.restype = C.c_int
        lib.GCGE_HIP_MultiVecCreate.argtypes = [C.c_int, C.c_int]
        lib.GCGE_HIP_MultiVecCreate.restype = C.c_void_p
        lib.GCGE_HIP_MultiVecDestroy.argtypes = [C.c_void_p]
        lib.GCGE_HIP_MultiVecSetHost.argtypes = [C.c_void_p, C.c_int, C.c_int, dptr, C.c_int]
        lib.GCGE_HIP_MultiVecGetHost.argtypes = [C.c_void_p, C.c_int, C.c_int, dptr, C.c_int]
        lib.GCGE_HIP_SpMM.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                      C.c_int]
        lib.GCGE_HIP_Gram.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int,
                                      dptr, C.c_int]
        lib.GCGE_HIP_LinearComb.argtypes = [C.c_void_p, C.c_int, C.c_int, dptr, C.c_int,
                                            C.c_double, C.c_void_p, C.c_int, C.c_int]
        lib.GCGE_HIP_Axpby.argtypes = [C.c_double, C.c_void_p, C.c_int, C.c_double, C.c_void_p,
                                       C.c_int, C.c_int]
        lib.GCGE_HIP_RandomFill.argtypes = [C.c_void_p, C.c_int, C.c_int]
        lib.GCGE_HIP_SetLinearSolverParams.argtypes = [C.c_int, C.c_double, C.c_int, C.c_double,
                                                       C.c_void_p]
        lib.GCGE_HIP_GramNoAlloc.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                             C.c_int, dptr, C.c_int]
        lib.GCGE_HIP_GramNoAlloc.restype = C.c_int
        lib.GCGE_HIP_WorkCounts.argtypes = [dptr, C.c_int, C.c_int]
        lib.GCGE_HIP_WorkCounts.restype = C.c_int
        lib.GCGE_HIP_GetLinearSolverParams.argtypes = [iptr, dptr, iptr, dptr,
                                                       C.POINTER(C.c_void_p)]
        lib.GCGE_HIP_MultiLinearSolve.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                                  C.c_int, C.c_int]
        lib.GCGE_HIP_PatternGradient.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                                 C.c_int, C.c_int, dptr, dptr]
        lib.GCGE_HIP_Q1ElementGradient.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                                   C.c_int, C.c_int, dptr, dptr]
        lib.GCGE_HIP_Q1ElementGradientRows.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                                       C.c_int, C.c_int, dptr, iptr, dptr]
        lib.GCGE_HIP_Synchronize.argtypes = []
        lib.GCGE_HIP_KernelTime.argtypes = [C.c_char_p, dptr, iptr]
        lib.GCGE_HIP_KernelTime.restype = C.c_int
        lib.GCGE_HIP_KernelStats.argtypes = [C.c_char_p, dptr, iptr, dptr]
        lib.GCGE_HIP_KernelStats.restype = C.c_int
        lib.GCGE_HIP_ProfileEnable.argtypes = [C.c_int]
        lib.GCGE_HIP_ProfileReset.argtypes = []
        lib.GCGE_HIP_ProfileReserve.argtypes = [C.c_int]


def as_dptr(a):
    return a.ctypes.data_as(dptr)


def as_iptr(a):
    return a.ctypes.data_as(iptr)


def argv_array(flags: list[str]):
    arr = (C.c_char_p * (len(flags) + 1))()
    for i, f in enumerate(flags):
        arr[i] = f.encode()
    arr[len(flags)] = None
    return arr
