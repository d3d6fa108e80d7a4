"""The C-ABI libraries load and export every function their headers declare
(CPU; no device call is made)."""
import ctypes as C
import os
import re
import sys

import numpy as np
import pytest

from gcge_old_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = {
    "gcge_hip.h": "libgcge_hip.so",
    "gcge.h": "libgcge_amd.so",
}
DECL = re.compile(r"^[A-Za-z_][\w \*]*?\b(GCGE_\w+)\s*\(", re.M)


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)  # drop comments
    text = re.sub(r"typedef struct .*?\} \w+;", "", text, flags=re.S)  # drop struct bodies
    names = set(DECL.findall(text))
    return sorted(n for n in names if not n.isupper())


@pytest.mark.parametrize("header", sorted(HEADERS))
def test_every_declared_symbol_is_exported(header):
    names = declared(header)
    assert len(names) > 10
    lib = C.CDLL(_native.lib_path(HEADERS[header]))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"{HEADERS[header]} lacks {missing}"


def test_c_example_is_built_on_the_public_headers():
    """build() compiles gcge_old_amd/csrc/examples/q1_solve.c against
    include/ only (gcge_hip_solver.h is header-only) and links both
    libraries; the GPU suite runs it."""
    exe = _native.lib_path("q1_solve")
    assert os.path.exists(exe) and os.access(exe, os.X_OK)
    src = open(os.path.join(ROOT, "gcge_old_amd", "csrc", "examples", "q1_solve.c")).read()
    assert re.findall(r'#include "([^"]+)"', src) == ["gcge_hip_solver.h"]


def test_ops_table_layout_matches_reference_order():
    """The table members appear in the reference order (src/gcge_ops.h:29-150),
    so a driver built against the reference header sees the same offsets."""
    text = open(os.path.join(ROOT, "include", "gcge_ops.h")).read()
    members = re.findall(r"\(\*(\w+)\)\s*\(", text)
    expected = [
        "VecSetRandomValue", "MatDotVec", "MatTransposeDotVec", "VecAxpby", "VecInnerProd",
        "VecLocalInnerProd", "VecCreateByVec", "VecCreateByMat", "VecDestroy",
        "GetVecFromMultiVec", "RestoreVecForMultiVec", "MultiGridCreate", "MultiGridDestroy",
        "VecFromItoJ", "MultiVecFromItoJ", "LinearSolver", "MultiLinearSolver",
        "Orthonormalization", "DenseMatCreate", "DenseMatDestroy", "MultiVecCreateByVec",
        "MultiVecCreateByMat", "MultiVecCreateByMultiVec", "MultiVecDestroy",
        "MultiVecSetRandomValue", "MatDotMultiVec", "MatTransposeDotMultiVec", "MultiVecAxpby",
        "MultiVecAxpbyColumn", "MultiVecLinearComb", "MultiVecInnerProd",
        "MultiVecInnerProdLocal", "MultiVecSwap", "MultiVecPrint", "SetDirichletBoundary",
        "DenseMatEigenSolver", "DenseMatDotDenseMat", "DenseSymMatDotDenseMat", "ArrayDotArray",
        "ArrayNorm", "ArrayAXPBY", "ArrayCopy", "ArrayScale",
    ]
    assert members == expected
    assert "void *ops;" in text and "void *linear_solver_workspace;" in text


def test_core_parameters_and_block_size_rules():
    """Defaults and the block-size rule of the reference (src/gcge_para.c:25-113,
    :420-429) through the one-call driver on the oracle backend."""
    import numpy as np

    import oracle
    from gcge_old_amd import problems

    A = problems.fd_laplacian_2d(12)
    for nev, flags, bs in [(12, [], 10), (5, [], 5), (12, ["-gcge_block_size", "40"], 12),
                           (60, [], 12), (30, ["-gcge_block_size", "7"], 7)]:
        out = oracle.solve(A, None, nev, flags + ["-gcge_ev_max_it", "1"])
        assert out["block_size"] == bs
    # -gcge_cg_rate is read with atoi: "0.5" becomes 0 (CG runs to cg_max_it)
    out = oracle.solve(A, None, 12, ["-gcge_cg_rate", "0.5", "-gcge_ev_max_it", "3"])
    assert np.all(np.isfinite(out["eval"]))


def test_nev_flag_must_match_the_nev_argument():
    """The output arrays are sized by the nev argument: a -gcge_nev flag with
    another value is refused before any solve (it used to overwrite nev and
    write past the caller's arrays); the same value is accepted.  x0 must be
    n x nev."""
    import numpy as np

    import oracle
    from gcge_old_amd import problems

    A = problems.fd_laplacian_2d(12)
    with pytest.raises(ValueError, match="gcge_nev"):
        oracle.solve(A, None, 2, ["-gcge_nev", "40"])
    out = oracle.solve(A, None, 4, ["-gcge_nev", "4", "-gcge_ev_max_it", "2"])
    assert out["eval"].shape == (4,)
    with pytest.raises(ValueError, match="x0 must have shape"):
        oracle.solve(A, None, 4, ["-gcge_ev_max_it", "1"], x0=np.zeros((A.nrows, 3)))


def test_torch_import_after_the_libraries():
    """Loading both libraries and then importing torch in one process must
    work (the libraries are opened with local symbol binding; opened
    globally, torch's own ROCm copies bind to /opt/rocm's and the import
    crashes).  Runs in a subprocess; needs no GPU."""
    import subprocess
    import sys

    code = ("import gcge_old_amd._native as n; n.core(); n.hip(); import torch; "
            "print('ok', torch.__version__)")
    p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=600,
                       cwd=ROOT)
    assert p.returncode == 0 and p.stdout.startswith("ok"), (p.returncode, p.stderr[-2000:])


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU: --gpus 2 exits with status 2 and says so,
    before any device work (CPU)."""
    import subprocess
    import sys

    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and "--gpus 2" in p.stderr
    assert p.stdout == ""


def test_bench_cluster_cut_for_subspace_angles():
    """bench.py compares the two precisions' eigenvector subspaces only below
    the last eigenvalue cluster (a cluster that may continue beyond nev has
    no unique basis): the kappa = 1 cube's 50 smallest eigenvalues end in a
    pair, so the cut is at 48; a simple spectrum loses its last value only."""
    sys.path.insert(0, ROOT)
    import bench
    from gcge_old_amd import problems

    assert bench.leading_complete_clusters(problems.q1_eigs(64, 50)) == 48
    assert bench.leading_complete_clusters(np.array([1.0, 2.0, 3.0, 4.0])) == 3
    assert bench.leading_complete_clusters(np.array([1.0, 2.0, 3.0, 3.0, 3.0 + 1e-9])) == 2
