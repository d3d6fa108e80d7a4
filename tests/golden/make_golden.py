"""Writes the committed fixtures under tests/golden/ (run where the
reference tree is present; the GPU tests only read the results).

1. data/*.npz -- the CSR matrices the reference ships as test inputs
   (test/data/testA, testB, A_3.txt, M_3.txt, A_5.txt, M_5.txt), converted
   from the reference text format to npz.
   data/A_5.petsc.bin, M_5.petsc.bin -- the same matrices as the
   reference ships them in PETSc's binary format (copied unchanged).
2. ref_runs.npz -- outputs of the reference solver built from its own
   sources (oracle/_ref/gcge_ref) on small problems, each started from a
   fixed initial block: final eigenvalues, residuals, iteration counts and
   the per-iteration convergence counts.  These pin the GCG core's control
   flow (soft locking, unlock lists, block CG, P construction) without the
   reference tree at test time.  Runs named *_dirichlet start from the
   block with the Dirichlet rows zeroed as CSR_SetDirichletBoundary
   (app/csr/gcge_app_csr.c:130-183) zeroes them: the reference applies it
   only to randomly drawn blocks (src/gcge_eigsol.c:134-185), so the zeroed
   block is what a solve with dirichlet_boundary = 1 and this random block
   orthonormalises.

    python tests/golden/make_golden.py [/root/reference]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from gcge_old_amd import problems  # noqa: E402

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
DATA = ["testA", "testB", "A_3.txt", "M_3.txt", "A_5.txt", "M_5.txt"]

# (name, matrix builder, nev, flags, x0 seed)
RUNS = [
    ("A5M5_nev20", ("A_5", "M_5"), 20, ["-gcge_block_size", "10", "-gcge_ev_max_it", "40"], 7),
    ("A5M5_nev40", ("A_5", "M_5"), 40, ["-gcge_ev_max_it", "60"], 7),
    ("A3M3_nev8", ("A_3", "M_3"), 8, ["-gcge_ev_max_it", "60"], 5),
    ("A5M5_nev9", ("A_5", "M_5"), 9, [], 7),
    ("fd2d_24_nev12", "fd24", 12, ["-gcge_ev_max_it", "60"], 9),
    ("q1_e10_nev15", "q1e10", 15, ["-gcge_ev_max_it", "60", "-gcge_ev_tol", "1e-10"], 3),
    ("A5M5_nev20_dirichlet", ("A_5", "M_5"), 20, ["-gcge_block_size", "10", "-gcge_ev_max_it", "40"], 7),
    ("A5M5_nev30_dirichlet", ("A_5", "M_5"), 30, ["-gcge_block_size", "10", "-gcge_ev_max_it", "60"], 8),
]


def dirichlet_rows(A, B):
    """Rows whose B row has no entry above 1e-20 and whose A row exactly one
    (the rule of app/csr/gcge_app_csr.c:130-183)."""
    rows = []
    for i in range(A.nrows):
        b = B.vals[B.rowptr[i]:B.rowptr[i + 1]]
        a = A.vals[A.rowptr[i]:A.rowptr[i + 1]]
        if not np.any(np.abs(b) > 1e-20) and np.count_nonzero(np.abs(a) > 1e-20) == 1:
            rows.append(i)
    return np.array(rows, dtype=np.int64)


# (name, matrices, k, max_it, rate, if_shift, shift, zero column or -1, seed)
BCG_RUNS = [
    ("fd12_rate", "fd12", 4, 20, 1e-2, 0, 0.0, -1, 21),
    ("fd12_zero_col", "fd12", 4, 20, 1e-2, 0, 0.0, 1, 22),
    ("fd12_fixed_iters", "fd12", 3, 9, 1e-30, 0, 0.0, -1, 23),
    ("q1e6_shift_B", "q1e6", 4, 20, 1e-2, 1, 3.5, -1, 24),
    ("q1e6_shift_I", "q1e6A", 3, 20, 1e-3, 1, 2.0, -1, 25),
]


def bcg_case(spec, k, zero_col, seed):
    """Matrices and inputs of one block CG fixture."""
    if spec == "fd12":
        A, B = problems.fd_laplacian_2d(12), None
    elif spec == "q1e6":
        A, B = problems.q1_fem_3d(6)
    else:
        A, B = problems.q1_fem_3d(6)[0], None
    rng = np.random.default_rng(seed)
    n = A.nrows
    rhs = rng.standard_normal((n, k))
    x0 = 0.01 * rng.standard_normal((n, k))
    if zero_col >= 0:
        rhs[:, zero_col] = 0.0
        x0[:, zero_col] = 0.0
    return A, B, np.asfortranarray(rhs), np.asfortranarray(x0)


def load(name):
    d = np.load(os.path.join(HERE, "data", name + ".npz"))
    return problems.CSR(int(d["n"]), int(d["n"]), d["rowptr"], d["colidx"], d["vals"])


def build(spec):
    if isinstance(spec, tuple):
        return load(spec[0]), load(spec[1])
    if spec == "fd24":
        return problems.fd_laplacian_2d(24), None
    if spec == "q1e10":
        return problems.q1_fem_3d(10)
    raise ValueError(spec)


def copy_petsc_fixtures():
    import shutil

    for f in ("A_5.petsc.bin", "M_5.petsc.bin"):
        shutil.copyfile(os.path.join(REF, "test", "data", f), os.path.join(HERE, "data", f))
        print("fixture", f)


def main():
    os.makedirs(os.path.join(HERE, "data"), exist_ok=True)
    copy_petsc_fixtures()
    for f in DATA:
        m = problems.CSR.read_text(os.path.join(REF, "test", "data", f))
        name = f.replace(".txt", "")
        np.savez_compressed(os.path.join(HERE, "data", name + ".npz"), n=m.nrows, rowptr=m.rowptr,
                            colidx=m.colidx, vals=m.vals)
        print("fixture", name, m.nrows, m.nnz)
    if not oracle.ref_available():
        print("reference build missing (make -C oracle ref); skipping ref_runs.npz")
        return
    out = {}
    for name, spec, nev, flags, seed in RUNS:
        A, B = build(spec)
        x0 = oracle.random_fill(seed, 0, 0, nev, A.nrows)
        if name.endswith("_dirichlet"):
            x0[dirichlet_rows(A, B), :] = 0.0
        r = oracle.ref_solve(A, B, nev, flags, x0=x0, print_eval=True)
        out[name + "/eval"] = r["eval"]
        out[name + "/res"] = r["res"]
        out[name + "/iters"] = np.array([r["num_iter"], r["num_conv"]])
        out[name + "/conv_history"] = r["history"][:, :2]
        out[name + "/eval_history"] = r["history"][:, 2:]
        print("ref run", name, r["num_iter"], r["num_conv"])
    np.savez_compressed(os.path.join(HERE, "ref_runs.npz"), **out)
    out = {}
    for name, spec, k, max_it, rate, if_shift, shift, zero_col, seed in BCG_RUNS:
        A, B, rhs, x0 = bcg_case(spec, k, zero_col, seed)
        x, r = oracle.ref_block_cg(A, B, rhs, x0, max_it, rate, if_shift, shift)
        out[name + "/rhs"] = rhs
        out[name + "/x0"] = x0
        out[name + "/x"] = x
        out[name + "/r"] = r
        out[name + "/params"] = np.array([max_it, rate, if_shift, shift])
        out[name + "/spec"] = np.array(spec)
        print("ref block CG", name, int(np.isnan(x).any(axis=0).sum()), "NaN columns")
    np.savez_compressed(os.path.join(HERE, "bcg_runs.npz"), **out)


if __name__ == "__main__":
    main()
