"""Writes the committed fixtures under tests/golden/ (run in the container
that holds the reference tree; the GPU box only reads the results).

1. data/*.npz -- the CSR matrices the reference ships as test inputs
   (test/data/testA, testB, A_3.txt, M_3.txt, A_5.txt, M_5.txt), converted
   from the reference text format to npz.
2. ref_runs.npz -- outputs of the reference solver built from its own
   sources (oracle/_ref/gcge_ref) on small problems, each started from a
   fixed initial block: final eigenvalues, residuals, iteration counts and
   the per-iteration convergence counts.  These pin the GCG core's control
   flow (soft locking, unlock lists, block CG, P construction) without the
   reference tree at test time.

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
]


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


def main():
    os.makedirs(os.path.join(HERE, "data"), exist_ok=True)
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
        r = oracle.ref_solve(A, B, nev, flags, x0=x0, print_eval=True)
        out[name + "/eval"] = r["eval"]
        out[name + "/res"] = r["res"]
        out[name + "/iters"] = np.array([r["num_iter"], r["num_conv"]])
        out[name + "/conv_history"] = r["history"][:, :2]
        out[name + "/eval_history"] = r["history"][:, 2:]
        print("ref run", name, r["num_iter"], r["num_conv"])
    np.savez_compressed(os.path.join(HERE, "ref_runs.npz"), **out)


if __name__ == "__main__":
    main()
