"""Drop-in check (GPU): the reference solver core, compiled unchanged from its
sources, runs on the device backend through GCGE_HIP_SetOps
(oracle/ref_driver_hip.c) and reproduces its own CSR-backend run from the
same initial block: same iteration count and converged count, eigenvalues to
1e-10 relative, and the same per-iteration convergence counts where the
spectrum has no repeated eigenvalues."""
import os

import numpy as np
import pytest

import oracle
from gcge_old_amd import problems

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (oracle.ref_available() and oracle.ref_hip_available()),
                                 reason="reference binaries not built")]


def _fixture(golden_dir, name):
    d = np.load(os.path.join(golden_dir, "data", name + ".npz"))
    return problems.CSR(int(d["n"]), int(d["n"]), d["rowptr"], d["colidx"], d["vals"])


@pytest.mark.parametrize("case", ["A5M5_nev20", "A5M5_nev40", "fd2d_30_nev12", "q1_e12_nev20"])
def test_reference_core_on_device_backend(case, golden_dir):
    if case.startswith("A5M5"):
        A, B = _fixture(golden_dir, "A_5"), _fixture(golden_dir, "M_5")
        nev = int(case.split("nev")[1])
        flags = ["-gcge_block_size", "10", "-gcge_ev_max_it", "60"]
        simple = True
    elif case.startswith("fd2d"):
        A, B = problems.fd_laplacian_2d(30), None
        nev, flags, simple = 12, ["-gcge_ev_max_it", "60"], False
    else:
        A, B = problems.q1_fem_3d(12)
        nev, flags, simple = 20, ["-gcge_ev_max_it", "60", "-gcge_ev_tol", "1e-10"], False
    x0 = oracle.random_fill(21, 0, 0, nev, A.nrows)
    cpu = oracle.ref_solve(A, B, nev, flags, x0=x0, print_eval=True)
    dev = oracle.ref_solve(A, B, nev, flags, x0=x0, print_eval=True, on_device=True)
    assert dev["num_conv"] == cpu["num_conv"] == nev
    assert abs(dev["num_iter"] - cpu["num_iter"]) <= (0 if simple else 2)
    np.testing.assert_allclose(dev["eval"], cpu["eval"], rtol=1e-10)
    if simple:
        np.testing.assert_array_equal(dev["history"][:, :2], cpu["history"][:, :2])
