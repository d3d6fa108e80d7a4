"""In-tree build of the native libraries.

    python -m gcge_old_amd.build

libgcge_amd.so   GCG core, host C (gcc)
libgcge_hip.so   MI355X backend, HIP C++ for gfx950 (hipcc)
q1_solve         example C program on the public headers (gcc)

Every HIP translation unit is compiled with packed-FP32 instructions
disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).  Objects are
compiled in parallel and linked into gcge_old_amd/lib/.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gcge_old_amd")
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
OBJ = os.path.join(PKG, "lib", "obj")
INCLUDE = os.path.join(ROOT, "include")
ARCH = os.environ.get("GCGE_HIP_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CC = shutil.which("gcc") or "cc"

CORE_FLAGS = ["-O3", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter", "-I" + INCLUDE]
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-I" + INCLUDE,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result"]


def _run(cmd: list[str]) -> None:
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{p.stdout}\n{p.stderr}")


def _compile(job):
    cmd, = job
    _run(cmd)


def build_core() -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "core", "*.c")))
    objs, jobs = [], []
    for s in srcs:
        o = os.path.join(OBJ, "core_" + os.path.basename(s) + ".o")
        objs.append(o)
        jobs.append(([CC] + CORE_FLAGS + ["-c", s, "-o", o],))
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        list(ex.map(_compile, jobs))
    out = os.path.join(LIB, "libgcge_amd.so")
    _run([CC, "-shared", "-o", out] + objs + ["-lm", "-pthread"])
    return out


def build_hip() -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "hip", "*.hip")))
    objs, jobs = [], []
    for s in srcs:
        o = os.path.join(OBJ, "hip_" + os.path.basename(s) + ".o")
        objs.append(o)
        jobs.append(([HIPCC] + HIP_FLAGS + ["-c", s, "-o", o],))
    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        list(ex.map(_compile, jobs))
    out = os.path.join(LIB, "libgcge_hip.so")
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", out] + objs +
         ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"])
    return out


def build_examples() -> str:
    """C programs on the public headers only, linked to both libraries."""
    out = os.path.join(LIB, "q1_solve")
    _run([CC, "-O2", "-Wall", "-I" + INCLUDE, os.path.join(CSRC, "examples", "q1_solve.c"),
          "-o", out, "-L" + LIB, "-lgcge_amd", "-lgcge_hip", "-Wl,-rpath,$ORIGIN", "-lm"])
    return out


def build_oracle(with_reference: bool = True) -> None:
    """The checker: the CPU restatement, and the reference built from its own
    sources where the reference tree is present."""
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", odir])
    if with_reference and os.path.isdir("/root/reference/src"):
        _run(["make", "-s", "-j8", "-C", odir, "ref"])
        # the reference core on the device backend (the drop-in check)
        if os.path.exists(os.path.join(LIB, "libgcge_hip.so")):
            _run(["make", "-s", "-C", odir, "ref-hip"])


def build_all(verbose: bool = True) -> None:
    with cf.ThreadPoolExecutor(max_workers=2) as ex:
        futs = [ex.submit(build_core), ex.submit(build_hip)]
        for f in futs:
            r = f.result()
            if verbose and r:
                print("built", r)
    r = build_examples()
    if verbose:
        print("built", r)
    build_oracle()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
