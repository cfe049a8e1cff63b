"""Build the HIP library librajzel_amd.so in-tree (gfx950).

Flags that carry meaning for the results:
  -ffp-contract=off                 no FMA contraction: the reference's x86-64
                                    build evaluates every product and sum separately
  -fno-gpu-flush-denormals-to-zero  IEEE denormals, as on the CPU
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                    no packed FP32 instructions
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librajzel_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RZ_OFFLOAD_ARCH", "gfx950")

SOURCES = ["rz_kernels.hip", "rz_api.cpp", "rz_precond.hip", "rz_picks.cpp", "rz_taper.cpp", "rz_lbfgs.hip",
           "rz_smooth.hip"]
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
]


def build(verbose: bool = False) -> str:
    """Compile every source to an object and link the shared library."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs, cmds = [], []
    for src in SOURCES:
        obj = os.path.join(objdir, src.rsplit(".", 1)[0] + ".o")
        cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd[1:1] = ["-x", "hip"]
        cmds.append(cmd)
        objs.append(obj)
    # the sources compile independently: one process each, at most 8 at once
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        for f in [ex.submit(_run, c, verbose) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp], verbose)
    os.replace(tmp, LIB)
    build_fd_surface(verbose)
    return LIB


def _mpicc():
    for c in (os.environ.get("MPICC"), "/opt/conda/bin/mpicc", "/usr/bin/mpicc"):
        if c and os.path.exists(c):
            return c
    return None


def build_fd_surface(verbose: bool = False):
    """Compile the reference-surface shim csrc/rz_fd_surface.c (C, needs
    mpi.h) to _build/rz_fd_surface.o when an MPI C compiler is present.  The
    object is linked into the reference program together with its own
    sources; it is not part of librajzel_amd.so.  Returns the object path,
    or None without an MPI compiler."""
    mpicc = _mpicc()
    if mpicc is None:
        return None
    obj = os.path.join(HERE, "_build", "rz_fd_surface.o")
    env = dict(os.environ, MPICH_CC=os.environ.get("MPICH_CC", "gcc"))
    cmd = [mpicc, "-O2", "-fPIC", "-std=gnu99", "-ffp-contract=off", "-Wall", "-Wextra", "-Werror",
           "-I" + os.path.join(os.path.dirname(HERE), "include"),
           "-c", os.path.join(CSRC, "rz_fd_surface.c"), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, env=env)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stderr}")
    return obj


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{err}")
    if verbose and err:
        print(err, file=sys.stderr)


if __name__ == "__main__":
    print(build(verbose=True))
