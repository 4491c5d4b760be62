"""Builds liblogo2opt.so (the HIP C-ABI library) in this directory for gfx950.

    python logo-tsp-solver_amd/build.py

The library is the product: hand-written gfx950 kernels (csrc/l2o_kernels.hip),
the host engine behind include/logo2opt.h (csrc/l2o_engine.hip), the
workload generators (csrc/l2o_workload.cpp), the server (csrc/l2o_server.cpp)
and the TSPLIB file reader/writer (csrc/l2o_tsplib.cpp).  It links only the HIP runtime.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "liblogo2opt.so")
SOURCES = ["csrc/l2o_kernels.hip", "csrc/l2o_engine.hip", "csrc/l2o_workload.cpp", "csrc/l2o_server.cpp",
           "csrc/l2o_tsplib.cpp", "csrc/l2o_noalloc.hip"]
ARCH = os.environ.get("LOGO2OPT_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    cmd = [
        hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        # exact host-side arithmetic (generators must match bit for bit); the
        # kernels request FMA explicitly where they want it
        "-ffp-contract=off",
        # packed FP32 VALU forms are kept out of the code object
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-command-line-argument",
        "-I", os.path.join(ROOT, "include"),
        "-o", LIB + ".tmp",
    ] + [os.path.join(HERE, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building liblogo2opt.so")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
