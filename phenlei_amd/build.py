"""Build the HIP library (libphl.so) for gfx950 and the oracle (liborc.so).

The library is compiled in-tree with hipcc; the oracle with gcc.  Both use
-ffp-contract=off so that every multiply and add rounds as written.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libphl.so")
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liborc.so")
ORACLE_LIB_OMP = os.path.join(ORACLE_DIR, "liborc_omp.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
SOURCES = ["phl_kernels.hip", "phl_sa.hip", "phl_api.cpp", "phl_mesh.cpp", "phl_output.cpp", "phl_fts.cpp",
           "phl_mpnn.hip", "phl_mpnn_api.cpp"]


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return res


def build_library(verbose=False):
    cmd = [HIPCC] + HIP_FLAGS + ["-I", os.path.join(ROOT, "include"), "-o", LIB + ".tmp"]
    cmd += [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def build_oracle(verbose=False):
    """The sequential oracle, and the same source built with OpenMP (the CPU
    baseline over all host cores; identical results)."""
    for lib, extra in ((ORACLE_LIB, []), (ORACLE_LIB_OMP, ["-fopenmp"])):
        cmd = ["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-Wall", "-Wextra", *extra,
               "-o", lib + ".tmp", os.path.join(ORACLE_DIR, "phl_oracle.c"), "-lm"]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        os.replace(lib + ".tmp", lib)
    return ORACLE_LIB


def build_all(verbose=False):
    build_oracle(verbose)
    build_library(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
