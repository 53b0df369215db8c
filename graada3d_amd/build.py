"""Builds libgraeng.so (HIP, gfx950) in-tree.

    python -m graada3d_amd.build

Each .hip translation unit is compiled with hipcc for gfx950 only, then the
objects are linked into graada3d_amd/libgraeng.so.  Device code is built
with -ffp-contract=off (the prism kernels must round as written) and with
packed-FP32 instruction selection disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libgraeng.so")
BUILD_DIR = os.path.join(HERE, "_build")

SOURCES = ["prism_kernels.hip", "rectilinear_kernels.hip", "linalg_kernels.hip",
           "fused_kernels.hip", "wavelet_kernels.hip", "gemm_kernels.hip", "cgls.hip",
           "engine.hip"]
ARCH = os.environ.get("GRAENG_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-I", INCLUDE, "-I", CSRC,
]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the engine needs ROCm's hipcc to build")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("command failed: " + " ".join(cmd))
    # the host pass warns about the packed-fp32 feature it does not know
    msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if msg:
        sys.stderr.write(msg + "\n")


def build(verbose: bool = False) -> str:
    hipcc = _hipcc()
    os.makedirs(BUILD_DIR, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(BUILD_DIR, src.replace(".hip", ".o"))
        cmd = [hipcc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_hostcheck() -> str:
    """Test-support libraries: the prism math compiled for the host (g++), in
    the reference's operation order (libhostcheck.so) and with the GPU's
    formulation -- pairwise atan differences, table log, polynomial atan --
    (libhostcheck_pairs.so)."""
    src = os.path.join(ROOT, "tests", "hostcheck", "prism_hostcheck.cpp")
    out = os.path.join(ROOT, "tests", "hostcheck", "libhostcheck.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
          "-I", CSRC, "-o", out, src, "-lm"])
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
          "-DGRA_ATAN_PAIRS=1", "-DGRA_TABLE_LOG=1", "-DGRA_FAST_ATAN=1", "-I", CSRC,
          "-o", out.replace(".so", "_pairs.so"), src, "-lm"])
    return out


def build_adaptor_check() -> str:
    """Test-support program: a C++ host using include/graeng.hpp."""
    src = os.path.join(ROOT, "tests", "cpp", "adaptor_check.cpp")
    out = os.path.join(ROOT, "tests", "cpp", "adaptor_check")
    _run(["g++", "-O2", "-std=c++17", "-I", INCLUDE, "-o", out, src, "-L", HERE, "-lgraeng",
          "-Wl,-rpath,$ORIGIN/../../graada3d_amd"])
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_hostcheck())
    print(build_adaptor_check())
