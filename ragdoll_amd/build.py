"""Builds the in-tree native libraries.

* `ragdoll_amd/libragdoll_hip.so`: the HIP kernels and the C ABI of include/ragdoll_hip.h,
  compiled for gfx950 with hipcc (cross-compiles without a GPU).
* `oracle/liboracle.so`: the CPU oracle (test infrastructure), compiled with gcc by its own
  Makefile.
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libragdoll_hip.so")
SOURCES = ["rgh_analyse.hip", "rgh_deblock.hip", "rgh_api.hip"]
ARCH = os.environ.get("RGH_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build_library(verbose=False):
    cmd = [hipcc(), "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
           # packed FP32 instructions stay off: no code here wants them.
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wall", "-Wno-unused-function", "-Wno-unused-result",
           "-I", os.path.join(ROOT, "include"),
           "-o", LIB + ".tmp"] + [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def build_oracle():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    return os.path.join(ROOT, "oracle", "liboracle.so")


def main():
    build_library(verbose=True)
    build_oracle()


if __name__ == "__main__":
    sys.exit(main())
