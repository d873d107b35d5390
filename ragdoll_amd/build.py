"""Builds the in-tree native libraries.

* `ragdoll_amd/libragdoll_hip.so`: the HIP kernels and the C ABI of include/ragdoll_hip.h,
  compiled for gfx950 with hipcc (cross-compiles without a GPU).
* `oracle/liboracle.so`: the CPU oracle (test infrastructure), compiled with gcc by its own
  Makefile.
"""
import os
import re
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libragdoll_hip.so")
SOURCES = ["rgh_analyse.hip", "rgh_api.hip", "rgh_learned.hip"]
HOST_SOURCES = ["rgh_cavlc.cpp", "rgh_cabac.cpp"]
ARCH = os.environ.get("RGH_OFFLOAD_ARCH", "gfx950")
# Device instructions no kernel may contain. v_ashr_pk_{u8,i8}_i32: the backend assumes they
# zero bits 31:16 of their destination, which MI355X leaves unchanged. Packed FP32 arithmetic:
# kept out of all code (the compile flags below disable it; this catches a regression).
FORBIDDEN = re.compile(r"\b(v_ashr_pk_[ui]8_i32|v_pk_add_f32|v_pk_mul_f32|v_pk_fma_f32)\b")
# DPP lane broadcasts (quad_perm:[k,k,k,k]) folded into arithmetic gave wrong results on MI355X;
# broadcasts stay plain v_mov_b32_dpp.
FORBIDDEN_DPP = re.compile(r"\b(v_(?!mov_b32)\w+_dpp)\b[^\n]*quad_perm:\[(\d),\2,\2,\2\]")
FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17",
         # packed FP32 instructions stay off: no code here wants them.
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-result"]


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose, quiet=False):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, stderr=subprocess.DEVNULL if quiet else None)


def _jobs():
    return max(1, min(8, os.cpu_count() or 1))


def check_device_code(verbose=False):
    """Compiles every source's device code to assembly and rejects forbidden instructions."""
    with tempfile.TemporaryDirectory() as tmp:
        def one(src):
            out = os.path.join(tmp, src + ".s")
            _run([hipcc()] + FLAGS + ["--cuda-device-only", "-S", "-I", os.path.join(ROOT, "include"),
                                      "-o", out, os.path.join(CSRC, src)], verbose, quiet=True)
            with open(out) as f:
                text = f.read()
            bad = sorted({m.group(1) for m in FORBIDDEN.finditer(text)} |
                         {m.group(1) + " (broadcast)" for m in FORBIDDEN_DPP.finditer(text)})
            if bad:
                raise RuntimeError("%s: device code contains %s" % (src, ", ".join(bad)))
        with ThreadPoolExecutor(_jobs()) as ex:
            list(ex.map(one, SOURCES))


def build_library(verbose=False):
    """Compiles each source to an object (in parallel, with the device-code check beside them),
    then links the shared library."""
    inc = ["-I", os.path.join(ROOT, "include")]
    with tempfile.TemporaryDirectory() as tmp:
        def obj(src):
            out = os.path.join(tmp, src + ".o")
            _run([hipcc()] + FLAGS + ["-fPIC", "-c"] + inc + ["-o", out, os.path.join(CSRC, src)], verbose)
            return out
        with ThreadPoolExecutor(_jobs()) as ex:
            check = ex.submit(check_device_code, verbose)
            objs = list(ex.map(obj, SOURCES + HOST_SOURCES))
            check.result()
        _run([hipcc()] + FLAGS + ["-fPIC", "-shared", "-o", LIB + ".tmp"] + objs, verbose)
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
