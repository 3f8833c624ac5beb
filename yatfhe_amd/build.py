"""Build the engine library (and the C++ drop-in shim) in-tree for gfx950.

    python -m yatfhe_amd.build

Produces yatfhe_amd/libyatfhe_gpu.so (HIP kernels + runtime + host keygen) and
yatfhe_amd/dropin_test_t32, dropin_test_t56: the reference's bootstrapping tests
written against the header-only C++ drop-in (csrc/dropin/yatfhe_dropin.h).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = os.environ.get("YGPU_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wno-unused-value",
    # packed FP32 disabled for all device code (the engine is FP64 and integer)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}", f"-I{CSRC}",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-pthread", f"-I{INCLUDE}", f"-I{CSRC}"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in ("ygpu_kernels.hip", "ygpu_kernels_crt.hip", "ygpu_kernels_lat.hip", "ygpu_kernels_lazy.hip", "ygpu_keyexpand.hip",
                "ygpu_runtime.hip", "ygpu_lazy.hip", "ygpu_server.hip", "ygpu_keyfile.hip"):
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    obj = os.path.join(BUILD, "ygpu_keygen.o")
    _run([CXX, *CXX_FLAGS, "-c", os.path.join(CSRC, "ygpu_keygen.cpp"), "-o", obj], verbose)
    objs.append(obj)
    lib = os.path.join(HERE, "libyatfhe_gpu.so")
    tmp = lib + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    os.replace(tmp, lib)
    dropin = os.path.join(CSRC, "dropin")
    test_src = os.path.join(dropin, "dropin_test.cpp")
    if os.path.exists(test_src):
        # the header-only drop-in, compiled the way a reference user would: one
        # program per torus width
        for width in (32, 56):
            exe = os.path.join(HERE, f"dropin_test_t{width}")
            _run([CXX, *CXX_FLAGS, f"-DTORUS{width}", f"-I{dropin}", "-o", exe + ".tmp", test_src, f"-L{HERE}",
                  "-lyatfhe_gpu", "-Wl,-rpath,$ORIGIN"], verbose)
            os.replace(exe + ".tmp", exe)
    return lib


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
