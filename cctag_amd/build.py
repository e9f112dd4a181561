"""Builds the HIP library libcctag_amd.so (gfx950) in-tree.

    python -m cctag_amd.build          # or __graft_entry__.build()

hipcc compiles every .hip source of csrc/ for gfx950; the shared host stages
(csrc/host/*.cpp) are plain C++ built with the ROCm clang++.  Float
contraction is off everywhere, so kernels and host code keep the reference's
rounding order; packed FP32 instructions are disabled in device code.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcctag_amd.so")
ARCH = os.environ.get("CCTAGX_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CLANGXX = "/opt/rocm/llvm/bin/clang++" if os.path.exists("/opt/rocm/llvm/bin/clang++") else "clang++"
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
          "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unknown-pragmas",
          "-Wno-sign-compare"]
# host stages: AVX2-era x86-64 (vector widths change, IEEE results do not:
# no contraction, no reassociation)
HOST_FLAGS = ["-march=x86-64-v3"]
HIP_FLAGS = COMMON + [f"--offload-arch={ARCH}",
                      "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def sources():
    out = [os.path.join(CSRC, f) for f in sorted(os.listdir(CSRC)) if f.endswith(".hip")]
    host = os.path.join(CSRC, "host")
    out += [os.path.join(host, f) for f in sorted(os.listdir(host)) if f.endswith(".cpp")]
    return out


def _compile(src: str) -> str:
    rel = os.path.relpath(src, CSRC).replace(os.sep, "_")
    obj = os.path.join(BUILD, rel + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, *HIP_FLAGS, "-I", CSRC, "-c", src, "-o", obj]
    else:
        cmd = [CLANGXX, *COMMON, *HOST_FLAGS, "-I", CSRC, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed for %s:\n%s%s" % (src, r.stdout, r.stderr))
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
