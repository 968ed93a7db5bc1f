"""Builds libmrcc.so (HIP, gfx950) in the package directory.

hipcc compiles every csrc/*.hip for gfx950 with
  -ffp-contract=fast-honor-pragmas : binary32 bulk code may fuse; the binary64 predicates opt out
                                     with `#pragma clang fp contract(off)`
  -target-feature -packed-fp32-ops : no packed FP32 instructions are generated
and links them into one shared library with a plain C ABI (include/mrcc.h).
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(ROOT, "_build", "mrcc")
LIB = os.path.join(PKG, "libmrcc.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MRCC_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-ffp-contract=fast-honor-pragmas",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result",
    f"-I{os.path.join(ROOT, 'include')}",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    """Compiles every HIP source and links libmrcc.so; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    # C++ per-call driver over the C ABI (cpp/percall_bench.cpp)
    cpp = os.path.join(PKG, "cpp")
    r = subprocess.run([HIPCC, "-O2", "-std=c++17", f"-I{os.path.join(ROOT, 'include')}",
                        os.path.join(cpp, "percall_bench.cpp"), f"-L{PKG}", "-lmrcc",
                        "-Wl,-rpath,$ORIGIN/..", "-o", os.path.join(cpp, "percall_bench")],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"percall_bench failed:\n{r.stderr}")
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
