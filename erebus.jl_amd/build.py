"""Build libereb_hip.so (gfx950) in-tree: erebus.jl_amd/lib/libereb_hip.so.

Every source is compiled with hipcc for gfx950 only, with
  -ffp-contract=off   the Float64 path must round like the reference (Julia
                      never fuses a*b+c); the mixed-precision kernels call
                      fma explicitly where they want it
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      packed FP32 instructions disabled
With --sanitize it also builds lib/libereb_hip_san.so: the host side of
every source (argument validation, slot-table discovery, the field tables)
with AddressSanitizer and UndefinedBehaviorSanitizer (the device code is
not instrumented); the CPU tests load it under the sanitizer runtime, with no
GPU.
Usage: python erebus.jl_amd/build.py [-j N] [--sanitize]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libereb_hip.so")

SOURCES = [
    "ereb_ctx.hip", "ereb_scan.hip", "ereb_sort.hip", "ereb_scatter.hip",
    "ereb_nodes.hip", "ereb_markers.hip", "ereb_assembly.hip", "ereb_adjoint.hip",
    "ereb_hydration.hip", "ereb_backtrace.hip", "ereb_replenish.hip", "ereb_noalloc.hip",
    "ereb_pullback.hip", "ereb_save.hip",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-value",
]


LIB_SAN = os.path.join(LIBDIR, "libereb_hip_san.so")
# each sanitizer flag bound to the host compilation (the device code is built
# as usual, and never run from this library)
SAN_FLAGS = [
    "--offload-arch=gfx950", "-O1", "-g", "-fPIC", "-std=c++17",
    "-ffp-contract=off", "-fno-omit-frame-pointer",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
    "-Xarch_host", "-fno-sanitize-recover=all",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-value",
]


def _compile(src: str, san: bool = False) -> str:
    out = os.path.join(OBJ, src.replace(".hip", ".san.o" if san else ".o"))
    cmd = [HIPCC, *(SAN_FLAGS if san else FLAGS), "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(l for l in r.stderr.splitlines()
                     if "packed-fp32-ops" not in l and "not a recognized feature" not in l
                     and l.strip())
    if warn:
        sys.stderr.write(warn + "\n")
    return out


def build(jobs: int = 8) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


def build_sanitized(jobs: int = 8) -> str:
    """The host side of the library under ASan + UBSan (see the module
    docstring); it links the sanitizer runtime dynamically, which the
    loading process must preload."""
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, True), SOURCES))
    tmp = LIB_SAN + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC",
           "-shared-libsan", "-Xarch_host", "-fsanitize=address", "-Xarch_host",
           "-fsanitize=undefined", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"sanitizer link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB_SAN)
    return LIB_SAN


def sanitizer_runtime() -> str:
    """Path of the compiler's ASan runtime, which a process loading
    libereb_hip_san.so must preload."""
    r = subprocess.run([HIPCC, "-print-file-name=libclang_rt.asan-x86_64.so"],
                       capture_output=True, text=True)
    p = r.stdout.strip()
    if r.returncode != 0 or not os.path.isabs(p) or not os.path.exists(p):
        r = subprocess.run([HIPCC, "--print-runtime-dir"], capture_output=True, text=True)
        p = os.path.join(r.stdout.strip(), "libclang_rt.asan-x86_64.so")
    if not os.path.exists(p):
        raise RuntimeError("clang ASan runtime not found")
    return p


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("--sanitize", action="store_true")
    a = ap.parse_args()
    print(build(a.j))
    if a.sanitize:
        print(build_sanitized(a.j))
