"""Build libravesim_hip.so for gfx950 in-tree (hipcc, parallel per translation unit).

Usage: python 3d_ravesim_amd/build.py [--jobs N] [--verbose] [--sanitize]

--sanitize builds libravesim_hip_san.so instead: the same sources with the
host side instrumented by AddressSanitizer and UndefinedBehaviorSanitizer
(device code is built as usual).  It leaves the sanitizer runtimes to the
executable that loads it (tests/sanitize/abi_check.cpp, linked with
-fsanitize=address,undefined).
"""

from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libravesim_hip.so")
BUILD_SAN = os.path.join(HERE, "_build_san")
LIB_SAN = os.path.join(HERE, "libravesim_hip_san.so")
# the host-sanitizer checker of the C ABI and its executable
CHECK_SRC = os.path.join(os.path.dirname(HERE), "tests", "sanitize", "abi_check.cpp")
CHECK_EXE = os.path.join(BUILD_SAN, "abi_check")
CLANGXX = os.environ.get("RSIM_CLANGXX", "/opt/rocm/llvm/bin/clang++")
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: no v_pk_*_f32 in generated code.
COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", CSRC,
    "-I", os.path.join(os.path.dirname(HERE), "include"),
]


# host-only instrumentation: each sanitizer flag applies to the host pass
SANITIZE = [
    "-Xarch_host", "-fsanitize=address",
    "-Xarch_host", "-fsanitize=undefined",
    "-Xarch_host", "-fno-omit-frame-pointer",
    "-Xarch_host", "-g",
]


def sources() -> list[str]:
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src: str, verbose: bool, sanitize: bool = False) -> tuple[str, int, str]:
    obj = os.path.join(os.path.join(BUILD_SAN, "obj") if sanitize else BUILD,
                       os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC, *COMMON, *(SANITIZE if sanitize else []), "-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (p.stdout + p.stderr).splitlines() if "packed-fp32-ops" not in l)
    return obj, p.returncode, out


def build(jobs: int | None = None, verbose: bool = False, sanitize: bool = False) -> str:
    os.makedirs(os.path.join(BUILD_SAN, "obj") if sanitize else BUILD, exist_ok=True)
    lib = LIB_SAN if sanitize else LIB
    srcs = sources()
    jobs = jobs or min(len(srcs), os.cpu_count() or 4, 8)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda s: _compile(s, verbose, sanitize), srcs))
    failed = [(o, out) for o, rc, out in results if rc != 0]
    for o, rc, out in results:
        if out.strip() and (verbose or rc != 0):
            print(out, file=sys.stderr)
    if failed:
        raise RuntimeError("HIP compilation failed: " + ", ".join(os.path.basename(o) for o, _ in failed))
    tmp = lib + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp,
           *[o for o, _, _ in results]]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stdout + p.stderr)
    os.replace(tmp, lib)
    if sanitize:
        build_checker(verbose)
    return lib


def build_checker(verbose: bool = False) -> str:
    """Host executable that drives the C ABI of libravesim_hip_san.so; it
    carries the ASan and UBSan runtimes (host code only)."""
    os.makedirs(BUILD_SAN, exist_ok=True)
    cmd = [CLANGXX, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-I", os.path.join(os.path.dirname(HERE), "include"),
           CHECK_SRC, "-o", CHECK_EXE, LIB_SAN, f"-Wl,-rpath,{HERE}"]
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("checker build failed:\n" + p.stdout + p.stderr)
    return CHECK_EXE


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--sanitize", action="store_true",
                    help="host ASan + UBSan build: libravesim_hip_san.so")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose, a.sanitize))
