"""Build libaccelerint_hip.so in-tree with hipcc for gfx950.

    python -m accelerint_amd.build          # compile the library and the CLI (always rebuilds)

The library is the product: hand-written HIP kernels plus the accelerInt C ABI
(include/accelerint.h).  No torch types cross the boundary.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libaccelerint_hip.so")
SOURCES = [os.path.join(CSRC, "accelerint.hip"), os.path.join(CSRC, "mlp.hip"), os.path.join(CSRC, "factor_test.hip")]
HEADERS = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hpp")] + [
    os.path.join(ROOT, "include", "accelerint.h")]

HIPCC_FLAGS = [
    "-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions stay off on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # the LU / solve loops are fully unrolled over compile-time N
    "-mllvm", "-pragma-unroll-threshold=2000000",
    "-Wno-unused-result", "-Wno-unused-value",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build libaccelerint_hip.so)")


def build(verbose: bool = False, extra_flags=()) -> str:
    """Compiles the sources in parallel (one hipcc per file) and links the library."""
    tmp = LIB + ".tmp%d" % os.getpid()
    objdir = tempfile.mkdtemp(prefix="accelerint_build_")
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    t0 = time.time()
    try:
        jobs = []
        for src in SOURCES:
            obj = os.path.join(objdir, os.path.splitext(os.path.basename(src))[0] + ".o")
            cmd = [hipcc(), *compile_flags, *extra_flags, "-I", os.path.join(ROOT, "include"), "-I", CSRC,
                   "-c", "-o", obj, src]
            jobs.append((src, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                                    text=True)))
        failed = []
        for src, obj, proc in jobs:
            out, _ = proc.communicate()
            if proc.returncode != 0:
                sys.stderr.write(out)
                failed.append(os.path.basename(src))
        if failed:
            raise RuntimeError("hipcc failed compiling %s" % ", ".join(failed))
        cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *[obj for _, obj, _ in jobs]]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError("hipcc failed linking %s" % LIB)
    finally:
        shutil.rmtree(objdir, ignore_errors=True)
    os.replace(tmp, LIB)
    if verbose:
        print("built %s in %.1fs" % (LIB, time.time() - t0))
    return LIB


CLI_SRC = os.path.join(HERE, "cli", "accelerint_main.cpp")
CLI = os.path.join(HERE, "accelerint_main")


def build_cli(verbose: bool = False) -> str:
    """The command-line driver (host C++ over the library's C ABI)."""
    tmp = CLI + ".tmp%d" % os.getpid()
    cmd = [shutil.which("g++") or "g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), CLI_SRC,
           "-o", tmp, "-L", HERE, "-laccelerint_hip", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed building %s" % CLI)
    os.replace(tmp, CLI)
    if verbose:
        print("built %s" % CLI)
    return CLI


def is_stale() -> bool:
    if not os.path.exists(LIB):
        return True
    t = os.path.getmtime(LIB)
    return any(os.path.getmtime(f) > t for f in SOURCES + HEADERS)


if __name__ == "__main__":
    build(verbose=True)
    build_cli(verbose=True)
