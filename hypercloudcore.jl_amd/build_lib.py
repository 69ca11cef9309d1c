"""Build recipe of libhcc.so (gfx950): every HIP kernel and the C ABI, compiled in-tree with hipcc.

Flags: -ffp-contract=off keeps each multiply and add separately rounded (the reference's Float64 order, bit for
bit); the packed-FP32 target feature is switched off for all device code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libhcc.so")
SOURCES = ["hcc_runtime.hip", "hcc_step.hip", "hcc_mlp.hip", "hcc_host.cpp"]
ARCH = os.environ.get("HCC_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall", "-Wno-unused-function",
          "-Wno-unused-command-line-argument", f"-I{CSRC}"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose: bool = False, jobs: int = 3) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    objs, procs = [], []
    for src in SOURCES:
        obj = os.path.join(LIBDIR, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, "-c", os.path.join(CSRC, src), "-o", obj] + COMMON
        if src.endswith(".hip"):
            cmd += ["-x", "hip"] + DEVICE
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        objs.append(obj)
        if len(procs) >= jobs:
            if procs.pop(0).wait() != 0:
                raise RuntimeError(f"compile failed: {src}")
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("compile failed")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", LIB + ".tmp"] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
