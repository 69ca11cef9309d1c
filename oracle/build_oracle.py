"""Build recipe of the CPU oracle (test infrastructure only).

Compiles oracle/hcc_oracle.cpp into oracle/liboracle.so with g++. `-ffp-contract=off` keeps every
multiply and add separately rounded, as in the Julia reference; OpenMP is used only to time the CPU
baseline (every per-node loop gives the same bits for any thread count).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "hcc_oracle.cpp")
OUT = os.path.join(HERE, "liboracle.so")


def build(verbose: bool = False) -> str:
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
           "-fno-fast-math", "-Wall", "-Wno-unused-function", SRC, "-o", OUT + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
