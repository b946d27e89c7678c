"""Build recipe for the oracle's C restatement (test infrastructure).

Produces oracle/liboracle.so with gcc.  -ffp-contract=off keeps every
multiply-add separately rounded, matching the reference's evaluation.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "pssfss_oracle.c")
LIB = os.path.join(HERE, "liboracle.so")


def build(verbose: bool = False) -> str:
    cmd = ["gcc", "-O2", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-fPIC",
           "-shared", "-std=gnu11", SRC, "-o", LIB, "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.check_call(cmd)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
