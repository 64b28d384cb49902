"""Builds the oracle's host-C library (oracle/_build/liboracle.so) with gcc.

Test infrastructure: the library restates the reference's fixed OpenCL programs and the
CPU baselines (see reference_kernels.c). The refcl-host kernels of generated expressions
are compiled on demand by refcl.py into oracle/_build/refcl/.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "liboracle.so")


def build():
    os.makedirs(OUT, exist_ok=True)
    cmd = ["gcc", "-O2", "-std=gnu11", "-fopenmp", "-fPIC", "-shared", "-ffp-contract=off",
           "-o", LIB + ".tmp", os.path.join(HERE, "reference_kernels.c"), "-lm"]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stderr)
        raise RuntimeError("oracle build failed")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build())
