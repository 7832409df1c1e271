"""Build recipe for the oracle (test infrastructure only).

Compiles oracle/gallatin_oracle.c into oracle/liboracle.so with gcc.  The
reference itself (CUDA C++, needs nvcc and CUDA cooperative groups) cannot be
compiled in this image, so there is no oracle/_ref build.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build(verbose: bool = False) -> str:
    src = os.path.join(HERE, "gallatin_oracle.c")
    cmd = ["gcc", "-O2", "-std=c11", "-Wall", "-Wextra", "-Wno-unused-parameter",
           "-shared", "-fPIC", "-o", LIB + ".tmp", src]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
