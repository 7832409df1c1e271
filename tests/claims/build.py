"""Build recipe of tests/claims/libclaim_kernels.so: test kernels
of the batched forms compiled against the allocator headers and linked with
libgallatin_hip.so.

    python -m tests.claims.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(HERE, "libclaim_kernels.so")


def build(verbose=False):
    sys.path.insert(0, ROOT)
    from gallatin_amd import build as hip_build
    tmp = LIB + ".tmp"
    cmd = [hip_build._hipcc(), *hip_build.HIP_FLAGS, "-shared",
           "-I", hip_build.INCLUDE, os.path.join(HERE, "claim_kernels.hip"),
           "-L", hip_build.HERE, "-lgallatin_hip",
           "-Wl,-rpath,$ORIGIN/../../gallatin_amd", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
