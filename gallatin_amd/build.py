"""Build recipe of libgallatin_hip.so (gfx950) and the host CPU baseline.

    python -m gallatin_amd.build [-v]

Sources: gallatin_amd/csrc/*.hip, headers under include/.  The libraries are
written next to this file so they travel with the source tree.
"""
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libgallatin_hip.so")
CPU_LIB = os.path.join(HERE, "libgallatin_cpu.so")
ARCH = "gfx950"

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed fp32 ops stay disabled on these boxes
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-Wno-unused-command-line-argument",
]


def _hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found")
    return path


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_hip(verbose=False, jobs=4):
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for src in srcs:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        cmd = [_hipcc(), *HIP_FLAGS, "-I", INCLUDE, "-I", os.path.join(HERE, "csrc"),
               "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd)))
        if len(procs) >= jobs:
            s, p = procs.pop(0)
            if p.wait() != 0:
                raise RuntimeError(f"hipcc failed on {s}")
    for s, p in procs:
        if p.wait() != 0:
            raise RuntimeError(f"hipcc failed on {s}")
    tmp = LIB + ".tmp"
    # export only the C interface (exports.map): the host-side handles of
    # template kernels instantiated from the public headers stay local, so
    # a user's instantiations of the same templates are not bound to them
    _run([_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC",
          "-Wl,--version-script=" + os.path.join(HERE, "exports.map"), "-o", tmp, *objs],
         verbose)
    os.replace(tmp, LIB)
    return LIB


def build_cpu(verbose=False):
    src = os.path.join(HERE, "csrc", "cpu_baseline.cpp")
    if not os.path.exists(src):
        return None
    tmp = CPU_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-I", INCLUDE,
          "-o", tmp, src], verbose)
    os.replace(tmp, CPU_LIB)
    return CPU_LIB


def build(verbose=False):
    out = [build_hip(verbose)]
    c = build_cpu(verbose)
    if c:
        out.append(c)
    return out


if __name__ == "__main__":
    for p in build(verbose="-v" in sys.argv):
        print(p)
