"""Build recipe of libgallatin_hip.so (gfx950) and the host CPU baseline.

    python -m gallatin_amd.build [-v]

Sources: gallatin_amd/csrc/*.hip, headers under include/.  The libraries are
written next to this file so they travel with the source tree.
"""
import glob
import json
import os
import re
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libgallatin_hip.so")
# per-kernel registers, scratch and occupancy reported by the compiler
RESOURCES = os.path.join(HERE, "_build", "kernel_resources.json")
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
    resources = {}

    def finish(src, proc):
        # the remarks go to the resource table, other diagnostics through
        _, err = proc.communicate()
        resources.update(parse_resource_usage(err))
        sys.stderr.write("".join(l + "\n" for l in err.splitlines()
                                 if "kernel-resource-usage" not in l))
        if proc.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}")

    for src in srcs:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        cmd = [_hipcc(), *HIP_FLAGS, "-Rpass-analysis=kernel-resource-usage",
               "-I", INCLUDE, "-I", os.path.join(HERE, "csrc"), "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stderr=subprocess.PIPE, text=True)))
        if len(procs) >= jobs:
            finish(*procs.pop(0))
    for src, proc in procs:
        finish(src, proc)
    with open(RESOURCES, "w") as f:
        json.dump(resources, f, indent=1, sort_keys=True)
    tmp = LIB + ".tmp"
    # export only the C interface (exports.map): the host-side handles of
    # template kernels instantiated from the public headers stay local, so
    # a user's instantiations of the same templates are not bound to them
    _run([_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC",
          "-Wl,--version-script=" + os.path.join(HERE, "exports.map"), "-o", tmp, *objs],
         verbose)
    os.replace(tmp, LIB)
    return LIB


def parse_resource_usage(text):
    """Kernel name -> {"vgprs", "agprs", "sgprs", "scratch", "occupancy"}
    from hipcc's -Rpass-analysis=kernel-resource-usage remarks."""
    keys = {"VGPRs": "vgprs", "AGPRs": "agprs", "TotalSGPRs": "sgprs",
            "ScratchSize [bytes/lane]": "scratch", "Occupancy [waves/SIMD]": "occupancy"}
    out, name = {}, None
    for line in text.splitlines():
        m = re.search(r"remark:\s+Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            out[name] = {}
            continue
        m = re.search(r"remark:\s+([A-Za-z /\[\]]+): (\d+) \[", line)
        if m and name and m.group(1).strip() in keys:
            out[name][keys[m.group(1).strip()]] = int(m.group(2))
    return out


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
