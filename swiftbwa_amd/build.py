"""Builds libhipsw.so (HIP, gfx950) and libswwork.so (C workload generator) in-tree.

Run as ``python -m swiftbwa_amd.build`` or through ``__graft_entry__.build()``.
Each call recompiles every source.
"""
import os
import subprocess
import sys
import concurrent.futures as cf

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
INCLUDE = os.path.join(ROOT, "include")

HIP_SOURCES = ["engine.hip", "ext_host.hip", "ext_kernel.hip", "nw_host.hip", "nw_kernel.hip",
               "local_host.hip", "local_kernel.hip", "server.hip", "cigar_host.hip",
               "driver_host.hip", "noalloc.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # packed FP32 instructions stay disabled in every HIP build of this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
    "-I" + INCLUDE, "-I" + CSRC,
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build failed: " + " ".join(cmd))
    return r.stdout


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def build_hipsw():
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(HIP_SOURCES))) as ex:
        objs = list(ex.map(_compile, HIP_SOURCES))
    out = os.path.join(HERE, "libhipsw.so")
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, out)
    return out


def build_workload():
    out = os.path.join(HERE, "libswwork.so")
    tmp = out + ".tmp"
    _run(["gcc", "-O3", "-std=c11", "-fPIC", "-shared", "-Wall", "-o", tmp,
          os.path.join(CSRC, "workload.c"), "-lm"])
    os.replace(tmp, out)
    return out


def build_all():
    return [build_hipsw(), build_workload()]


if __name__ == "__main__":
    for p in build_all():
        print(p)
