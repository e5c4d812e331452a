"""Builds libaphi_amd.so in-tree: host C++ with g++, HIP sources with hipcc for gfx950.

Usage: python hybrid-aphi-solver_amd/build.py [--jobs N]
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib", "libaphi_amd.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

HOST_SOURCES = [
    "host/mesh.cpp",
    "host/problem.cpp",
    "host/sparsity.cpp",
    "host/ordering.cpp",
    "host/symbolic.cpp",
    "host/gauge_free.cpp",
    "host/writers.cpp",
    "capi/problem_capi.cpp",
]
HIP_SOURCES = [
    "device/assembly.hip",
    "device/random.hip",
    "device/factor.hip",
    "device/solve.hip",
    "capi/device_capi.hip",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter"]
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # packed FP32 VALU ops stay off (compiler-generated v_pk_*_f32 only)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-parameter", "-Wno-unused-result",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed:\n  " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _obj(src):
    return os.path.join(BUILD, src.replace("/", "_") + ".o")


def build(jobs=8, verbose=False):
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    cmds = []
    for s in HOST_SOURCES:
        cmds.append(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, s), "-o", _obj(s)])
    for s in HIP_SOURCES:
        cmds.append([HIPCC, *HIP_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, s), "-o", _obj(s)])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for r in ex.map(_run, cmds):
            if verbose and (r.stdout or r.stderr):
                sys.stderr.write(r.stdout + r.stderr)
    objs = [_obj(s) for s in HOST_SOURCES + HIP_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp, *objs])
    shutil.move(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))
