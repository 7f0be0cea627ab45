"""Build libmoris_gpu.so in-tree: host set-up in C++ (g++), kernels in HIP for gfx950 (hipcc).

Run as ``python -m moris_amd.build`` or through ``__graft_entry__.build()``.  Every call compiles;
outputs go to ``moris_amd/lib/`` (object files to ``moris_amd/_build/``).
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib")
LIBNAME = "libmoris_gpu.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HOST_SOURCES = [
    "setup/problem.cpp",
    "setup/enrich.cpp",
    "set/set_host.cpp",
    "set/extraction.cpp",
    "learned/dataset.cpp",
]
HIP_SOURCES = [
    "set/assemble.hip",
    "set/batch.hip",
    "set/generic.hip",
    "set/sideset.hip",
    "learned/lq_assemble.hip",
    "learned/mlp.hip",
    "solve/cg.hip",
    "solve/gmres.hip",
    "solve/ilu.hip",
    "setup/gpu_setup.hip",
]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-I" + os.path.join(ROOT, "include")]
# Host arithmetic of the set-up (cut vertices, level sets) must not be contracted into FMAs so
# that it is reproducible across compilers.
HOST_FLAGS = COMMON + ["-ffp-contract=off", "-Wall", "-Wno-unused-variable", "-Wno-unused-function",
                       "-pthread"]
HIP_FLAGS = COMMON + [
    "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wno-unused-variable", "-Wno-unused-function",
]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), p.stdout))
    return p.stdout


def _obj(src):
    return os.path.join(BUILD, src.replace("/", "__").rsplit(".", 1)[0] + ".o")


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    jobs = []
    for s in HOST_SOURCES:
        jobs.append([CXX] + HOST_FLAGS + ["-c", os.path.join(CSRC, s), "-o", _obj(s)])
    for s in HIP_SOURCES:
        jobs.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", _obj(s)])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    tmp = os.path.join(LIB, LIBNAME + ".tmp")
    objs = [_obj(s) for s in HOST_SOURCES + HIP_SOURCES]
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp] + objs)
    os.replace(tmp, os.path.join(LIB, LIBNAME))
    # HBM counter calibration program (copy / write-only kernels of a known size)
    _run([HIPCC] + HIP_FLAGS + ["-x", "hip", os.path.join(CSRC, "calib", "hbm_calibrate.cpp"), "-o",
                                os.path.join(LIB, "mgpu_hbm_calibrate")])
    return os.path.join(LIB, LIBNAME)


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
