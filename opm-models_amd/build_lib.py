"""Build recipe of libomodels_amd.so (gfx950) and of the oracle library.

Usage: python opm-models_amd/build_lib.py [--oracle-only|--lib-only]

The HIP library is compiled with hipcc for gfx950 only, one object per
source, with FP contraction off (the assembly kernels are compared bit for bit
with a sequential CPU evaluation of the same formulas) and packed-FP32
instructions disabled.
"""
import os
import subprocess
import sys
import concurrent.futures as cf

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libomodels_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["bo_kernels.hip", "linalg_kernels.hip", "ilu_tiled.hip", "ilu_colored.hip", "amg.hip", "ens_kernels.hip", "lens_kernels.hip", "noalloc.hip", "lens_host.cpp", "problem.cpp", "host_model.cpp", "io.cpp",
           "ensemble.cpp", "capi.cpp"]
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I", os.path.join(ROOT, "include"),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: %s" % cmd[-1])
    return r


def build_lib():
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        objs.append(obj)
        lang = ["-x", "hip"]
        cmds.append([HIPCC] + lang + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-C", odir, "-B", "liboracle.so"])
    return os.path.join(odir, "liboracle.so")


if __name__ == "__main__":
    args = set(sys.argv[1:])
    if "--lib-only" not in args:
        print(build_oracle())
    if "--oracle-only" not in args:
        print(build_lib())
