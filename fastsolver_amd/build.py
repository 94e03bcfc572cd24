"""Build recipe for the native libraries (hipcc for gfx950, g++ for host code).

    python -m fastsolver_amd.build          # builds everything in-tree

Produces, next to this file:
  libfastsolver_amd.so      HIP kernels + plan + C-ABI (include/fastsolver_amd.h)
  libfastsolver_amd_gen.so  host input generation (include/fastsolver_amd_gen.h)
  examples/single_launch    C++ driver over include/fastsolver_amd.hpp
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
ARCH = os.environ.get("HODLR_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["kernels.hip", "plan.cc", "device.cc"]
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wno-unused-result",
             # packed FP32 instructions stay disabled for all device code
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
DEVICE_LIB = os.path.join(HERE, "libfastsolver_amd.so")
GEN_LIB = os.path.join(HERE, "libfastsolver_amd_gen.so")
EXAMPLES = os.path.join(HERE, "examples")
EXAMPLE_BINS = ["single_launch"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if p.returncode != 0:
        raise RuntimeError("build step failed: " + " ".join(cmd) + "\n" + out)
    return out


def build_device():
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        objs.append(obj)
        lang = [] if src.endswith(".hip") else ["-x", "hip"]
        jobs.append([HIPCC] + HIP_FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        list(ex.map(_run, jobs))
    tmp = DEVICE_LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, DEVICE_LIB)
    return DEVICE_LIB


def build_gen():
    tmp = GEN_LIB + ".tmp"
    _run(["g++", "-O2", "-fPIC", "-shared", "-fopenmp", "-std=c++17", "-Wall",
          os.path.join(CSRC, "gen", "hodlr_gen.cc"), "-o", tmp, "-ldl"])
    os.replace(tmp, GEN_LIB)
    return GEN_LIB


def build_examples():
    """C++ drivers linked against the two libraries (rpath $ORIGIN/..)."""
    out = []
    for name in EXAMPLE_BINS:
        exe = os.path.join(EXAMPLES, name)
        tmp = exe + ".tmp"
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
              os.path.join(EXAMPLES, name + ".cc"), "-L", HERE, "-lfastsolver_amd", "-lfastsolver_amd_gen",
              "-Wl,-rpath,$ORIGIN/..", "-o", tmp])
        os.replace(tmp, exe)
        out.append(exe)
    return out


def build_all():
    build_gen()
    build_device()
    build_examples()


if __name__ == "__main__":
    build_all()
    print("built", DEVICE_LIB, GEN_LIB)
    sys.exit(0)
