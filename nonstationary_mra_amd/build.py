"""Build libmra_hip.so for gfx950 in-tree (nonstationary_mra_amd/_build/).

hipcc cross-compiles the device code without a GPU.  Packed-FP32 instructions
are disabled for every HIP translation unit.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(BUILD, "libmra_hip.so")
SHIM = os.path.join(BUILD, "shim_example")
REFMAIN = os.path.join(BUILD, "reference_main")
ROOT = os.path.dirname(HERE)

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MRA_OFFLOAD_ARCH", "gfx950")
# -amdgpu-mfma-vgpr-form keeps MFMA accumulators in VGPRs (no AGPR copies
# around every matrix instruction in loops whose accumulators are carried)
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-mllvm", "-amdgpu-mfma-vgpr-form", "-I", os.path.join(ROOT, "include")]
HIP_SOURCES = ["mra_engine.hip", "mra_mlp.hip", "mra_partition_gpu.hip"]
CPP_SOURCES = ["mra_partition.cpp", "mra_optimize.cpp", "mra_io.cpp"]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose=True):
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for s in CPP_SOURCES:
        o = os.path.join(BUILD, os.path.splitext(s)[0] + ".o")
        jobs.append((["g++", "-O2", "-std=c++17", "-fPIC", "-c", os.path.join(CSRC, s), "-o", o], o))
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, os.path.splitext(s)[0] + ".o")
        jobs.append(([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o], o))
    # the translation units compile independently (the engine is the long one)
    with ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 1)) as ex:
        list(ex.map(lambda j: run(j[0]), jobs))
    objs = [o for _, o in jobs]
    tmp = LIB + ".tmp"
    run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    # the reference-style C++ driver over include/mra_approximation.hpp
    for src, exe in (("shim_example.cpp", SHIM), ("reference_main.cpp", REFMAIN)):
        run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
             os.path.join(HERE, "examples", src), "-o", exe, "-L", BUILD, "-lmra_hip", "-Wl,-rpath,$ORIGIN"])
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
