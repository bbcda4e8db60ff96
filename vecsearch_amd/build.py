"""In-tree build of the engine.

    python vecsearch_amd/build.py

1. ``libvsg_hip.so``: the HIP kernels and the C ABI (include/vsg.h), compiled by hipcc for
   gfx950. Packed-FP32 instructions are disabled and FP contraction is off: every fused
   multiply-add in the kernels is written out, so the arithmetic order is the one the code
   states.
2. ``_vecsearch*.so``: the pybind11 module; host C++ only (g++), linked to libvsg_hip.so.
"""
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("VSG_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["vsg_api.hip", "vsg_scan.hip", "vsg_lists.hip", "vsg_short.hip"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
LIB = os.path.join(HERE, "libvsg_hip.so")


def _ext_suffix():
    return sysconfig.get_config_var("EXT_SUFFIX") or ".so"


EXT = os.path.join(HERE, "_vecsearch" + _ext_suffix())


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")
    return r


def build_hip():
    os.makedirs(BUILD, exist_ok=True)
    hipcc = os.path.join(ROCM, "bin", "hipcc")

    def compile_one(src):
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([hipcc, *HIP_FLAGS, "-I", INCLUDE, "-I", CSRC, "-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=len(HIP_SOURCES)) as ex:
        objs = list(ex.map(compile_one, HIP_SOURCES))
    tmp = LIB + ".tmp"
    _run([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
          f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(tmp, LIB)
    return LIB


def build_ext():
    import pybind11

    py_inc = sysconfig.get_paths()["include"]
    tmp = EXT + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden",
          "-I", INCLUDE, "-I", pybind11.get_include(), "-I", py_inc,
          os.path.join(CSRC, "gpu_brute_force.cpp"), os.path.join(CSRC, "bindings.cpp"),
          "-o", tmp, "-L", HERE, "-lvsg_hip", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{ROCM}/lib"])
    os.replace(tmp, EXT)
    return EXT


def build():
    build_hip()
    build_ext()
    return LIB, EXT


if __name__ == "__main__":
    for p in build():
        print(p)
