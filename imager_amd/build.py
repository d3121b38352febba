"""Build recipe of the gfx950 library and the C++ host layer.

``python -m imager_amd.build`` (or ``build()``) compiles, in-tree:
  * ``imager_amd/libblink_mi355x.so``  -- HIP kernels + the C ABI of include/blink_mi355x.h
  * ``imager_amd/libpacer_imager_mi355x.so`` -- the C++ drop-in imager class over that ABI
  * ``tests/cpp/bin/*``                 -- C++ test programs mirroring the reference's tests
with hipcc / g++ directly (no cmake).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libblink_mi355x.so")
HOST_LIB = os.path.join(HERE, "libpacer_imager_mi355x.so")
TESTS_CPP = os.path.join(ROOT, "tests", "cpp")
BIN = os.path.join(TESTS_CPP, "bin")
ORACLE_DIR = os.path.join(ROOT, "oracle")

ARCH = os.environ.get("BLINK_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]

KERNEL_SOURCES = ["blink_kernels.hip", "blink_batch.hip", "blink_dft_mfma.hip", "blink_capi.hip"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_library():
    srcs = [os.path.join(CSRC, s) for s in KERNEL_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *srcs])
    os.replace(tmp, LIB)
    return LIB


def build_host():
    """C++ host layer (CPacerImagerMI355X) and the C++ test programs, which
    link the CPU oracle (oracle/_build/libblink_oracle.so) as their checker."""
    srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp"))
    common = ["-O2", "-std=gnu++17", "-fPIC", "-I", INCLUDE, "-I", HOST,
              "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include", "-Wall"]
    tmp = HOST_LIB + ".tmp"
    _run(["g++", *common, "-shared", "-o", tmp, *srcs, "-L", HERE, "-lblink_mi355x",
          "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, HOST_LIB)
    tests = sorted(f for f in os.listdir(TESTS_CPP) if f.endswith(".cpp")) if os.path.isdir(TESTS_CPP) else []
    if not tests:
        return HOST_LIB
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    oracle_lib = os.path.join(ORACLE_DIR, "_build")
    os.makedirs(BIN, exist_ok=True)
    for t in tests:
        exe = os.path.join(BIN, os.path.splitext(t)[0])
        _run(["g++", *common, "-I", ORACLE_DIR, "-o", exe, os.path.join(TESTS_CPP, t),
              "-L", HERE, "-lpacer_imager_mi355x", "-lblink_mi355x", "-Wl,-rpath," + HERE,
              "-L", oracle_lib, "-lblink_oracle", "-Wl,-rpath," + oracle_lib,
              "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib", "-fopenmp"])
    return HOST_LIB


def build():
    build_library()
    build_host()


if __name__ == "__main__":
    build()
    sys.exit(0)
