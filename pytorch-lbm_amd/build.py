#!/usr/bin/env python3
"""Build liblbm_amd.so (HIP, gfx950) in-tree with hipcc.

    python pytorch-lbm_amd/build.py            # build
    python pytorch-lbm_amd/build.py --check    # print the command only
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "lbm_api.hip")
OUT_DIR = os.path.join(HERE, "lbm", "_native")
OUT = os.path.join(OUT_DIR, "liblbm_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3", "-std=c++17", "-fPIC", "-shared",
    "-mcode-object-version=5",
    # bit-level parity with the reference's separately rounded tensor ops
    "-ffp-contract=off",
    # packed fp32 math off for the device compiler
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
DRIVER_SRC = os.path.join(os.path.dirname(HERE), "tests", "c_abi", "abi_driver.cpp")
DRIVER = os.path.join(os.path.dirname(DRIVER_SRC), "abi_driver")


def build(verbose=True):
    os.makedirs(OUT_DIR, exist_ok=True)
    cmd = [HIPCC] + FLAGS + ["-o", OUT + ".tmp", SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    build_driver(verbose)
    return OUT


def build_driver(verbose=True):
    """Host-only C++ caller of the C ABI (tests/c_abi): g++, linked against
    the library above and the HIP runtime, no torch."""
    if not os.path.exists(DRIVER_SRC):
        return None
    rel = os.path.relpath(OUT_DIR, os.path.dirname(DRIVER))
    cmd = ["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__",
           "-I" + os.path.join(ROCM, "include"), DRIVER_SRC, "-o", DRIVER + ".tmp",
           "-L" + OUT_DIR, "-llbm_amd", "-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
           "-Wl,-rpath,$ORIGIN/" + rel, "-Wl,-rpath," + os.path.join(ROCM, "lib")]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


if __name__ == "__main__":
    if "--check" in sys.argv:
        print(" ".join([HIPCC] + FLAGS + ["-o", OUT, SRC]))
    else:
        build()
