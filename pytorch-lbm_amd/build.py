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


def build(verbose=True):
    os.makedirs(OUT_DIR, exist_ok=True)
    cmd = [HIPCC] + FLAGS + ["-o", OUT + ".tmp", SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    if "--check" in sys.argv:
        print(" ".join([HIPCC] + FLAGS + ["-o", OUT, SRC]))
    else:
        build()
