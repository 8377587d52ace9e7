#!/usr/bin/env python3
"""Build liblbm_amd.so (HIP, gfx950) in-tree with hipcc.

    python pytorch-lbm_amd/build.py            # build
    python pytorch-lbm_amd/build.py --check    # print the command only
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
SRC = os.path.join(CSRC, "lbm_api.hip")
INST = os.path.join(CSRC, "lbm_inst.hip")
OUT_DIR = os.path.join(HERE, "lbm", "_native")
OUT = os.path.join(OUT_DIR, "liblbm_amd.so")
OBJ_DIR = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# device and host code flags of every translation unit
CFLAGS = [
    "--offload-arch=gfx950",
    "-O3", "-std=c++17", "-fPIC",
    "-mcode-object-version=5",
    # bit-level parity with the reference's separately rounded tensor ops
    "-ffp-contract=off",
    # packed fp32 math off for the device compiler
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
FLAGS = CFLAGS + ["-shared"]

# the entry points, plus the launch sequences of each storage type (the bulk
# of the device code) in objects of their own, compiled in parallel
UNITS = [("lbm_api", SRC, [])] + [
    (f"lbm_inst_{name}", INST, [f"-DLBM_INST_PREC={k}"])
    for k, name in enumerate(("f32", "f64", "bf16", "f16"))]


ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
DRIVER_SRC = os.path.join(os.path.dirname(HERE), "tests", "c_abi", "abi_driver.cpp")
DRIVER = os.path.join(os.path.dirname(DRIVER_SRC), "abi_driver")


def build(verbose=True, jobs=None, out=None, extra=(), obj_dir=None):
    """Compile every unit (in parallel) and link `out` (default: the
    in-tree library).  `extra`: additional compiler flags, e.g. -D tuning
    macros for a variant build written elsewhere."""
    out = out or OUT
    obj_dir = obj_dir or OBJ_DIR
    os.makedirs(os.path.dirname(out), exist_ok=True)
    os.makedirs(obj_dir, exist_ok=True)
    jobs = jobs or int(os.environ.get("MAX_JOBS", "0")) or min(5, os.cpu_count() or 1)
    cmds = []
    for name, src, defs in UNITS:
        obj = os.path.join(obj_dir, name + ".o")
        cmds.append((obj, [HIPCC] + CFLAGS + list(extra) + defs + ["-c", "-o", obj, src]))
    pending = list(cmds)
    running = []
    while pending or running:
        while pending and len(running) < jobs:
            obj, cmd = pending.pop(0)
            if verbose:
                print(" ".join(cmd), flush=True)
            running.append((cmd, subprocess.Popen(cmd)))
        cmd, proc = running.pop(0)
        if proc.wait() != 0:
            for _, other in running:
                other.wait()
            raise subprocess.CalledProcessError(proc.returncode, cmd)
    link = [HIPCC] + FLAGS + ["-o", out + ".tmp"] + [obj for obj, _ in cmds]
    if verbose:
        print(" ".join(link), flush=True)
    subprocess.run(link, check=True)
    os.replace(out + ".tmp", out)
    if out == OUT:
        build_driver(verbose)
    return out


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
        for name, src, defs in UNITS:
            print(" ".join([HIPCC] + CFLAGS + defs + ["-c", "-o", name + ".o", src]))
    else:
        build()
