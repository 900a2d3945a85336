"""Build recipe of the HIP library (gfx950) and the FP64 oracle.

    python physicell-mac-metal_amd/build.py          # library + oracle (+ reference build if present)

The library libphysicell_hip.so is written next to the Python package
(physicell_amd/), in-tree, so that it travels with the source tree.
Flags: -ffp-contract=off keeps every FP64 multiply and add separately
rounded (the operation sequence written in the kernels is the one executed);
packed FP32 instructions are disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "physicell_amd", "libphysicell_hip.so")
ARCH = os.environ.get("PHYSICELL_HIP_ARCH", "gfx950")

SOURCES = ["context.hip", "diffusion.hip", "sweep_part.hip", "cells.hip", "batch.hip", "motility.hip", "phenotype.hip", "neighbors.hip", "host/physicell_host.cpp", "host/snapshot.cpp"]
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         f"--offload-arch={ARCH}", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
         "-I" + os.path.join(CSRC, "host")]


def run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_library(jobs=4):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs, objs = [], []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.basename(src).replace(".hip", ".o").replace(".cpp", ".o"))
        objs.append(obj)
        cmd = ["hipcc"] + FLAGS + (["-x", "hip"] if src.endswith(".cpp") else []) + \
              ["-c", os.path.join(CSRC, src), "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= jobs:
            if procs.pop(0).wait() != 0:
                raise SystemExit("compile failed")
    for p in procs:
        if p.wait() != 0:
            raise SystemExit("compile failed")
    tmp = OUT + ".tmp"
    run(["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, OUT)
    return OUT


def build_oracle():
    run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    run(["sh", os.path.join(ROOT, "oracle", "build_ref.sh")])


if __name__ == "__main__":
    build_library()
    if "--no-oracle" not in sys.argv:
        build_oracle()
