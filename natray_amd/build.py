"""Builds libnatray.so (HIP kernels for gfx950 + the C-ABI host library) in-tree.

    python -m natray_amd.build

The kernels are compiled without floating-point contraction and with IEEE f32 division and
square root, so every float expression rounds the way the source writes it; packed-FP32
instructions are disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libnatray.so")
DRIVE = os.path.join(HERE, "natray_drive")
COPY = os.path.join(HERE, "natray_copy")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("NATRAY_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
    f"-I{os.path.join(ROOT, 'include')}",
    f"-I{CSRC}",
]

SOURCES = ["natray_kernels.hip", "natray_net.hip", "natray_sim.cpp"]


def build(verbose: bool = False) -> str:
    """Compiles every source and links libnatray.so; returns its path."""
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(CSRC, src + ".o")
        cmd = [HIPCC, *FLAGS, "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for obj in objs:
        os.remove(obj)
    # The native driver a profiler wraps (natray_drive.cpp), linked against the library.
    cmd = [HIPCC, "-O2", "-std=c++17", f"-I{os.path.join(ROOT, 'include')}", os.path.join(CSRC, "natray_drive.cpp"),
           "-o", DRIVE + ".tmp", f"-L{HERE}", "-lnatray", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(DRIVE + ".tmp", DRIVE)
    # The plain copy that calibrates the HBM traffic counters (natray_copy.hip).
    cmd = [HIPCC, *FLAGS, os.path.join(CSRC, "natray_copy.hip"), "-o", COPY + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(COPY + ".tmp", COPY)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
