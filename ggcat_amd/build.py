"""Builds the native libraries in-tree.

* ``ggcat_amd/_build/libggcat_gpu.so``   -- the C-ABI library (HIP, gfx950)
* ``ggcat_amd/_build/libggcat_synth.so`` -- seeded synthetic workloads (host C++)
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
OUT = os.path.join(HERE, "_build")
CSRC = os.path.join(HERE, "csrc")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GGCAT_GPU_ARCH", "gfx950")

GPU_LIB = os.path.join(OUT, "libggcat_gpu.so")
SYNTH_LIB = os.path.join(OUT, "libggcat_synth.so")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed fp32 instructions are disabled for every HIP object we build
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build failed: {' '.join(cmd)}")


def build_gpu() -> str:
    os.makedirs(OUT, exist_ok=True)
    tmp = GPU_LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), os.path.join(CSRC, "api.hip"), "-o", tmp, "-lz"])
    os.replace(tmp, GPU_LIB)
    return GPU_LIB


def build_synth() -> str:
    os.makedirs(OUT, exist_ok=True)
    tmp = SYNTH_LIB + ".tmp"
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", os.path.join(CSRC, "synth.cpp"), "-o", tmp, "-lpthread"])
    os.replace(tmp, SYNTH_LIB)
    return SYNTH_LIB


def build_all() -> None:
    build_synth()
    build_gpu()


if __name__ == "__main__":
    build_all()
    print("built", GPU_LIB, SYNTH_LIB)
