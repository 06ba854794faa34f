"""Builds a CPU emulation of the C-ABI library (tests only).

The kernel and host sources of ggcat_amd/csrc are compiled with g++ against
tests/emu/hip/hip_runtime.h, which runs every workgroup with real threads,
barriers and wavefront collectives.  The only source rewrite is the
`extern __shared__` LDS declaration, pointed at the emulated LDS buffer.
Used to check kernel logic (termination, indexing, results) on the CPU
before a GPU run; the product path never loads this library.
"""
from __future__ import annotations

import os
import re
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libggcat_gpu_emu.so")
LIB_ASAN = os.path.join(OUT, "libggcat_gpu_emu_asan.so")  # built with host sanitizers


def _newest_input() -> float:
    """Latest modification time of everything the library is built from."""
    paths = [os.path.abspath(__file__), os.path.join(HERE, "hip", "hip_runtime.h")]
    for d in (os.path.join(ROOT, "ggcat_amd", "csrc"), os.path.join(ROOT, "include")):
        paths += [os.path.join(d, n) for n in os.listdir(d)]
    return max(os.path.getmtime(p) for p in paths)


def build(sanitize: bool = False) -> str:
    # a library built from the current sources (without sanitizers) is reused
    if not sanitize and os.path.exists(LIB) and os.path.getmtime(LIB) > _newest_input():
        return LIB
    src = os.path.join(OUT, "src")
    if os.path.isdir(src):
        shutil.rmtree(src)
    os.makedirs(os.path.join(src, "ggcat_amd", "csrc"))
    shutil.copytree(os.path.join(ROOT, "include"), os.path.join(src, "include"))
    for name in os.listdir(os.path.join(ROOT, "ggcat_amd", "csrc")):
        if not (name.endswith(".hip") or name.endswith(".h") or name.endswith(".cpp")):
            continue
        text = open(os.path.join(ROOT, "ggcat_amd", "csrc", name)).read()
        text = re.sub(r"extern __shared__ __align__\(16\) uint8_t lds\[\];", "uint8_t* lds = emu::lds;", text)
        open(os.path.join(src, "ggcat_amd", "csrc", name), "w").write(text)
    glob = os.path.join(src, "emu_globals.cpp")
    open(glob, "w").write("#include <hip/hip_runtime.h>\nHIP_EMU_DEFINE_GLOBALS\n")
    flags = ["-std=c++20", "-O1", "-g", "-fPIC", "-shared", "-I", HERE, "-Wno-unused-result"]
    if sanitize:
        flags += ["-fsanitize=address,undefined", "-fno-omit-frame-pointer"]
    out = LIB_ASAN if sanitize else LIB
    cmd = ["g++", *flags, "-x", "c++", os.path.join(src, "ggcat_amd", "csrc", "api.hip"), "-x", "none", glob,
           "-o", out + ".tmp", "-lpthread", "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("emulation build failed")
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build("--asan" in sys.argv))
