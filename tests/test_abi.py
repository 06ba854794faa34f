"""The C-ABI library loads on a machine without a GPU and exports every
function that include/ggcat_gpu.h declares (no compute calls here)."""
from __future__ import annotations

import ctypes
import os
import re

from ggcat_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ggcat_gpu.h")


def declared_functions() -> list[str]:
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(ggcat_gpu_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("ggcat_gpu_create", "ggcat_gpu_bucketing", "ggcat_gpu_kmers_merge", "ggcat_gpu_solid_copy",
                 "ggcat_gpu_read_files", "ggcat_gpu_build_from_files", "ggcat_gpu_plan_build"):
        assert must in names


def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(native.GPU_LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_contextless_error_text():
    lib = native.gpu_lib()
    assert lib.ggcat_gpu_last_error(None)  # never NULL
