"""ctypes binding of the C ABI in include/ggcat_gpu.h and of the synthetic
workload generator.

There is no CPU fallback: if the HIP library is missing or a call fails, an
exception is raised.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
GPU_LIB = os.path.join(BUILD, "libggcat_gpu.so")
SYNTH_LIB = os.path.join(BUILD, "libggcat_synth.so")

OK, EINVAL, EHIP, ENOMEM, ESTATE, EAGAIN = 0, 1, 2, 3, 4, 5
HASH_SEQ, HASH_RK128 = 0, 1

_u8p = C.POINTER(C.c_uint8)
_u16p = C.POINTER(C.c_uint16)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_f64p = C.POINTER(C.c_double)


class GGCATGpuError(RuntimeError):
    pass


class Params(C.Structure):
    _fields_ = [
        ("k", C.c_uint32),
        ("m", C.c_uint32),
        ("buckets_log", C.c_uint32),
        ("second_buckets_log", C.c_uint32),
        ("forward_only", C.c_uint32),
        ("hash_kind", C.c_uint32),
        ("min_multiplicity", C.c_uint32),
    ]


class BuildOptions(C.Structure):
    """ggcat_gpu_build_options: m = 0 and buckets_log = 0 select the automatic values."""
    _fields_ = [
        ("k", C.c_uint32),
        ("m", C.c_uint32),
        ("min_multiplicity", C.c_uint32),
        ("buckets_log", C.c_uint32),
        ("forward_only", C.c_uint32),
        ("hash_kind", C.c_uint32),
        ("colors", C.c_uint32),
    ]


class Tuning(C.Structure):
    """ggcat_gpu_tuning: schedule knobs (zero = default); only read_block changes results."""
    _fields_ = [
        ("item_occ", C.c_uint32),
        ("lds_max_occ", C.c_uint32),
        ("merge_wgs_per_cu", C.c_uint32),
        ("part_occ", C.c_uint32),
        ("lds_table_log", C.c_uint32),
        ("table_pct", C.c_uint32),
        ("no_dedup", C.c_uint32),
        ("unit_bytes", C.c_uint32),
        ("use_graphs", C.c_uint32),
        ("unit_cap", C.c_uint32),
        ("rec_cap", C.c_uint32),
        ("merge_threads", C.c_uint32),
        ("read_block", C.c_uint32),
        ("host_reader", C.c_uint32),
        ("upload_chunk", C.c_uint32),
    ]


_gpu: dict = {}
_synth = None


def _ptr(a: np.ndarray | None, t):
    if a is None:
        return C.cast(None, t)
    return a.ctypes.data_as(t)


def gpu_lib(path: str = GPU_LIB) -> C.CDLL:
    """Loads libggcat_gpu.so (built by ggcat_amd.build); raises if absent."""
    if path in _gpu:
        return _gpu[path]
    if not os.path.exists(path):
        raise GGCATGpuError(f"{path} is missing: run `python -m ggcat_amd.build` (no CPU fallback exists)")
    lib = C.CDLL(path)
    V = C.c_void_p
    lib.ggcat_gpu_create.argtypes = [C.c_int, C.POINTER(V)]
    lib.ggcat_gpu_destroy.argtypes = [V]
    lib.ggcat_gpu_destroy.restype = None
    lib.ggcat_gpu_last_error.argtypes = [V]
    lib.ggcat_gpu_last_error.restype = C.c_char_p
    lib.ggcat_gpu_bucketing.argtypes = [V, V, C.c_uint64, C.c_int, _u64p, _u32p, C.c_uint32, C.POINTER(Params)]
    lib.ggcat_gpu_records_info.argtypes = [V, _u64p, _u32p, _u64p]
    lib.ggcat_gpu_records_copy.argtypes = [V, _u32p, _u32p, _u8p, _u16p, _u32p, _u64p, _u8p]
    lib.ggcat_gpu_kmers_merge.argtypes = [V]
    lib.ggcat_gpu_solid_info.argtypes = [V, _u64p, _u32p]
    lib.ggcat_gpu_solid_copy.argtypes = [V, _u64p, _u32p, _u8p, _u32p]
    lib.ggcat_gpu_timings.argtypes = [V, _f64p, C.c_uint32]
    lib.ggcat_gpu_counters.argtypes = [V, _u64p, C.c_uint32]
    lib.ggcat_gpu_device_alloc.argtypes = [V, C.c_uint64, C.POINTER(V)]
    lib.ggcat_gpu_device_free.argtypes = [V, V]
    lib.ggcat_gpu_copy_to_device.argtypes = [V, V, V, C.c_uint64]
    lib.ggcat_gpu_copy_device.argtypes = [V, V, V, C.c_uint64]
    lib.ggcat_gpu_copy_to_host.argtypes = [V, V, V, C.c_uint64]
    lib.ggcat_gpu_query_device_noalloc.argtypes = [V, V, C.c_uint64, C.c_uint32, V, V, V]
    lib.ggcat_gpu_color_subsets.argtypes = [V, _u64p, _u32p]
    lib.ggcat_gpu_colormap_write.argtypes = [V, C.c_char_p, C.POINTER(C.c_char_p), C.c_uint32, _u64p]
    lib.ggcat_gpu_colormap_write_sets.argtypes = [C.c_char_p, C.POINTER(C.c_char_p), C.c_uint32, _u64p, _u32p,
                                                  C.c_uint64]
    lib.ggcat_gpu_colormap_open.argtypes = [C.c_char_p, C.POINTER(V)]
    lib.ggcat_gpu_colormap_info.argtypes = [V, _u32p, _u64p]
    lib.ggcat_gpu_colormap_color_name.argtypes = [V, C.c_uint32]
    lib.ggcat_gpu_colormap_color_name.restype = C.c_char_p
    lib.ggcat_gpu_colormap_subset.argtypes = [V, C.c_uint64, _u32p, C.c_uint32, _u32p]
    lib.ggcat_gpu_colormap_close.argtypes = [V]
    lib.ggcat_gpu_partial_unitigs.argtypes = [V, C.c_uint32, _u8p, C.POINTER(C.c_uint16), _u8p, _u8p, _u8p]
    lib.ggcat_gpu_outliers.argtypes = [V, C.c_uint32, _u32p, _u32p, _u32p, _u64p, C.c_uint32]
    lib.ggcat_gpu_resplit_outliers.argtypes = [V, C.c_uint32, _u32p, _u32p, _u32p, _u32p, C.c_uint32]
    lib.ggcat_gpu_resplit.argtypes = [V, C.c_uint32, _u32p, _u32p, _u32p]
    lib.ggcat_gpu_query_graph.argtypes = [V, C.c_char_p, C.c_char_p, C.c_char_p, C.c_uint32, C.c_uint32, C.c_int,
                                          C.c_uint32, C.c_int, C.c_uint32, C.c_uint32, C.c_char_p, C.c_uint64]
    lib.ggcat_gpu_colormap_close.restype = None
    lib.ggcat_gpu_index_build.argtypes = [V]
    lib.ggcat_gpu_index_colors.argtypes = [V, _u32p]
    lib.ggcat_gpu_query.argtypes = [V, V, C.c_uint64, C.c_uint32, _u64p, _u64p, _u32p]
    lib.ggcat_gpu_serve_open.argtypes = [V, C.c_uint32, C.c_uint64]
    lib.ggcat_gpu_serve_submit.argtypes = [V, V, C.c_uint64, C.c_uint32, _u64p]
    lib.ggcat_gpu_serve_poll.argtypes = [V, C.c_int, _u64p]
    lib.ggcat_gpu_serve_fetch.argtypes = [V, C.c_uint64, _u64p, _u64p, _u32p]
    lib.ggcat_gpu_serve_stats.argtypes = [V, _u32p, _u32p]
    lib.ggcat_gpu_serve_close.argtypes = [V]
    lib.ggcat_gpu_unitigs.argtypes = [V, _u64p, _u64p]
    lib.ggcat_gpu_unitigs_copy.argtypes = [V, _u8p, _u32p, _u8p]
    lib.ggcat_gpu_maximal_unitigs.argtypes = [V, _u64p, _u64p]
    lib.ggcat_gpu_records_serialize.argtypes = [V, _u8p, C.c_uint64, _u64p, _u64p]
    lib.ggcat_gpu_unitigs_write_fasta.argtypes = [V, C.c_char_p, _u64p]
    lib.ggcat_gpu_extension_copy.argtypes = [V, _u8p]
    lib.ggcat_gpu_read_files.argtypes = [V, C.POINTER(C.c_char_p), C.c_uint32, C.c_uint32, C.c_int,
                                         C.POINTER(_u8p), _u64p, C.POINTER(_u64p), C.POINTER(_u32p), _u32p, _u64p]
    lib.ggcat_gpu_free_host.argtypes = [V]
    lib.ggcat_gpu_free_host.restype = None
    lib.ggcat_gpu_set_tuning.argtypes = [V, C.POINTER(Tuning)]
    lib.ggcat_gpu_solid_colors.argtypes = [V, _u32p, _u32p]
    lib.ggcat_gpu_plan_build.argtypes = [C.POINTER(BuildOptions), C.c_uint64, C.POINTER(Params)]
    lib.ggcat_gpu_build_from_files.argtypes = [V, C.POINTER(C.c_char_p), C.c_uint32, C.POINTER(BuildOptions),
                                               C.POINTER(Params)]
    for name in ("ggcat_gpu_device_alloc", "ggcat_gpu_device_free", "ggcat_gpu_copy_to_device",
                 "ggcat_gpu_create", "ggcat_gpu_bucketing", "ggcat_gpu_records_info", "ggcat_gpu_records_copy",
                 "ggcat_gpu_kmers_merge", "ggcat_gpu_solid_info", "ggcat_gpu_solid_copy", "ggcat_gpu_timings",
                 "ggcat_gpu_counters", "ggcat_gpu_read_files", "ggcat_gpu_build_from_files",
                 "ggcat_gpu_plan_build", "ggcat_gpu_set_tuning", "ggcat_gpu_solid_colors",
                 "ggcat_gpu_copy_device", "ggcat_gpu_index_build", "ggcat_gpu_index_colors", "ggcat_gpu_query",
                 "ggcat_gpu_serve_open", "ggcat_gpu_serve_submit", "ggcat_gpu_serve_poll", "ggcat_gpu_serve_fetch",
                 "ggcat_gpu_serve_stats", "ggcat_gpu_serve_close", "ggcat_gpu_unitigs", "ggcat_gpu_unitigs_copy",
                 "ggcat_gpu_maximal_unitigs", "ggcat_gpu_unitigs_write_fasta", "ggcat_gpu_records_serialize",
                 "ggcat_gpu_extension_copy", "ggcat_gpu_copy_to_host", "ggcat_gpu_query_device_noalloc",
                 "ggcat_gpu_color_subsets", "ggcat_gpu_colormap_write", "ggcat_gpu_colormap_write_sets",
                 "ggcat_gpu_colormap_open", "ggcat_gpu_colormap_info", "ggcat_gpu_colormap_subset",
                 "ggcat_gpu_query_graph", "ggcat_gpu_partial_unitigs", "ggcat_gpu_outliers",
                 "ggcat_gpu_resplit_outliers", "ggcat_gpu_resplit"):
        getattr(lib, name).restype = C.c_int
    _gpu[path] = lib
    return lib


def synth_lib() -> C.CDLL:
    global _synth
    if _synth is not None:
        return _synth
    if not os.path.exists(SYNTH_LIB):
        raise GGCATGpuError(f"{SYNTH_LIB} is missing: run `python -m ggcat_amd.build`")
    lib = C.CDLL(SYNTH_LIB)
    lib.synth_genome.argtypes = [C.c_uint64, C.c_uint64, C.c_double, C.c_uint32, C.c_uint32, C.c_double, _u8p, C.c_int]
    lib.synth_genome.restype = None
    lib.synth_reads.argtypes = [_u8p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32, C.c_double, _u8p, C.c_int]
    lib.synth_reads.restype = None
    lib.synth_descendant.argtypes = [_u8p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_double, C.c_double, C.c_uint64,
                                     _u8p, C.c_uint64]
    lib.synth_descendant.restype = C.c_uint64
    lib.synth_lognormal_len.argtypes = [C.c_uint64, C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_double]
    lib.synth_lognormal_len.restype = C.c_uint64
    lib.synth_queries.argtypes = [_u8p, _u64p, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32,
                                  C.c_double, _u8p, _u32p]
    lib.synth_queries.restype = C.c_uint64
    _synth = lib
    return lib


def _paths(paths) -> tuple:
    enc = [os.fsencode(p) for p in paths]
    arr = (C.c_char_p * max(1, len(enc)))(*enc)
    return arr, len(enc)


@dataclass
class ReadInput:
    joined: np.ndarray  # uint8, records each followed by '\n'
    color_starts: np.ndarray  # uint64 byte offsets (empty when uncolored)
    colors: np.ndarray  # uint32
    estimated_bases: int


def read_files(paths, k: int, colors: bool = False, lib_path: str = GPU_LIB) -> ReadInput:
    """FASTA/FASTQ(.gz) files -> joined layout (host code; needs no GPU)."""
    lib = gpu_lib(lib_path)
    arr, n = _paths(paths)
    jp, nb = _u8p(), C.c_uint64()
    csp, cvp = _u64p(), _u32p()
    nc, est = C.c_uint32(), C.c_uint64()
    rc = lib.ggcat_gpu_read_files(None, arr, n, k, int(colors), C.byref(jp), C.byref(nb), C.byref(csp),
                                  C.byref(cvp), C.byref(nc), C.byref(est))
    if rc != OK:
        raise GGCATGpuError(f"ggcat_gpu_read_files failed (status {rc}): "
                            f"{lib.ggcat_gpu_last_error(None).decode(errors='replace')}")
    try:
        joined = np.ctypeslib.as_array(jp, shape=(nb.value,)).copy() if nb.value else np.zeros(0, np.uint8)
        cs = np.ctypeslib.as_array(csp, shape=(nc.value,)).copy() if nc.value else np.zeros(0, np.uint64)
        cv = np.ctypeslib.as_array(cvp, shape=(nc.value,)).copy() if nc.value else np.zeros(0, np.uint32)
    finally:
        for p in (jp, csp, cvp):
            lib.ggcat_gpu_free_host(C.cast(p, C.c_void_p))
    return ReadInput(joined, cs, cv, est.value)


def _names(names) -> tuple:
    arr = (C.c_char_p * max(len(names), 1))(*[n.encode() for n in names])
    return arr, len(names)


def colormap_write_sets(path: str, names: list[str], subsets: list[list[int]], lib_path: str = GPU_LIB) -> None:
    """Writes a color map of explicit subsets (host code; needs no GPU)."""
    lib = gpu_lib(lib_path)
    off = np.zeros(len(subsets) + 1, np.uint64)
    off[1:] = np.cumsum([len(s) for s in subsets]) if subsets else []
    cols = np.array([c for s in subsets for c in s], np.uint32)
    arr, n = _names(names)
    rc = lib.ggcat_gpu_colormap_write_sets(str(path).encode(), arr, n, _ptr(off, _u64p),
                                           _ptr(cols, _u32p) if cols.size else C.cast(None, _u32p), len(subsets))
    if rc != OK:
        raise GGCATGpuError(f"ggcat_gpu_colormap_write_sets failed (status {rc}): "
                            f"{lib.ggcat_gpu_last_error(None).decode(errors='replace')}")


def colormap_read(path: str, lib_path: str = GPU_LIB) -> tuple[list[str], list[list[int]]]:
    """Reads a color map: (color names, subsets in id order) (host code)."""
    lib = gpu_lib(lib_path)
    h = C.c_void_p()
    rc = lib.ggcat_gpu_colormap_open(str(path).encode(), C.byref(h))
    if rc != OK:
        raise GGCATGpuError(f"ggcat_gpu_colormap_open failed (status {rc}): "
                            f"{lib.ggcat_gpu_last_error(None).decode(errors='replace')}")
    try:
        nc, ns = C.c_uint32(), C.c_uint64()
        lib.ggcat_gpu_colormap_info(h, C.byref(nc), C.byref(ns))
        names = [lib.ggcat_gpu_colormap_color_name(h, i).decode() for i in range(nc.value)]
        subsets = []
        buf = np.zeros(max(nc.value, 1), np.uint32)
        cnt = C.c_uint32()
        for s in range(ns.value):
            if lib.ggcat_gpu_colormap_subset(h, s, _ptr(buf, _u32p), buf.size, C.byref(cnt)) != OK:
                raise GGCATGpuError("ggcat_gpu_colormap_subset failed")
            subsets.append(buf[:cnt.value].tolist())
        return names, subsets
    finally:
        lib.ggcat_gpu_colormap_close(h)


def plan_build(options: BuildOptions, estimated_bases: int, lib_path: str = GPU_LIB) -> Params:
    """Parameters the reference derives for inputs of this size (host code)."""
    out = Params()
    rc = gpu_lib(lib_path).ggcat_gpu_plan_build(C.byref(options), estimated_bases, C.byref(out))
    if rc != OK:
        raise GGCATGpuError(f"ggcat_gpu_plan_build failed (status {rc})")
    return out


@dataclass
class Records:
    sub: np.ndarray
    length: np.ndarray
    flags: np.ndarray
    minpos: np.ndarray
    color: np.ndarray
    off: np.ndarray
    bases: np.ndarray  # ASCII


@dataclass
class Solid:
    keys: np.ndarray  # (n, key_words) uint64
    mult: np.ndarray
    flags: np.ndarray
    sub: np.ndarray


class GpuContext:
    """One device context (one HIP stream, its buffers)."""

    def __init__(self, device: int = 0, lib_path: str = GPU_LIB):
        self.lib = gpu_lib(lib_path)
        self._params: Params | None = None  # of the last bucketing call
        h = C.c_void_p()
        rc = self.lib.ggcat_gpu_create(device, C.byref(h))
        if rc != OK:
            raise GGCATGpuError(f"ggcat_gpu_create failed with status {rc}")
        self.h = h

    def close(self) -> None:
        if self.h:
            self.lib.ggcat_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str) -> None:
        if rc != OK:
            msg = self.lib.ggcat_gpu_last_error(self.h).decode(errors="replace")
            raise GGCATGpuError(f"{what} failed (status {rc}): {msg}")

    def bucketing(self, data, params: Params, colors: tuple[np.ndarray, np.ndarray] | None = None,
                  device_ptr: int | None = None, n_bytes: int | None = None) -> None:
        """data: uint8 numpy array (host) or None with device_ptr/n_bytes (device-resident)."""
        self._params = params
        if device_ptr is not None:
            ptr, n, on_dev = C.c_void_p(device_ptr), int(n_bytes), 1
        else:
            data = np.ascontiguousarray(data, dtype=np.uint8)
            ptr, n, on_dev = data.ctypes.data_as(C.c_void_p), data.size, 0
        cs = cv = None
        nc = 0
        if colors is not None:
            cs = np.ascontiguousarray(colors[0], dtype=np.uint64)
            cv = np.ascontiguousarray(colors[1], dtype=np.uint32)
            nc = cs.size
        self._check(self.lib.ggcat_gpu_bucketing(self.h, ptr, n, on_dev, _ptr(cs, _u64p), _ptr(cv, _u32p), nc,
                                                 C.byref(params)), "ggcat_gpu_bucketing")

    def build_from_files(self, paths, options: BuildOptions) -> Params:
        """Reads the files, buckets and merges; returns the parameters used."""
        arr, n = _paths(paths)
        used = Params()
        self._check(self.lib.ggcat_gpu_build_from_files(self.h, arr, n, C.byref(options), C.byref(used)),
                    "ggcat_gpu_build_from_files")
        return used

    def set_tuning(self, item_occ: int = 0, lds_max_occ: int = 0, merge_wgs_per_cu: int = 0,
                   part_occ: int = 0, lds_table_log: int = 0, table_pct: int = 0, no_dedup: int = 0,
                   unit_bytes: int = 0, use_graphs: int = 0, unit_cap: int = 0, rec_cap: int = 0,
                   merge_threads: int = 0, read_block: int = 0, host_reader: int = 0,
                   upload_chunk: int = 0) -> None:
        t = Tuning(item_occ, lds_max_occ, merge_wgs_per_cu, part_occ, lds_table_log, table_pct, no_dedup, unit_bytes,
                   use_graphs, unit_cap, rec_cap, merge_threads, read_block, host_reader, upload_chunk)
        self._check(self.lib.ggcat_gpu_set_tuning(self.h, C.byref(t)), "ggcat_gpu_set_tuning")

    def records(self) -> Records:
        nr, ns, nb = C.c_uint64(), C.c_uint32(), C.c_uint64()
        self._check(self.lib.ggcat_gpu_records_info(self.h, C.byref(nr), C.byref(ns), C.byref(nb)), "records_info")
        n = nr.value
        r = Records(np.empty(n, np.uint32), np.empty(n, np.uint32), np.empty(n, np.uint8), np.empty(n, np.uint16),
                    np.empty(n, np.uint32), np.empty(n, np.uint64), np.empty(nb.value, np.uint8))
        self._check(self.lib.ggcat_gpu_records_copy(self.h, _ptr(r.sub, _u32p), _ptr(r.length, _u32p),
                                                    _ptr(r.flags, _u8p), _ptr(r.minpos, _u16p), _ptr(r.color, _u32p),
                                                    _ptr(r.off, _u64p), _ptr(r.bases, _u8p)), "records_copy")
        return r

    def records_serialize(self) -> tuple[bytes, np.ndarray]:
        """The last bucketing result in the reference's bucket-file record
        layout: (all bucket streams concatenated, bucket offsets)."""
        nr, ns, nb = C.c_uint64(), C.c_uint32(), C.c_uint64()
        self._check(self.lib.ggcat_gpu_records_info(self.h, C.byref(nr), C.byref(ns), C.byref(nb)), "records_info")
        sb = self._params.second_buckets_log if self._params else 0
        off = np.zeros((ns.value >> sb) + 1, np.uint64)
        n = C.c_uint64()
        self._check(self.lib.ggcat_gpu_records_serialize(self.h, None, 0, _ptr(off, _u64p), C.byref(n)),
                    "ggcat_gpu_records_serialize")
        out = np.zeros(max(n.value, 1), np.uint8)
        self._check(self.lib.ggcat_gpu_records_serialize(self.h, _ptr(out, _u8p), out.size, _ptr(off, _u64p),
                                                         C.byref(n)), "ggcat_gpu_records_serialize")
        return out[:n.value].tobytes(), off

    def kmers_merge(self) -> None:
        self._check(self.lib.ggcat_gpu_kmers_merge(self.h), "ggcat_gpu_kmers_merge")

    def solid(self) -> Solid:
        n, kw = C.c_uint64(), C.c_uint32()
        self._check(self.lib.ggcat_gpu_solid_info(self.h, C.byref(n), C.byref(kw)), "solid_info")
        s = Solid(np.empty((n.value, kw.value), np.uint64), np.empty(n.value, np.uint32), np.empty(n.value, np.uint8),
                  np.empty(n.value, np.uint32))
        self._check(self.lib.ggcat_gpu_solid_copy(self.h, _ptr(s.keys, _u64p), _ptr(s.mult, _u32p),
                                                  _ptr(s.flags, _u8p), _ptr(s.sub, _u32p)), "solid_copy")
        return s

    def solid_colors(self) -> np.ndarray:
        """(n_solid, color_words) uint32 bitmaps, solid() order."""
        n, kw, w = C.c_uint64(), C.c_uint32(), C.c_uint32()
        self._check(self.lib.ggcat_gpu_solid_info(self.h, C.byref(n), C.byref(kw)), "solid_info")
        self._check(self.lib.ggcat_gpu_solid_colors(self.h, C.byref(w), C.cast(None, _u32p)), "solid_colors")
        out = np.zeros((n.value, w.value), np.uint32)
        self._check(self.lib.ggcat_gpu_solid_colors(self.h, C.byref(w), _ptr(out, _u32p)), "solid_colors")
        return out

    # ---- unitigs -------------------------------------------------------
    def _build_unitigs(self, joined: bool) -> tuple[int, int]:
        n, nb = C.c_uint64(), C.c_uint64()
        fn = "ggcat_gpu_maximal_unitigs" if joined else "ggcat_gpu_unitigs"
        self._check(getattr(self.lib, fn)(self.h, C.byref(n), C.byref(nb)), fn)
        return int(n.value), int(nb.value)

    def unitig_counts(self, joined: bool = False) -> tuple[int, int]:
        """Builds the intra-sub-bucket unitigs (joined=True: the maximal
        unitigs of the whole solid set) on the device and returns (unitigs,
        bytes) without copying them out."""
        return self._build_unitigs(joined)

    def write_unitigs_fasta(self, path: str) -> int:
        """Writes the last unitig result as FASTA ('>id LN:i:len'); returns
        the number of records."""
        n = C.c_uint64()
        self._check(self.lib.ggcat_gpu_unitigs_write_fasta(self.h, os.fsencode(path), C.byref(n)),
                    "ggcat_gpu_unitigs_write_fasta")
        return int(n.value)

    def unitigs(self, joined: bool = False):
        """Intra-sub-bucket unitigs: (list of ASCII sequences, sub-bucket per
        unitig, flags: bit0 open start, bit1 open end, bit2 circular).
        joined=True: the maximal unitigs of the whole solid set (sub-bucket 0)."""
        n, nb = C.c_uint64(), C.c_uint64()
        n.value, nb.value = self._build_unitigs(joined)
        seq = np.zeros(nb.value, np.uint8)
        sub, fl = np.zeros(n.value, np.uint32), np.zeros(n.value, np.uint8)
        self._check(self.lib.ggcat_gpu_unitigs_copy(self.h, _ptr(seq, _u8p), _ptr(sub, _u32p), _ptr(fl, _u8p)),
                    "ggcat_gpu_unitigs_copy")
        return seq.tobytes().split(b"\n")[:-1], sub, fl

    def outliers(self, threads: int) -> list[tuple[int, int, int]]:
        """(sub-bucket, resplit bucket count, records) of every outlier sub-bucket."""
        n = C.c_uint32()
        self._check(self.lib.ggcat_gpu_outliers(self.h, threads, C.byref(n), C.cast(None, _u32p),
                                                C.cast(None, _u32p), C.cast(None, _u64p), 0), "outliers")
        s, r, c = np.zeros(n.value + 1, np.uint32), np.zeros(n.value + 1, np.uint32), np.zeros(n.value + 1, np.uint64)
        self._check(self.lib.ggcat_gpu_outliers(self.h, threads, C.byref(n), _ptr(s, _u32p), _ptr(r, _u32p),
                                                _ptr(c, _u64p), n.value), "outliers")
        return [(int(s[i]), int(r[i]), int(c[i])) for i in range(n.value)]

    def resplit_outliers(self, threads: int) -> list[tuple[int, int, int]]:
        """Resplits the outlier sub-buckets of the last bucketing (found as
        outliers(threads) finds them); returns (sub-bucket, resplit bucket
        count, first resplit sub-bucket id) of each.  The next merge counts
        every resplit bucket as a sub-bucket of its own."""
        nr, ns, nb = C.c_uint64(), C.c_uint32(), C.c_uint64()
        self._check(self.lib.ggcat_gpu_records_info(self.h, C.byref(nr), C.byref(ns), C.byref(nb)), "records_info")
        cap = int(ns.value) + 1
        n = C.c_uint32()
        s, r, f = np.zeros(cap, np.uint32), np.zeros(cap, np.uint32), np.zeros(cap, np.uint32)
        self._check(self.lib.ggcat_gpu_resplit_outliers(self.h, threads, C.byref(n), _ptr(s, _u32p), _ptr(r, _u32p),
                                                        _ptr(f, _u32p), cap), "resplit_outliers")
        return [(int(s[i]), int(r[i]), int(f[i])) for i in range(n.value)]

    def resplit(self, subs: list[tuple[int, int]]) -> list[int]:
        """Resplits the listed (sub-bucket, bucket count) pairs of the last
        bucketing; returns the first resplit sub-bucket id of each."""
        s = np.array([x[0] for x in subs] + [0], np.uint32)
        c = np.array([x[1] for x in subs] + [0], np.uint32)
        f = np.zeros(len(subs) + 1, np.uint32)
        self._check(self.lib.ggcat_gpu_resplit(self.h, len(subs), _ptr(s, _u32p), _ptr(c, _u32p), _ptr(f, _u32p)),
                    "resplit")
        return [int(x) for x in f[:len(subs)]]

    def partial_unitigs(self, buckets_log: int, n: int) -> dict:
        """Routing of the last per-sub-bucket unitigs (n of them) as the
        reference's merge writes them: kind, bucket, rc, last_align, flags."""
        out = {"kind": np.zeros(n, np.uint8), "bucket": np.zeros(n, np.uint16), "rc": np.zeros(n, np.uint8),
               "last_align": np.zeros(n, np.uint8), "flags": np.zeros(n, np.uint8)}
        self._check(self.lib.ggcat_gpu_partial_unitigs(
            self.h, buckets_log, _ptr(out["kind"], _u8p), out["bucket"].ctypes.data_as(C.POINTER(C.c_uint16)),
            _ptr(out["rc"], _u8p), _ptr(out["last_align"], _u8p), _ptr(out["flags"], _u8p)),
            "ggcat_gpu_partial_unitigs")
        return out

    def extension_flags(self) -> np.ndarray:
        n, kw = C.c_uint64(), C.c_uint32()
        self._check(self.lib.ggcat_gpu_solid_info(self.h, C.byref(n), C.byref(kw)), "solid_info")
        out = np.zeros(n.value, np.uint8)
        self._check(self.lib.ggcat_gpu_extension_copy(self.h, _ptr(out, _u8p)), "ggcat_gpu_extension_copy")
        return out

    # ---- queries -------------------------------------------------------
    def index_build(self) -> None:
        self._check(self.lib.ggcat_gpu_index_build(self.h), "ggcat_gpu_index_build")

    def index_colors(self) -> int:
        n = C.c_uint32()
        self._check(self.lib.ggcat_gpu_index_colors(self.h, C.byref(n)), "ggcat_gpu_index_colors")
        return n.value

    def query(self, queries: np.ndarray):
        """Joined queries -> (query_kmers, matched, color_counts (n, n_colors) or None)."""
        q = np.ascontiguousarray(queries, np.uint8)
        nq = int(np.count_nonzero(q == 10))
        nc = self.index_colors()
        qk, mt = np.zeros(nq, np.uint64), np.zeros(nq, np.uint64)
        cc = np.zeros((nq, max(nc, 1)), np.uint32)
        self._check(self.lib.ggcat_gpu_query(self.h, q.ctypes.data_as(C.c_void_p), q.size, nq, _ptr(qk, _u64p),
                                             _ptr(mt, _u64p), _ptr(cc, _u32p) if nc else C.cast(None, _u32p)),
                    "ggcat_gpu_query")
        return qk, mt, (cc if nc else None)

    def serve_open(self, streams: int = 4, memory_budget: int = 1 << 30) -> None:
        self._check(self.lib.ggcat_gpu_serve_open(self.h, streams, memory_budget), "ggcat_gpu_serve_open")

    def serve_submit(self, queries: np.ndarray) -> int:
        q = np.ascontiguousarray(queries, np.uint8)
        t = C.c_uint64()
        self._check(self.lib.ggcat_gpu_serve_submit(self.h, q.ctypes.data_as(C.c_void_p), q.size,
                                                    int(np.count_nonzero(q == 10)), C.byref(t)),
                    "ggcat_gpu_serve_submit")
        return t.value

    def serve_poll(self, block: bool = False) -> int | None:
        t = C.c_uint64()
        rc = self.lib.ggcat_gpu_serve_poll(self.h, int(block), C.byref(t))
        if rc == EAGAIN:
            return None
        self._check(rc, "ggcat_gpu_serve_poll")
        return t.value

    def serve_fetch(self, ticket: int, n_queries: int):
        nc = self.index_colors()
        qk, mt = np.zeros(n_queries, np.uint64), np.zeros(n_queries, np.uint64)
        cc = np.zeros((n_queries, max(nc, 1)), np.uint32)
        self._check(self.lib.ggcat_gpu_serve_fetch(self.h, ticket, _ptr(qk, _u64p), _ptr(mt, _u64p),
                                                   _ptr(cc, _u32p) if nc else C.cast(None, _u32p)),
                    "ggcat_gpu_serve_fetch")
        return qk, mt, (cc if nc else None)

    def serve_stats(self) -> tuple[int, int]:
        a, b = C.c_uint32(), C.c_uint32()
        self._check(self.lib.ggcat_gpu_serve_stats(self.h, C.byref(a), C.byref(b)), "ggcat_gpu_serve_stats")
        return a.value, b.value

    def serve_close(self) -> None:
        self._check(self.lib.ggcat_gpu_serve_close(self.h), "ggcat_gpu_serve_close")

    def upload(self, data: np.ndarray) -> tuple[int, int]:
        """Copies a host array into a new device allocation; returns (ptr, nbytes)."""
        data = np.ascontiguousarray(data)
        p = C.c_void_p()
        self._check(self.lib.ggcat_gpu_device_alloc(self.h, data.nbytes, C.byref(p)), "device_alloc")
        self._check(self.lib.ggcat_gpu_copy_to_device(self.h, p, data.ctypes.data_as(C.c_void_p), data.nbytes),
                    "copy_to_device")
        return p.value, data.nbytes

    def alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        self._check(self.lib.ggcat_gpu_device_alloc(self.h, nbytes, C.byref(p)), "device_alloc")
        return p.value

    def copy_device(self, dst: int, src: int, nbytes: int) -> None:
        self._check(self.lib.ggcat_gpu_copy_device(self.h, C.c_void_p(dst), C.c_void_p(src), nbytes), "copy_device")

    def color_subsets(self) -> tuple[int, np.ndarray]:
        """(number of distinct color sets, subset id of every solid entry in solid() order)."""
        n = C.c_uint64()
        self._check(self.lib.ggcat_gpu_color_subsets(self.h, C.byref(n), C.cast(None, _u32p)), "color_subsets")
        out = np.zeros(int(self.counters()[5]), np.uint32)
        self._check(self.lib.ggcat_gpu_color_subsets(self.h, C.byref(n), _ptr(out, _u32p)), "color_subsets")
        return n.value, out

    def query_graph(self, graph: str, queries: str, out_prefix: str, k: int, colors: bool = False,
                    names: bool = False, hash_kind: int = HASH_SEQ) -> str:
        """ggcat_gpu_query_graph: query a unitig FASTA (+ color map) with a query
        file; returns the output file path."""
        buf = C.create_string_buffer(4096)
        self._check(self.lib.ggcat_gpu_query_graph(self.h, os.fsencode(graph), os.fsencode(queries),
                                                   os.fsencode(out_prefix), k, 1, 0, 0, int(colors), int(names),
                                                   hash_kind, buf, len(buf)), "ggcat_gpu_query_graph")
        return buf.value.decode()

    def colormap_write(self, path: str, names: list[str]) -> int:
        arr, n = _names(names)
        ns = C.c_uint64()
        self._check(self.lib.ggcat_gpu_colormap_write(self.h, str(path).encode(), arr, n, C.byref(ns)),
                    "colormap_write")
        return ns.value

    def download(self, ptr: int, out: np.ndarray) -> np.ndarray:
        """Copies device bytes at ptr into the host array `out` (its whole size)."""
        self._check(self.lib.ggcat_gpu_copy_to_host(self.h, out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr),
                                                    out.nbytes), "copy_to_host")
        return out

    def query_device_noalloc(self, d_queries: int, n_bytes: int, n_queries: int, d_query_kmers: int | None,
                             d_matched: int, d_color_counts: int | None) -> None:
        """ggcat_gpu_query_device_noalloc on caller-owned device buffers (pointers)."""
        self._check(self.lib.ggcat_gpu_query_device_noalloc(self.h, C.c_void_p(d_queries), n_bytes, n_queries,
                                                            C.c_void_p(d_query_kmers), C.c_void_p(d_matched),
                                                            C.c_void_p(d_color_counts)),
                    "ggcat_gpu_query_device_noalloc")

    def free(self, ptr: int) -> None:
        self._check(self.lib.ggcat_gpu_device_free(self.h, C.c_void_p(ptr)), "device_free")

    def timings(self) -> np.ndarray:
        out = np.zeros(9, np.float64)
        self._check(self.lib.ggcat_gpu_timings(self.h, _ptr(out, _f64p), 9), "timings")
        return out

    def counters(self) -> np.ndarray:
        out = np.zeros(11, np.uint64)
        self._check(self.lib.ggcat_gpu_counters(self.h, _ptr(out, _u64p), 11), "counters")
        return out
