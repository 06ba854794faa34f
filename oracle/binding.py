"""ctypes binding of the CPU oracle (oracle/ggcat_oracle.cpp).

Test infrastructure only: imported by tests/, by __graft_entry__.smoke() and
by bench.py's CPU-baseline leg -- never by the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libggcat_oracle.so")

_u8p = C.POINTER(C.c_uint8)
_u16p = C.POINTER(C.c_uint16)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    L = C.CDLL(LIB)
    V = C.c_void_p
    L.orc_nthash.argtypes = [_u8p, C.c_size_t, C.c_uint, _u64p, _u64p]
    L.orc_nthash.restype = None
    L.orc_minqueue.argtypes = [_u64p, _u32p, _u8p, C.c_size_t, C.c_size_t, _u64p, _u32p, _u8p]
    L.orc_minqueue.restype = C.c_size_t
    L.orc_bucketing.argtypes = [_u8p, _u64p, C.c_size_t, _u32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                C.c_int, C.c_int]
    L.orc_bucketing.restype = V
    L.orc_compact.argtypes = [V]
    L.orc_compact.restype = None
    L.orc_rec_count.argtypes = [V]
    L.orc_rec_count.restype = C.c_size_t
    L.orc_rec_bases.argtypes = [V]
    L.orc_rec_bases.restype = C.c_size_t
    L.orc_rec_get.argtypes = [V, _u32p, _u32p, _u8p, _u16p, _u32p, _u64p, _u8p]
    L.orc_rec_get.restype = None
    L.orc_merge.argtypes = [V, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_uint32, C.c_int]
    L.orc_merge.restype = None
    L.orc_resplit.argtypes = [V, C.c_uint32, C.c_uint32, _u32p, _u32p, C.c_size_t, C.c_uint32]
    L.orc_resplit.restype = None
    L.orc_solid_count.argtypes = [V]
    L.orc_solid_count.restype = C.c_size_t
    L.orc_solid_get.argtypes = [V, _u32p, _u64p, _u64p, _u32p, _u8p]
    L.orc_solid_get.restype = None
    L.orc_times.argtypes = [V, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.orc_times.restype = None
    L.orc_free.argtypes = [V]
    L.orc_free.restype = None
    L.orc_kmer_keys.argtypes = [_u8p, C.c_size_t, C.c_uint, C.c_int, _u64p, _u64p, _u8p]
    L.orc_kmer_keys.restype = None
    L.orc_bruteforce_counts.argtypes = [_u8p, _u64p, C.c_size_t, _u32p, C.c_uint, C.c_uint32]
    L.orc_query.argtypes = [V, _u8p, C.c_uint64, C.c_uint, C.c_int, C.c_uint32, _u64p, _u64p, _u64p, C.c_uint64]
    L.orc_query.restype = C.c_uint64
    L.orc_extension.argtypes = [V, C.c_uint32, _u8p]
    L.orc_extension.restype = None
    L.orc_unitigs.argtypes = [V, C.c_uint32, C.c_int, C.c_int]
    L.orc_unitigs.restype = C.c_size_t
    L.orc_unitigs_bytes.argtypes = [V]
    L.orc_unitigs_bytes.restype = C.c_size_t
    L.orc_unitigs_get.argtypes = [V, _u8p, _u32p, _u8p]
    L.orc_unitigs_get.restype = None
    L.orc_rec_colors.argtypes = [V, _u32p]
    L.orc_rec_colors.restype = None
    L.orc_solid_colors_total.argtypes = [V]
    L.orc_solid_colors_total.restype = C.c_size_t
    L.orc_solid_colors_get.argtypes = [V, _u32p, _u32p]
    L.orc_solid_colors_get.restype = None
    L.orc_merge_sorting.argtypes = [V, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_uint32, C.c_int,
                                    _u64p]
    L.orc_merge_sorting.restype = None
    L.orc_sorting_count.argtypes = [V]
    L.orc_sorting_count.restype = C.c_size_t
    L.orc_sorting_get.argtypes = [V, _u32p, _u64p, _u64p, _u32p, _u8p]
    L.orc_sorting_get.restype = None
    L.orc_bruteforce_counts.restype = V
    _lib = L
    return L


def _p(a, t):
    return a.ctypes.data_as(t)


def nthash(seq: bytes | np.ndarray, m: int) -> tuple[np.ndarray, np.ndarray]:
    s = np.frombuffer(bytes(seq), np.uint8) if not isinstance(seq, np.ndarray) else seq
    n = max(0, s.size - m + 1)
    fw, rc = np.zeros(n, np.uint64), np.zeros(n, np.uint64)
    if n:
        lib().orc_nthash(_p(s, _u8p), s.size, m, _p(fw, _u64p), _p(rc, _u64p))
    return fw, rc


def minqueue(v: np.ndarray, index: np.ndarray, fwd: np.ndarray, w: int):
    v = np.ascontiguousarray(v, np.uint64)
    index = np.ascontiguousarray(index, np.uint32)
    fwd = np.ascontiguousarray(fwd, np.uint8)
    n = v.size
    ov, oi, of = np.zeros(n, np.uint64), np.zeros(n, np.uint32), np.zeros(n, np.uint8)
    cnt = lib().orc_minqueue(_p(v, _u64p), _p(index, _u32p), _p(fwd, _u8p), n, w, _p(ov, _u64p), _p(oi, _u32p),
                             _p(of, _u8p))
    return ov[:cnt], oi[:cnt], of[:cnt]


class Result:
    def __init__(self, h):
        self.h = h

    def __del__(self):
        try:
            lib().orc_free(self.h)
        except Exception:
            pass

    def compact(self) -> None:
        lib().orc_compact(self.h)

    def records(self) -> dict:
        L = lib()
        n = L.orc_rec_count(self.h)
        nb = L.orc_rec_bases(self.h)
        r = dict(sub=np.empty(n, np.uint32), length=np.empty(n, np.uint32), flags=np.empty(n, np.uint8),
                 minpos=np.empty(n, np.uint16), mult=np.empty(n, np.uint32), off=np.empty(n, np.uint64),
                 bases=np.empty(nb, np.uint8))
        L.orc_rec_get(self.h, _p(r["sub"], _u32p), _p(r["length"], _u32p), _p(r["flags"], _u8p),
                      _p(r["minpos"], _u16p), _p(r["mult"], _u32p), _p(r["off"], _u64p), _p(r["bases"], _u8p))
        r["color"] = np.zeros(n, np.uint32)
        L.orc_rec_colors(self.h, _p(r["color"], _u32p))
        return r

    def merge(self, k: int, b: int, sb: int, kind: int, min_mult: int, threads: int = 1) -> None:
        lib().orc_merge(self.h, k, b, sb, kind, min_mult, threads)

    def resplit(self, k: int, m: int, outliers: list[tuple[int, int]], first_extra: int) -> None:
        """Replaces the records of each (sub-bucket, resplit bucket count)
        outlier by its resplit records, numbered from first_extra (each
        outlier's count normal buckets, then its duplicates bucket)."""
        subs = np.array([o[0] for o in outliers], np.uint32)
        cnts = np.array([o[1] for o in outliers], np.uint32)
        lib().orc_resplit(self.h, k, m, _p(subs, _u32p), _p(cnts, _u32p), len(outliers), first_extra)

    def merge_sorting(self, k: int, m: int, b: int, sb: int, kind: int, min_mult: int, threads: int = 1) -> dict:
        """The sorting extender's rule on the normal sub-buckets: entries
        (sub, key_lo, key_hi, mult, flags = extra bases, bit 0 backward, bit 1
        forward) and stats (both_extra, groups, centred)."""
        L = lib()
        st = np.zeros(3, np.uint64)
        L.orc_merge_sorting(self.h, k, m, b, sb, kind, min_mult, threads, _p(st, _u64p))
        n = L.orc_sorting_count(self.h)
        s = dict(sub=np.empty(n, np.uint32), key_lo=np.empty(n, np.uint64), key_hi=np.empty(n, np.uint64),
                 mult=np.empty(n, np.uint32), flags=np.empty(n, np.uint8))
        L.orc_sorting_get(self.h, _p(s["sub"], _u32p), _p(s["key_lo"], _u64p), _p(s["key_hi"], _u64p),
                          _p(s["mult"], _u32p), _p(s["flags"], _u8p))
        s["stats"] = dict(both_extra=int(st[0]), groups=int(st[1]), centred=int(st[2]))
        return s

    def solid(self) -> dict:
        L = lib()
        n = L.orc_solid_count(self.h)
        s = dict(sub=np.empty(n, np.uint32), key_lo=np.empty(n, np.uint64), key_hi=np.empty(n, np.uint64),
                 mult=np.empty(n, np.uint32), flags=np.empty(n, np.uint8))
        L.orc_solid_get(self.h, _p(s["sub"], _u32p), _p(s["key_lo"], _u64p), _p(s["key_hi"], _u64p),
                        _p(s["mult"], _u32p), _p(s["flags"], _u8p))
        return s

    def solid_colors(self) -> tuple[np.ndarray, np.ndarray]:
        """Per solid entry (solid() order): offsets [n+1] into sorted color lists."""
        L = lib()
        n = L.orc_solid_count(self.h)
        off = np.zeros(n + 1, np.uint32)
        cols = np.zeros(L.orc_solid_colors_total(self.h), np.uint32)
        L.orc_solid_colors_get(self.h, _p(off, _u32p), _p(cols, _u32p))
        return off, cols

    def query(self, queries: np.ndarray, k: int, kind: int = 0, n_colors: int = 0):
        """Per query of a joined buffer: (query_kmers, matched, color_counts or None)
        against this handle's solid k-mers (after merge())."""
        qd = np.ascontiguousarray(queries, np.uint8)
        nq = int(np.count_nonzero(qd == 10))
        qk, mt = np.zeros(nq, np.uint64), np.zeros(nq, np.uint64)
        cc = np.zeros(max(1, nq * n_colors), np.uint64)
        got = lib().orc_query(self.h, _p(qd, _u8p), qd.size, k, kind, n_colors, _p(qk, _u64p), _p(mt, _u64p),
                              _p(cc, _u64p), nq)
        assert got == nq
        return qk, mt, (cc[:nq * n_colors].reshape(nq, n_colors) if n_colors else None)

    def extension(self, k: int) -> np.ndarray:
        """In/out extension flags per solid entry (solid() order)."""
        L = lib()
        out = np.zeros(L.orc_solid_count(self.h), np.uint8)
        L.orc_extension(self.h, k, _p(out, _u8p))
        return out

    def unitigs(self, k: int, sequential: bool = False, joined: bool = False, threads: int = 1):
        """Intra-sub-bucket unitigs of the merged table: (list of ASCII sequences,
        sub-bucket per unitig, flags: bit0 open start, bit1 open end, bit2 circular).
        joined=True: the maximal unitigs of the whole solid k-mer set (one
        graph: single-threaded); otherwise sub-buckets go to `threads` workers."""
        L = lib()
        n = L.orc_unitigs(self.h, k, 2 if joined else (1 if sequential else 0), threads)
        seq = np.zeros(L.orc_unitigs_bytes(self.h), np.uint8)
        sub, fl = np.zeros(n, np.uint32), np.zeros(n, np.uint8)
        L.orc_unitigs_get(self.h, _p(seq, _u8p), _p(sub, _u32p), _p(fl, _u8p))
        return seq.tobytes().split(b"\n")[:-1], sub, fl

    def times(self) -> tuple[float, float]:
        a, b = C.c_double(), C.c_double()
        lib().orc_times(self.h, C.byref(a), C.byref(b))
        return a.value, b.value


def bucketing(data: np.ndarray, k: int, m: int, b: int, sb: int, canonical: bool = True, threads: int = 1,
              rec_off: np.ndarray | None = None, rec_color: np.ndarray | None = None) -> Result:
    """Joined input: by default the whole buffer is one record (separators act as N)."""
    data = np.ascontiguousarray(data, np.uint8)
    if rec_off is None:
        # split the joined buffer into `threads` record ranges at separators
        cuts = [0]
        for t in range(1, threads):
            c = data.size * t // threads
            while c < data.size and data[c] != ord("\n"):
                c += 1
            cuts.append(min(c + 1, data.size))
        cuts.append(data.size)
        rec_off = np.array(sorted(set(cuts)), np.uint64)
    rec_off = np.ascontiguousarray(rec_off, np.uint64)
    rc = None if rec_color is None else np.ascontiguousarray(rec_color, np.uint32)
    h = lib().orc_bucketing(_p(data, _u8p), _p(rec_off, _u64p), rec_off.size - 1,
                            None if rc is None else _p(rc, _u32p), k, m, b, sb, 1 if canonical else 0, threads)
    return Result(h)


def kmer_keys(seq: bytes, k: int, kind: int = 0):
    s = np.frombuffer(bytes(seq), np.uint8).copy()
    n = max(0, s.size - k + 1)
    lo, hi, fw = np.zeros(n, np.uint64), np.zeros(n, np.uint64), np.zeros(n, np.uint8)
    if n:
        lib().orc_kmer_keys(_p(s, _u8p), s.size, k, kind, _p(lo, _u64p), _p(hi, _u64p), _p(fw, _u8p))
    return lo, hi, fw


def color_records(n_bytes: int, color_starts, colors) -> tuple[np.ndarray, np.ndarray]:
    """Record ranges and record colors for a joined input whose color ranges
    start at the byte offsets color_starts (the C ABI's color layout)."""
    cs = np.asarray(color_starts, np.uint64)
    rec_off = np.concatenate([cs, np.array([n_bytes], np.uint64)])
    return rec_off, np.ascontiguousarray(colors, np.uint32)


def bucketing_colored(data: np.ndarray, k: int, m: int, b: int, sb: int, color_starts, colors,
                      canonical: bool = True, threads: int = 1) -> Result:
    rec_off, rec_color = color_records(np.asarray(data).size, color_starts, colors)
    return bucketing(data, k, m, b, sb, canonical, threads, rec_off, rec_color)


def bruteforce_counts(data: np.ndarray, k: int, min_mult: int, color_starts=None, colors=None):
    """Exact canonical counts (and, with colors, color sets) over the input;
    returns the solid dict, plus (offsets, colors) when colored."""
    data = np.ascontiguousarray(data, np.uint8)
    if color_starts is None:
        off = np.array([0, data.size], np.uint64)
        r = Result(lib().orc_bruteforce_counts(_p(data, _u8p), _p(off, _u64p), 1, None, k, min_mult))
        return r.solid()
    off, rc = color_records(data.size, color_starts, colors)
    r = Result(lib().orc_bruteforce_counts(_p(data, _u8p), _p(off, _u64p), off.size - 1, _p(rc, _u32p), k, min_mult))
    return r.solid(), r.solid_colors()
