"""Seeded synthetic workloads shaped like the benchmark configurations.

Every read set is returned in the joined layout (record bytes + b'\\n').
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

from .native import _ptr, _u8p, _u32p, _u64p, synth_lib

THREADS = max(1, min(16, os.cpu_count() or 1))


def genome(seed: int, length: int, repeat_frac: float = 0.0, rep_min: int = 2000, rep_max: int = 8000,
           max_div: float = 0.03) -> np.ndarray:
    g = np.empty(length, np.uint8)
    synth_lib().synth_genome(seed, length, repeat_frac, rep_min, rep_max, max_div, _ptr(g, _u8p), THREADS)
    return g


def reads(g: np.ndarray, seed: int, n_reads: int, read_len: int, err: float) -> np.ndarray:
    out = np.empty(n_reads * (read_len + 1), np.uint8)
    synth_lib().synth_reads(_ptr(g, _u8p), g.size, seed, n_reads, read_len, err, _ptr(out, _u8p), THREADS)
    return out


@dataclass
class Workload:
    name: str
    data: np.ndarray  # joined layout
    k: int
    m: int
    min_mult: int
    bases: int  # ACGT bases in records (excludes separators)
    buckets_log: int
    second_buckets_log: int
    hash_kind: int = 0
    colors: tuple | None = None


def auto_bucket_logs(estimated_bases: int) -> tuple[int, int]:
    """compute_stats_from_input_blocks (crates/io/src/lib.rs:67-140) with the
    input size as the base estimate (crates/io/src/sequences_stream/fasta.rs:12-30)."""
    MIN_BUCKET_SIZE, MAX_BUCKET_SIZE = 512 * 1024, 1024 ** 3

    def npow2(x: int) -> int:
        return 1 if x <= 1 else 1 << (x - 1).bit_length()

    b = npow2(max(min(1 << 10, estimated_bases // MIN_BUCKET_SIZE), estimated_bases // MAX_BUCKET_SIZE))
    b = min(max(b, 1 << 2), 1 << 13)
    per = estimated_bases // b
    sb = npow2(max(min(1 << 6, per // (2 * 1024)), per // (4 * 1024 * 1024)))
    sb = min(max(sb, 1 << 1), 1 << 8)
    return b.bit_length() - 1, sb.bit_length() - 1


def config1(scale: float = 1.0) -> Workload:
    """5 Mbp i.i.d. genome, 150 bp reads at 30x, 0.5% substitutions, both
    strands; k=31, m=15, s=2, b=9 (scale < 1 shrinks genome and reads)."""
    G = int(5_000_000 * scale)
    n = int(1_000_000 * scale)
    g = genome(1, G)
    d = reads(g, 101, n, 150, 0.005)
    _, sb = auto_bucket_logs(d.size + n * 4)  # FASTA with ">r<i>\n" headers
    return Workload("config1-5Mbp-150bp-30x" + ("" if scale == 1.0 else f"-x{scale}"), d, 31, 15, 2, n * 150, 9, sb)


def config2(scale: float = 1.0) -> Workload:
    """100 Mbp genome with 20% 2-8 kbp repeat copies (0-3% divergence), 250 bp
    reads at 40x, 1% substitutions, random strand; k=31, m=15, s=2, b auto."""
    G = int(100_000_000 * scale)
    n = int(16_000_000 * scale)
    g = genome(2, G, repeat_frac=0.2)
    d = reads(g, 202, n, 250, 0.01)
    b, sb = auto_bucket_logs(d.size + n * 10)
    return Workload("config2-100Mbp-250bp-40x" + ("" if scale == 1.0 else f"-x{scale}"), d, 31, 15, 2, n * 250, b, sb)


def config4(scale: float = 1.0) -> Workload:
    """500 Mbp genome with 30% repeats, 150 bp paired reads at 30x, 0.5%
    substitutions; k=63, m=21, s=2, rabin-karp128."""
    G = int(500_000_000 * scale)
    n = int(100_000_000 * scale)
    g = genome(4, G, repeat_frac=0.3)
    d = reads(g, 404, n, 150, 0.005)
    b, sb = auto_bucket_logs(d.size + n * 10)
    return Workload("config4-500Mbp-150bp-30x-k63" + ("" if scale == 1.0 else f"-x{scale}"), d, 63, 21, 2,
                    n * 150, b, sb, hash_kind=1)
