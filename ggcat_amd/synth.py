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


def descendant(anc: np.ndarray, seed: int, snp: float, indel: float, target_len: int | None = None,
               stream: int = 0) -> np.ndarray:
    """Ancestor copy with substitutions (rate snp) and 1-50 bp insertions or
    deletions (rate indel); a longer target appends mutated ancestor segments,
    a shorter one keeps a random window of the mutated copy."""
    target = anc.size if target_len is None else int(target_len)
    cap = max(target, 1)
    out = np.empty(cap, np.uint8)
    n = synth_lib().synth_descendant(_ptr(anc, _u8p), anc.size, seed, stream, snp, indel, target, _ptr(out, _u8p),
                                     cap)
    return out[:n]


def lognormal_len(seed: int, stream: int, median: float, sigma: float, lo: float, hi: float) -> int:
    return int(synth_lib().synth_lognormal_len(seed, stream, median, sigma, lo, hi))


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


def config3_genomes(scale: float = 1.0, n_genomes: int = 256) -> list[np.ndarray]:
    """256 genomes from one 4 Mbp ancestor (seed 3): 1% substitutions and 0.1%
    1-50 bp indels each; lengths log-normal (median 2 Mbp, sigma 0.8) clipped
    to 0.5-10 Mbp -- longer genomes append mutated ancestor segments, shorter
    ones are a random window of the mutated ancestor."""
    from concurrent.futures import ThreadPoolExecutor
    A = int(4_000_000 * scale)
    anc = genome(3, A)
    lens = [lognormal_len(3, i, 2_000_000 * scale, 0.8, 500_000 * scale, 10_000_000 * scale) for i in range(n_genomes)]
    with ThreadPoolExecutor(THREADS) as ex:
        return list(ex.map(lambda i: descendant(anc, 3, 0.01, 0.001, lens[i], stream=1 + i), range(n_genomes)))


def contigs_joined(genomes: list[np.ndarray], seed: int, lo: int = 20_000, hi: int = 100_000):
    """Each genome cut into error-free contigs (lengths uniform in [lo, hi]);
    joined layout plus the byte offset where each genome (color) starts."""
    rng = np.random.default_rng(seed)
    parts, starts, pos = [], [], 0
    for g in genomes:
        starts.append(pos)
        cuts = [0]
        while cuts[-1] < g.size:
            cuts.append(min(g.size, cuts[-1] + int(rng.integers(lo, hi + 1))))
        buf = np.empty(g.size + len(cuts) - 1, np.uint8)
        o = 0
        for a, b in zip(cuts, cuts[1:]):
            buf[o:o + b - a] = g[a:b]
            buf[o + b - a] = 10
            o += b - a + 1
        parts.append(buf)
        pos += buf.size
    return np.concatenate(parts) if parts else np.zeros(0, np.uint8), np.array(starts, np.uint64)


def config3(scale: float = 1.0, n_genomes: int = 256) -> Workload:
    """Colored build: one color per genome file, k=31, m=15, s=2 (the CLI
    default; the reference's C++ API defaults to 1)."""
    gs = config3_genomes(scale, n_genomes)
    data, starts = contigs_joined(gs, 303)
    bases = int(sum(g.size for g in gs))
    b, sb = auto_bucket_logs(int(data.size + (data == 10).sum() * 10))
    colors = np.arange(len(gs), dtype=np.uint32)
    return Workload(f"config3-{len(gs)}genomes-colored" + ("" if scale == 1.0 else f"-x{scale}"), data, 31, 15, 2,
                    bases, b, sb, colors=(starts, colors))


def queries(genomes: list[np.ndarray], seed: int, n: int, lmin: int = 100, lmax: int = 10_000, mut: float = 0.02,
            stream: int = 0) -> np.ndarray:
    """n queries in the joined layout: lengths log-uniform in [lmin, lmax]; half
    are substrings of the genomes with substitution rate mut, half i.i.d."""
    cat = np.concatenate(genomes) if genomes else np.zeros(1, np.uint8)
    off = np.zeros(len(genomes) + 1, np.uint64)
    off[1:] = np.cumsum([g.size for g in genomes])
    out = np.empty(n * (lmax + 1), np.uint8)
    lens = np.empty(max(n, 1), np.uint32)
    nb = synth_lib().synth_queries(_ptr(cat, _u8p), _ptr(off, _u64p), len(genomes), seed, stream, n, lmin, lmax, mut,
                                   _ptr(out, _u8p), _ptr(lens, _u32p))
    return out[:nb].copy()


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
