"""Maximal unitigs of the whole solid set (the partial unitigs of all buckets
joined, as build_final_unitigs outputs them): known answers for constructed
genomes (a single genome, a SNP bubble, a circular genome), the textbook
definition checked by brute force on noisy reads, and the device path and its
FASTA writer on the CPU emulation."""
from __future__ import annotations

import re

import numpy as np
import pytest

from ggcat_amd import synth
from oracle import binding as O
from tests.test_unitigs import COMP, canon_unitigs

K = 31
HASH_SEQ = 0


def _rc(s: bytes) -> bytes:
    return s.translate(COMP)[::-1]


def _canon(s: bytes) -> bytes:
    return min(s, _rc(s))


def _random_genome(seed: int, n: int) -> bytes:
    rng = np.random.default_rng(seed)
    return np.frombuffer(b"ACGT", np.uint8)[rng.integers(0, 4, n)].tobytes()


def _distinct_kmers(*seqs: bytes) -> bool:
    seen = set()
    for s in seqs:
        for i in range(len(s) - K + 1):
            c = _canon(s[i:i + K])
            if c in seen:
                return False
            seen.add(c)
    return True


def _tiled(seqs, read_len: int = 100, step: int = 7) -> np.ndarray:
    """Error-free reads tiling each sequence on both strands, the first and
    last window included, one read per line."""
    out = []
    for s in seqs:
        starts = list(range(0, len(s) - read_len + 1, step))
        if starts[-1] != len(s) - read_len:
            starts.append(len(s) - read_len)
        for i in starts:
            out += [s[i:i + read_len], _rc(s[i:i + read_len])]
    return np.frombuffer(b"\n".join(out) + b"\n", np.uint8).copy()


def _joined(data: np.ndarray, min_mult: int = 1, b: int = 3, sb: int = 1):
    r = O.bucketing(data, K, 15, b, sb, threads=2)
    r.merge(K, b, sb, HASH_SEQ, min_mult, 2)
    return r, r.unitigs(K, joined=True)


def test_single_genome_is_one_unitig():
    g = _random_genome(11, 3000)
    assert _distinct_kmers(g)
    r, (seqs, subs, fl) = _joined(_tiled([g]))
    assert len(set(r.solid()["sub"].tolist())) > 1  # the genome's k-mers span several sub-buckets
    assert [_canon(s) for s in seqs] == [_canon(g)]
    assert fl.tolist() == [0] and subs.tolist() == [0]


def test_snp_bubble_gives_four_unitigs():
    g = _random_genome(12, 4000)
    p = 2000
    alt = {ord("A"): b"C", ord("C"): b"G", ord("G"): b"T", ord("T"): b"A"}[g[p]]
    h = g[:p] + alt + g[p + 1:]
    assert _distinct_kmers(g, h[p - K + 1:p + K])
    _, (seqs, _, fl) = _joined(_tiled([g, h]))
    want = sorted(_canon(x) for x in (g[:p], g[p - K + 1:p + K], h[p - K + 1:p + K], g[p + 1:]))
    assert sorted(_canon(s) for s in seqs) == want
    assert fl.tolist() == [0] * 4


def test_circular_genome_is_one_cycle():
    g = _random_genome(13, 2500)
    assert _distinct_kmers(g + g[:K - 1])
    circ = g + g[:200]  # every read is a substring of the circular sequence
    data = _tiled([circ])
    _, (seqs, _, fl) = _joined(data)
    assert fl.tolist() == [4] and len(seqs) == 1
    s = seqs[0]
    n = len(g)
    assert len(s) == n + K - 1
    assert s[n:] == s[:K - 1]  # closes on itself
    rot = s[:n]
    assert rot in (g + g) or rot in (_rc(g) + _rc(g))


def _succ(x: bytes, S) -> list[bytes]:
    return [x[1:] + c for c in (b"A", b"C", b"G", b"T") if _canon(x[1:] + c) in S]


def _pred(x: bytes, S) -> list[bytes]:
    return [c + x[:-1] for c in (b"A", b"C", b"G", b"T") if _canon(c + x[:-1]) in S]


def _kmer_bytes(lo: int) -> bytes:
    return bytes(b"ACTG"[(lo >> (2 * j)) & 3] for j in range(K))


@pytest.mark.parametrize("err,rep", [(0.005, 0.0), (0.005, 0.3)])
def test_joined_unitigs_are_the_maximal_unitigs(err, rep):
    """Textbook compaction, checked directly on the solid k-mer set: every
    solid k-mer lies on exactly one unitig; consecutive k-mers x, y of a
    unitig satisfy succ(x) = {y} and pred(y) = {x} with y not x in either
    orientation; no unitig end can be extended that way; a circular unitig
    closes on itself."""
    g = synth.genome(81, 20_000, repeat_frac=rep, rep_min=300, rep_max=1000)
    data = synth.reads(g, 82, 4000, 150, err)
    r, (seqs, _, fl) = _joined(data, min_mult=2)
    S = {_kmer_bytes(lo) for lo in r.solid()["key_lo"].tolist()}
    S = {_canon(x) for x in S}
    seen = []
    for s, f in zip(seqs, fl.tolist()):
        km = [s[i:i + K] for i in range(len(s) - K + 1)]
        if f & 4:
            km = km[:len(s) - 2 * (K - 1)]  # the last k-1 bases repeat the first
        seen += [_canon(x) for x in km]
        chain = km + ([km[0]] if f & 4 else [])
        for x, y in zip(chain, chain[1:]):
            assert _succ(x, S) == [y] and _pred(y, S) == [x] and _canon(x) != _canon(y)
        if not f & 4:
            last, first = km[-1], km[0]
            sx = _succ(last, S)
            assert not (len(sx) == 1 and _pred(sx[0], S) == [last] and _canon(sx[0]) != _canon(last))
            px = _pred(first, S)
            assert not (len(px) == 1 and _succ(px[0], S) == [first] and _canon(px[0]) != _canon(first))
    assert sorted(seen) == sorted(S)


def test_joined_unitigs_emulated(tmp_path):
    from ggcat_amd.native import GpuContext, Params
    from tests.emu import build as emu_build
    ctx = GpuContext(0, lib_path=emu_build.build())
    try:
        ctx.set_tuning(merge_wgs_per_cu=2)
        g = synth.genome(83, 12_000, repeat_frac=0.2, rep_min=300, rep_max=1000)
        circ = _random_genome(84, 1500)
        data = np.concatenate([synth.reads(g, 85, 800, 150, 0.005), _tiled([circ + circ[:200]])])
        ctx.bucketing(data, Params(K, 15, 3, 1, 0, HASH_SEQ, 2))
        ctx.kmers_merge()
        gu = ctx.unitigs(joined=True)
        o = O.bucketing(data, K, 15, 3, 1, threads=2)
        o.merge(K, 3, 1, HASH_SEQ, 2, 2)
        want = canon_unitigs(*o.unitigs(K, joined=True))
        assert canon_unitigs(*gu) == want
        assert any(f == 4 for _, _, f in want)
        # the per-sub-bucket result is unchanged after a joined build and back
        assert canon_unitigs(*ctx.unitigs()) == canon_unitigs(*o.unitigs(K))
        gu = ctx.unitigs(joined=True)
        path = str(tmp_path / "unitigs.fa")
        assert ctx.write_unitigs_fasta(path) == len(gu[0])
        lines = open(path, "rb").read().split(b"\n")
        assert lines[-1] == b""
        for i, s in enumerate(gu[0]):
            m = re.fullmatch(rb">(\d+) LN:i:(\d+)", lines[2 * i])
            assert m and int(m.group(1)) == i and int(m.group(2)) == len(s) and lines[2 * i + 1] == s
    finally:
        ctx.close()
