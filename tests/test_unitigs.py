"""Unitig chains inside each sub-bucket: the order-free rule the device
implements against the reference's own greedy loop (both restated in the
oracle), the k-mer cover of the chains, and the kernels on the CPU emulation."""
from __future__ import annotations

import numpy as np
import pytest

from ggcat_amd import synth
from oracle import binding as O

COMP = bytes.maketrans(b"ACGT", b"TGCA")


def canon_unitigs(seqs, subs, fls):
    """Orientation-free form: a linear unitig as the smaller of itself and its
    reverse complement (open-end flags swapped with it); cycles as written."""
    out = []
    for s, sb, f in zip(seqs, subs.tolist(), fls.tolist()):
        rc = s.translate(COMP)[::-1]
        if not (f & 4) and rc < s:
            s, f = rc, (f & 4) | ((f & 1) << 1) | ((f >> 1) & 1)
        out.append((sb, s, f))
    return sorted(out)


@pytest.mark.parametrize("err,rep", [(0.0, 0.0), (0.01, 0.0), (0.0, 0.3), (0.005, 0.2)])
def test_order_free_rule_equals_reference_loop(err, rep):
    g = synth.genome(73, 40_000, repeat_frac=rep, rep_min=300, rep_max=1000)
    data = synth.reads(g, 74, 10_000, 150, err)
    r = O.bucketing(data, 31, 15, 3, 1, threads=2)
    r.merge(31, 3, 1, 0, 2, 2)
    assert canon_unitigs(*r.unitigs(31)) == canon_unitigs(*r.unitigs(31, sequential=True))


def test_unitigs_cover_solid_kmers_once():
    g = synth.genome(75, 30_000, repeat_frac=0.2, rep_min=300, rep_max=1000)
    data = synth.reads(g, 76, 6000, 150, 0.005)
    k = 31
    r = O.bucketing(data, k, 15, 3, 1, threads=2)
    r.merge(k, 3, 1, 0, 2, 2)
    seqs, subs, _ = r.unitigs(k)
    sol = r.solid()
    want = sorted(zip(sol["sub"].tolist(), sol["key_lo"].tolist()))
    code = {ord("A"): 0, ord("C"): 1, ord("T"): 2, ord("G"): 3}
    got = []
    for s, sb in zip(seqs, subs.tolist()):
        for i in range(len(s) - k + 1):
            km = s[i:i + k]
            f = sum(code[c] << (2 * j) for j, c in enumerate(km))
            rc = km.translate(COMP)[::-1]
            b = sum(code[c] << (2 * j) for j, c in enumerate(rc))
            got.append((sb, min(f, b)))
    assert sorted(got) == want


def _kmer_str(lo: int, k: int) -> bytes:
    # seq-hash key: base j at bits 2j, A0 C1 T2 G3
    return bytes(b"ACTG"[(lo >> (2 * j)) & 3] for j in range(k))


@pytest.mark.parametrize("err,rep", [(0.0, 0.2), (0.01, 0.0)])
def test_unitigs_are_maximal_nonbranching_paths(err, rep):
    """Pins the chain rule from its definition, independently of both
    restatements: consecutive k-mers of a unitig are each other's only
    neighbour in the sub-bucket (one base shift, either strand) with neither
    facing side flagged (flags exactly INCL_BEGIN / INCL_END), and no unitig
    end can be extended that way."""
    k = 21
    g = synth.genome(91, 8000, repeat_frac=rep, rep_min=200, rep_max=600)
    data = synth.reads(g, 92, 1500, 120, err)
    r = O.bucketing(data, k, 9, 2, 1, threads=2)
    r.merge(k, 2, 1, 0, 2, 2)
    sol = r.solid()
    canon = {}  # (sub, canonical k-mer) -> flags
    for sb, lo, fl in zip(sol["sub"].tolist(), sol["key_lo"].tolist(), sol["flags"].tolist()):
        canon[(sb, _kmer_str(lo, k))] = fl

    def rc(x):
        return x.translate(COMP)[::-1]

    code = {ord("A"): 0, ord("C"): 1, ord("T"): 2, ord("G"): 3}

    def cmin(x):
        # canonical orientation: the smaller packed key (base j at bits 2j)
        y = rc(x)
        vx = sum(code[c] << (2 * j) for j, c in enumerate(x))
        vy = sum(code[c] << (2 * j) for j, c in enumerate(y))
        return x if vx <= vy else y

    def key(sb, x):
        c = cmin(x)
        return (sb, c) if (sb, c) in canon else None

    def flagged(sb, x, succ):
        # side of x (as written) facing its successor (succ) or predecessor
        c = cmin(x)
        side = (1 if succ else 0) if x == c else (0 if succ else 1)
        fl = canon[(sb, c)]
        return (fl == 1 and side == 0) or (fl == 2 and side == 1)

    def nbrs(sb, x, succ):
        out = []
        for b in b"ACGT":
            y = x[1:] + bytes([b]) if succ else bytes([b]) + x[:-1]
            if key(sb, y):
                out.append(y)
        return out

    def linked(sb, x, y):
        # x followed by y, one base shift
        if key(sb, x) == key(sb, y):
            return False
        if flagged(sb, x, True) or flagged(sb, y, False):
            return False
        return nbrs(sb, x, True) == [y] and nbrs(sb, y, False) == [x]

    seqs, subs, fls = r.unitigs(k)
    assert len(seqs) > 0
    n_kmers = 0
    for s, sb, f in zip(seqs, subs.tolist(), fls.tolist()):
        kms = [s[i:i + k] for i in range(len(s) - k + 1)]
        n_kmers += len(kms)
        for x, y in zip(kms, kms[1:]):
            assert linked(sb, x, y)
        if f & 4:
            continue  # circular: walked from its smallest k-mer
        for y in nbrs(sb, kms[-1], True):
            assert not linked(sb, kms[-1], y)
        for y in nbrs(sb, kms[0], False):
            assert not linked(sb, y, kms[0])
        # open-end flags: the end side is flagged exactly when bit 0 / bit 1 says so
        assert bool(f & 1) == flagged(sb, kms[0], False)
        assert bool(f & 2) == flagged(sb, kms[-1], True)
    assert n_kmers == len(canon)
    # extension flags: neighbour counts of each solid k-mer in its canonical
    # orientation, predecessors in bits 0-1 and successors in bits 2-3 (2 = more)
    ext = r.extension(k)
    for sb, lo, e in zip(sol["sub"].tolist(), sol["key_lo"].tolist(), ext.tolist()):
        x = _kmer_str(lo, k)
        d0, d1 = len(nbrs(sb, x, False)), len(nbrs(sb, x, True))
        assert e == min(d0, 2) | (min(d1, 2) << 2)


def test_unitigs_and_extension_emulated():
    from ggcat_amd.native import GpuContext, Params
    from tests.emu import build as emu_build
    ctx = GpuContext(0, lib_path=emu_build.build())
    try:
        ctx.set_tuning(merge_wgs_per_cu=2)
        g = synth.genome(73, 12_000, repeat_frac=0.2, rep_min=300, rep_max=1000)
        data = synth.reads(g, 74, 1500, 150, 0.005)
        ctx.bucketing(data, Params(31, 15, 3, 1, 0, 0, 2))
        ctx.kmers_merge()
        gu = ctx.unitigs()
        ext = ctx.extension_flags()
        s = ctx.solid()
        o = O.bucketing(data, 31, 15, 3, 1, threads=2)
        o.merge(31, 3, 1, 0, 2, 2)
        assert canon_unitigs(*gu) == canon_unitigs(*o.unitigs(31))
        osol = o.solid()
        assert sorted(zip(s.sub.tolist(), s.keys[:, 0].tolist(), ext.tolist())) == \
            sorted(zip(osol["sub"].tolist(), osol["key_lo"].tolist(), o.extension(31).tolist()))
    finally:
        ctx.close()


@pytest.mark.parametrize("sequential", [False, True])
def test_unitigs_threads_same_output(sequential):
    # sub-buckets are independent: worker threads give the same unitigs, in
    # the same order, as one thread
    g = synth.genome(44, 60_000, repeat_frac=0.2, rep_min=300, rep_max=1500)
    data = synth.reads(g, 45, 6_000, 150, 0.005)
    r = O.bucketing(data, 31, 15, 3, 1, threads=2)
    r.merge(31, 3, 1, 0, 2, 2)
    one = r.unitigs(31, sequential=sequential)
    many = r.unitigs(31, sequential=sequential, threads=5)
    assert one[0] == many[0] and np.array_equal(one[1], many[1]) and np.array_equal(one[2], many[2])
    assert len(one[0]) > 100
