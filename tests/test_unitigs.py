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
