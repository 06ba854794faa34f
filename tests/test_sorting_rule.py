"""The reference counts normal sub-buckets with the sorting extender
(crates/assembler_kmers_merge/src/unitigs_extender/sorting.rs:641-1050, on
records grouped by minimizer, crates/assembler_kmers_merge/src/map_processor.rs:127-166),
not with the hash map the device and the oracle's merge restate
(crates/assembler_kmers_merge/src/unitigs_extender/hashmap.rs:361-431).
These tests pin the oracle's restatement of the sorting rule
(merge_subbucket_sorting) and show how its output relates to the hash-map
rule's solid table:

* the (sub-bucket, k-mer, multiplicity) sets are equal;
* the sorting rule holds a k-mer twice exactly when the k-mer is a group
  boundary whose two groups fall in the same sub-bucket (the hash-map entry
  then has both flag bits set and is halved); each copy carries the same
  multiplicity and one extra base;
* no centred k-mer ever holds both extra bases, so the reference's halving
  of a one-k-mer supertig (sorting.rs:454-486) never applies on these inputs.
"""
from __future__ import annotations

from collections import defaultdict

import numpy as np
import pytest

from ggcat_amd import synth
from oracle import binding as O
from tests.helpers import adversarial_records, joined


def _compare(data, k, m, b, sb, s, kind=0, compact=False):
    o = O.bucketing(data, k, m, b, sb, threads=4)
    if compact:
        o.compact()
    o.merge(k, b, sb, kind, s, 4)
    h = o.solid()
    st = o.merge_sorting(k, m, b, sb, kind, s, 4)
    assert st["stats"]["both_extra"] == 0
    normal = (1 << b) << sb
    sel = h["sub"] < normal
    H = {}
    for a, x, y, mu, f in zip(h["sub"][sel].tolist(), h["key_lo"][sel].tolist(), h["key_hi"][sel].tolist(),
                              h["mult"][sel].tolist(), h["flags"][sel].tolist()):
        H[(a, x, y)] = (mu, f)
    S = defaultdict(list)
    for a, x, y, mu, f in zip(st["sub"].tolist(), st["key_lo"].tolist(), st["key_hi"].tolist(), st["mult"].tolist(),
                              st["flags"].tolist()):
        S[(a, x, y)].append((mu, f))
    # equal (sub-bucket, k-mer, multiplicity) sets
    assert set(H) == set(S)
    for key, copies in S.items():
        assert all(mu == H[key][0] for mu, _ in copies)
    # duplicates: exactly the hash-map entries with both flag bits (halved),
    # two copies with one extra base each (in the stored orientation of its
    # record, which may differ between the two)
    dup = {key for key, copies in S.items() if len(copies) > 1}
    both = {key for key, (mu, f) in H.items() if f == 3}
    assert dup == both
    for key in dup:
        assert len(S[key]) == 2 and all(f in (1, 2) for _, f in S[key])
    return len(H), len(dup), st["stats"]


@pytest.mark.parametrize("b,sb", [(9, 6), (4, 1)])
def test_sorting_rule_config1_shape(b, sb):
    wl = synth.config1(0.01)
    n, d, st = _compare(wl.data, 31, 15, b, sb, 2)
    assert n > 50_000 and st["groups"] > 1000
    if b == 4:
        assert d > 0  # few sub-buckets: many boundaries inside one sub-bucket


def test_sorting_rule_config1_compacted():
    wl = synth.config1(0.005)
    _compare(wl.data, 31, 15, wl.buckets_log, wl.second_buckets_log, 2, compact=True)


def test_sorting_rule_config2_shape():
    wl = synth.config2(0.001)
    _compare(wl.data, 31, 15, 4, 6, 2)


def test_sorting_rule_config3_shape():
    gs = synth.config3_genomes(0.002, 8)
    data, _ = synth.contigs_joined(gs, 5, 3000, 9000)
    _compare(data, 31, 15, 5, 2, 2)


def test_sorting_rule_config4_shape():
    g = synth.genome(44, 150_000, repeat_frac=0.3, rep_min=500, rep_max=2000)
    data = synth.reads(g, 45, 30_000, 150, 0.005)
    _compare(data, 63, 21, 5, 3, 2, kind=1)
    _compare(data, 63, 21, 5, 3, 2, kind=0)


@pytest.mark.parametrize("k,m", [(31, 15), (21, 9), (11, 5), (63, 21), (27, 11)])
def test_sorting_rule_adversarial(k, m):
    rng = np.random.default_rng(3000 + k)
    data = joined(adversarial_records(rng))
    for b, sb in ((4, 2), (2, 0)):
        _compare(data, k, m, b, sb, 1)


# ---- independent restatement with explicit supertigs -----------------------
def _supertig_rule(records, k, m, s, b, sb):
    """Pure-Python restatement on the oracle's records, written from the
    reference's supertig description rather than from the C++ code:
    records of a sub-bucket are grouped by their minimizer m-mer (stored
    bases at minimizer_pos); in a group, every record contributes the
    k-mers [i, i+k) centred at offset i - minimizer_pos; a supertig is a
    maximal run of consecutive offsets whose k-mers are held by the same set
    of records; its multiplicity is the sum of their multiplicities, halved
    when it is one k-mer that one record ends without INCL_END and another
    starts without INCL_BEGIN."""
    sub, ln, fl, mp, mult = (records[x].tolist() for x in ("sub", "length", "flags", "minpos", "mult"))
    off, bases = records["off"].tolist(), records["bases"].tobytes()
    normal = (1 << b) << sb
    groups = defaultdict(list)
    for r in range(len(sub)):
        if sub[r] >= normal:
            continue
        seq = bases[off[r]:off[r] + ln[r]]
        groups[(sub[r], seq[mp[r]:mp[r] + m])].append(r)
    out = []  # (sub, kmer ascii, mult, extra flags)
    halved = 0
    for (sb_id, _), rs in groups.items():
        holders = defaultdict(set)  # (offset, kmer) -> records
        for r in rs:
            seq = bases[off[r]:off[r] + ln[r]]
            for i in range(ln[r] - k + 1):
                holders[(i - mp[r], seq[i:i + k])].add(r)
        by_offset = defaultdict(list)
        for (o, km), rset in holders.items():
            by_offset[o].append((km, frozenset(rset)))
        # supertigs: chain (o, km) -> (o + 1, km') when the same records hold both
        nxt = {}
        index = {(o, km): rset for (o, km), rset in holders.items()}
        for (o, km), rset in holders.items():
            for km2, rset2 in by_offset.get(o + 1, []):
                if rset2 == rset and km2[:-1] == km[1:]:
                    nxt[(o, km)] = (o + 1, km2)
        has_prev = set(nxt.values())
        for start in holders:
            if start in has_prev:
                continue
            tig = [start]
            while tig[-1] in nxt:
                tig.append(nxt[tig[-1]])
            rset = index[start]
            mu = sum(mult[r] for r in rset)
            fwd = any(fl[r] & 2 == 0 and tig[-1][0] == ln[r] - k - mp[r] for r in rset)
            bwd = any(fl[r] & 1 == 0 and tig[0][0] == -mp[r] for r in rset)
            flags = [0] * len(tig)
            flags[0] |= 1 if bwd else 0
            flags[-1] |= 2 if fwd else 0
            if len(tig) == 1 and fwd and bwd and len(rs) > 1:
                mu //= 2
                flags = [0]
                halved += 1
            if mu < s:
                continue
            for (o, km), f in zip(tig, flags):
                out.append((sb_id, km, mu, f))
    return out, halved


@pytest.mark.parametrize("seed,k,m", [(1, 31, 15), (2, 21, 9), (3, 11, 5)])
def test_sorting_rule_matches_supertig_restatement(seed, k, m):
    g = synth.genome(70 + seed, 3000, repeat_frac=0.3, rep_min=100, rep_max=400)
    data = np.concatenate([synth.reads(g, 80 + seed, 600, 120, 0.01),
                           joined(adversarial_records(np.random.default_rng(seed))[:25])])
    b, sb, s = 3, 1, 1 + seed % 2
    o = O.bucketing(data, k, m, b, sb, threads=2)
    o.compact()
    rec = o.records()
    want, halved = _supertig_rule(rec, k, m, s, b, sb)
    assert halved == 0
    kms = b"".join(km for _, km, _, _ in want)
    lo, hi, fw = O.kmer_keys(kms, k)
    step = np.arange(len(want)) * k  # one k-mer per k bytes of the joined string
    lo, hi, fw = lo[step], hi[step], fw[step]
    # the C++ restatement reports extra bases per centred k-mer; compare the
    # multiset of (sub, key, mult) and, per k-mer, the union of extra bases
    got = o.merge_sorting(k, m, b, sb, 0, s, 2)
    gm = sorted(zip(got["sub"].tolist(), got["key_lo"].tolist(), got["key_hi"].tolist(), got["mult"].tolist()))
    wm = sorted(zip([w[0] for w in want], lo.tolist(), hi.tolist(), [w[2] for w in want]))
    assert gm == wm and len(gm) > 1000
