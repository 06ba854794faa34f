"""CPU tests that pin the oracle's outlier resplit (oracle/ggcat_oracle.cpp
resplit_records, after crates/kmers_transform/src/resplitter.rs:47-183 and
crates/assembler_minimizer_bucketing/src/lib.rs:152-270) without reusing its
code: an independent brute-force restatement of the splitting with the
record's include flags, and the solid k-mers of an outlier before and after
its resplit."""
import numpy as np
import pytest

from oracle import binding as O
from tests.helpers import adversarial_records, joined, record_tuples
from tests.test_oracle import brute_window, nt_direct, revcomp

M64 = (1 << 64) - 1


def resplit_m(k: int, m: int) -> int:
    # crates/assembler_kmers_merge/src/lib.rs:236-246 (RESPLITTING_MAX_K_M_DIFFERENCE = 10)
    return k - 10 if k > 11 else min(m, 2)


def bf_split(seq: bytes, k: int, m: int, b: int, include_first: bool, include_last: bool):
    """process_sequence by brute force: windows of w = k - m m-mers over the
    m-mers that are not skipped (the first one when the first k-mer is not
    included, the last one when the last k-mer is not), groups of equal
    window minimum (value, and position when it is unique); a group's
    record spans the k-mers that touch its windows, from base 0 for the
    first group and up to the last base for the last group.  Nothing is
    emitted when fewer than w m-mers remain."""
    L = len(seq)
    n = L - m + 1
    w = k - m
    sb, se = (0 if include_first else 1), (0 if include_last else 1)
    items = []
    for p in range(n):
        fw, rc = nt_direct(seq[p:p + m], m)
        items.append((((min(fw, rc) << 1) & M64) | (fw != rc), p, fw < rc))
    use = items[sb:n - se]
    if len(use) < w:
        return []
    win = brute_window([x[0] for x in use], [x[1] for x in use], [x[2] for x in use], w)
    nwin = len(win)
    starts = [0] + [j for j in range(1, nwin) if win[j][0] != win[j - 1][0] or
                    (win[j][0] & 1 and win[j][1:] != win[j - 1][1:])]
    out = []
    for gi, a in enumerate(starts):
        first, last = gi == 0, gi + 1 == len(starts)
        bb = nwin - 1 if last else starts[gi + 1] - 1
        val, pos, f = win[bb]
        start = 0 if first else a + sb - 1
        end = L if last else bb + sb + k
        uniq = val & 1
        rc = bool(uniq) and not f
        bucket = ((val >> 1) & ((1 << b) - 1)) if uniq else (1 << b)
        minpos = 0 if not uniq else (end - pos - m if rc else pos - start)
        inc_first = include_first and first
        inc_last = include_last and last
        flags = (int(inc_first) << int(rc)) | (int(inc_last) << int(not rc))
        s = seq[start:end]
        out.append((bucket, end - start, flags, minpos, revcomp(s) if rc else s))
    return out


def _records(r, i):
    o, L = int(r["off"][i]), int(r["length"][i])
    return bytes(r["bases"][o:o + L])


@pytest.mark.parametrize("k,m", [(31, 15), (21, 9), (15, 7)])
def test_resplit_matches_bruteforce(k, m):
    # every record of a few sub-buckets, with every include-flag combination
    # the first bucketing produces, resplit into 2^3 buckets + duplicates
    rng = np.random.default_rng(k * 7 + m)
    recs = [r for r in adversarial_records(rng) if len(r) <= 500]
    data = joined(recs)
    h = O.bucketing(data, k, m, 3, 1)
    r = h.records()
    subs = sorted(set(int(s) for s in r["sub"]))[:6]
    m2 = resplit_m(k, m)
    first_extra = ((1 << 3) + 1) << 1
    outl = [(s, 8) for s in subs]
    want = []
    for i in range(r["sub"].size):
        s = int(r["sub"][i])
        if s not in subs:
            continue
        base = first_extra + subs.index(s) * 9
        fl = int(r["flags"][i])
        for bucket, L, flags, minpos, seq in bf_split(_records(r, i), k, m2, 3, bool(fl & 1), bool(fl & 2)):
            want.append((base + bucket, L, flags, minpos, seq))
    seen_flags = {int(r["flags"][i]) for i in range(r["sub"].size) if int(r["sub"][i]) in subs}
    assert seen_flags >= {0, 1, 2}  # inner, first and last records all resplit
    h.resplit(k, m, outl, first_extra)
    got = h.records()
    got_t = [t for t in record_tuples(got) if t[0] >= first_extra]
    assert got_t == sorted(want)
    # the other sub-buckets keep their records
    assert [t for t in record_tuples(got) if t[0] < first_extra] == \
        [t for t in record_tuples(r) if t[0] not in subs]


def test_resplit_single_kmer_record_dropped_when_not_included():
    # a record of one k-mer included at neither end keeps w - 1 m-mers once
    # both ends are skipped, which is no complete window: nothing is emitted
    # (batch_minqueue.rs:46-50); with one more base, or one end included, the
    # record comes back whole
    k, m = 31, 15
    m2 = resplit_m(k, m)
    rng = np.random.default_rng(1)
    s = bytes(rng.choice(np.frombuffer(b"ACGT", np.uint8), k + 1))
    assert bf_split(s[:k], k, m2, 2, False, False) == []
    for seq, a, b_ in ((s, False, False), (s[:k], False, True), (s[:k], True, False), (s[:k], True, True)):
        got = bf_split(seq, k, m2, 2, a, b_)
        assert len(got) == 1 and got[0][1] == len(seq)


@pytest.mark.parametrize("k,m,kind", [(31, 15, 0), (45, 21, 0), (63, 21, 1)])
def test_resplit_keeps_solid_kmers(k, m, kind):
    # an outlier's solid k-mers and their multiplicities are the same before
    # and after its resplit: every occurrence of a k-mer lands in the resplit
    # bucket of its own window minimum, and a k-mer that becomes a boundary
    # between two resplit records is counted on each side
    from ggcat_amd import synth
    g = synth.genome(71, 40_000, repeat_frac=0.3, rep_min=500, rep_max=2000)
    data = synth.reads(g, 72, 6_000, 150, 0.005)
    b, sb = 3, 1
    h = O.bucketing(data, k, m, b, sb)
    h.merge(k, b, sb, kind, 2)
    before = h.solid()
    r = h.records()
    subs = [int(s) for s in np.argsort(-np.bincount(r["sub"]))[:3]]
    first_extra = ((1 << b) + 1) << sb
    h2 = O.bucketing(data, k, m, b, sb)
    h2.resplit(k, m, [(s, 16) for s in subs], first_extra)
    h2.merge(k, b, sb, kind, 2)
    after = h2.solid()

    def table(sol, keep):
        out = {}
        for i in range(sol["sub"].size):
            s = int(sol["sub"][i])
            if keep(s):
                key = (int(sol["key_lo"][i]), int(sol["key_hi"][i]))
                out.setdefault(key, set()).add(int(sol["mult"][i]))
        return out

    bef = table(before, lambda s: s in subs)
    aft = table(after, lambda s: s >= first_extra)
    assert len(bef) > 1000
    assert set(bef) == set(aft)
    assert all(len(v) == 1 for v in aft.values())
    assert all(bef[key] == aft[key] for key in bef)
    # untouched sub-buckets are unchanged
    assert table(before, lambda s: s not in subs) == table(after, lambda s: s < first_extra)
