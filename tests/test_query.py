"""Query semantics: the oracle's per-query counts against a brute force over
the graph's k-mer set, the reference's CSV / JSONL formats on known answers,
and the kernel logic on the CPU emulation."""
from __future__ import annotations

import numpy as np
import pytest

from ggcat_amd import query as Q
from ggcat_amd import synth
from oracle import binding as O

COMP = bytes.maketrans(b"ACGT", b"TGCA")


def revcomp(kmer: bytes) -> bytes:
    return kmer.translate(COMP)[::-1]


def brute_query(graph: dict, queries: bytes, k: int, n_colors: int):
    """graph: k-mer (ASCII, both orientations) -> color tuple."""
    qk, mt, cc = [], [], []
    for rec in queries.split(b"\n")[:-1]:
        up = rec.upper()
        qk.append(max(0, len(rec) - (k - 1)))
        m, c = 0, [0] * n_colors
        for i in range(len(up) - k + 1):
            km = up[i:i + k]
            if any(ch not in b"ACGT" for ch in km):
                continue
            cs = graph.get(km)
            if cs is None:
                continue
            m += 1
            for col in cs:
                c[col] += 1
        mt.append(m)
        cc.append(c)
    return np.array(qk, np.uint64), np.array(mt, np.uint64), np.array(cc, np.uint64)


def graph_from_bruteforce(data, k, s, starts, colors):
    sol, (off, cols) = O.bruteforce_counts(data, k, s, starts, colors)
    graph = {}
    for i, lo in enumerate(sol["key_lo"].tolist()):
        # seq-hash key: base j at bits 2j, A0 C1 T2 G3
        km = bytes(b"ACTG"[(lo >> (2 * j)) & 3] for j in range(k))
        graph[km] = graph[revcomp(km)] = tuple(cols[off[i]:off[i + 1]].tolist())
    return graph


@pytest.mark.parametrize("s", [1, 2])
def test_oracle_query_matches_bruteforce(s):
    k = 21
    anc = synth.genome(31, 6000)
    gs = [synth.descendant(anc, 40 + i, 0.02, 0.002) for i in range(5)]
    data, starts = synth.contigs_joined(gs, 3, 1500, 2500)
    colors = np.array([4, 0, 3, 1, 2], np.uint32)
    r = O.bucketing_colored(data, k, 9, 4, 1, starts, colors, threads=2)
    r.merge(k, 4, 1, 0, s, 2)
    q = synth.queries(gs, 7, 25, 10, 300, 0.03)
    q = np.concatenate([q, np.frombuffer(b"acgtacgtacgtacgtacgtacgtNNacgt\n\nA\n", np.uint8)])
    got = r.query(q, k, 0, 5)
    want = brute_query(graph_from_bruteforce(data, k, s, starts, colors), q.tobytes(), k, 5)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    assert np.array_equal(got[2], want[2])
    assert int(got[1].sum()) > 0


def test_lookup_query_matches_oracle():
    # the vectorized lookup counter the GPU tests use at full scale, pinned
    # against the oracle's query on a colored graph
    from tests.helpers import bitmaps_from_sets, graph_index, query_by_lookup
    k = 31
    anc = synth.genome(32, 8000)
    gs = [synth.descendant(anc, 50 + i, 0.02, 0.002) for i in range(6)]
    data, starts = synth.contigs_joined(gs, 3, 1500, 2500)
    colors = np.array([5, 0, 3, 1, 2, 4], np.uint32)
    r = O.bucketing_colored(data, k, 15, 4, 1, starts, colors, threads=2)
    r.merge(k, 4, 1, 0, 1, 2)
    q = synth.queries(gs, 9, 40, 10, 600, 0.03)
    q = np.concatenate([q, np.frombuffer(b"acgtacgtacgtacgtacgtacgtacgtacgNNacgt\n\nA\n", np.uint8)])
    s = r.solid()
    off, cols = r.solid_colors()
    keys, bm = graph_index(s["key_lo"], bitmaps_from_sets(off, cols, 1))
    got = query_by_lookup(q, keys, bm, k)
    want = r.query(q, k, 0, 6)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    assert np.array_equal(got[2][:, :6], want[2])
    assert int(got[1].sum()) > 0


def test_csv_format_known_answer():
    qk = np.array([120, 0, 3], np.uint64)
    mt = np.array([64, 0, 1], np.uint64)
    assert Q.format_csv(qk, mt) == ("query_index,matched_kmers,query_kmers,match_percentage\n"
                                   "0,64,120,0.53\n1,0,0,NaN\n2,1,3,0.33\n")


def test_jsonl_format_known_answer():
    qk = np.array([100, 50, 8], np.uint64)
    cc = np.array([[0, 25, 100], [0, 0, 0], [1, 0, 7]], np.uint32)
    assert Q.format_jsonl(qk, cc) == ('{"query_index":0, "matches":{"1": 0.25,"2": 1.00}}\n'
                                      '{"query_index":2, "matches":{"0": 0.12,"2": 0.88}}\n')
    assert Q.format_jsonl(qk, cc, ["a", "b", "c"]).startswith('{"query_index":0, "matches":{"b": 0.25,"c": 1.00}}')


def test_query_and_serving_emulated():
    from ggcat_amd.native import GpuContext, Params
    from tests.emu import build as emu_build
    ctx = GpuContext(0, lib_path=emu_build.build())
    try:
        ctx.set_tuning(merge_wgs_per_cu=2)
        gs = synth.config3_genomes(0.001, 6)
        data, starts = synth.contigs_joined(gs, 5, 2000, 5000)
        colors = np.arange(6, dtype=np.uint32)
        ctx.bucketing(data, Params(31, 15, 3, 2, 0, 0, 1), colors=(starts, colors))
        ctx.kmers_merge()
        ctx.index_build()
        q = synth.queries(gs, 9, 12, 40, 400, 0.02)
        qk, mt, cc = ctx.query(q)
        o = O.bucketing_colored(data, 31, 15, 3, 2, starts, colors, threads=2)
        o.merge(31, 3, 2, 0, 1, 2)
        oqk, omt, occ = o.query(q, 31, 0, 6)
        assert np.array_equal(qk, oqk) and np.array_equal(mt, omt) and np.array_equal(cc.astype(np.uint64), occ)
        ctx.serve_open(2, 1 << 20)
        reqs = [synth.queries(gs, 50 + i, 6, 40, 300, 0.02) for i in range(4)]
        results = {}
        rep = Q.serve_poisson(ctx, reqs, rate_hz=1e6, results=results)
        ctx.serve_close()
        assert len(rep.latencies_ms) == 4
        for i, rq in enumerate(reqs):
            alone = ctx.query(rq)
            assert all(np.array_equal(a, b) for a, b in zip(results[i], alone))
        # a request larger than the whole budget is refused at submission and
        # does not block the requests after it
        ctx.serve_open(1, 8192)
        from ggcat_amd.native import GGCATGpuError
        with pytest.raises(GGCATGpuError, match="budget"):
            ctx.serve_submit(synth.queries(gs, 70, 50, 400, 600, 0.02))
        small = synth.queries(gs, 71, 2, 40, 60, 0.02)
        t = ctx.serve_submit(small)
        assert ctx.serve_poll(block=True) == t
        got = ctx.serve_fetch(t, 2)
        assert all(np.array_equal(a, b) for a, b in zip(got, ctx.query(small)))
        assert ctx.serve_stats() == (0, 0)
        ctx.serve_close()
    finally:
        ctx.close()
