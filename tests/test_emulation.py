"""Kernel logic on the CPU: the HIP sources built against a threaded stub
runtime (tests/emu) run the whole bucketing + merge path, compared with the
oracle.  This covers the cases that are kept off the GPU (long records in the
record pool, items split into hash-partitioned passes) at small sizes; it is
evidence about the kernels' control flow, not a substitute for the GPU run.
"""
from __future__ import annotations

import numpy as np
import pytest

from ggcat_amd import synth
from ggcat_amd.native import HASH_RK128, HASH_SEQ, GpuContext, Params
from oracle import binding as O
from tests.emu import build as emu_build
from tests.helpers import (adversarial_records, bitmaps_from_sets, colored_solid_array, fasta_text, joined,
                           build_both_readers, parse_case_files, random_bases, record_tuples, solid_array,
                           wrap_lines)


@pytest.fixture(scope="module")
def emu_ctx():
    ctx = GpuContext(0, lib_path=emu_build.build())
    yield ctx
    ctx.close()


def _check(ctx, data, k, m, b, sb, s, kind=HASH_SEQ):
    ctx.bucketing(data, Params(k, m, b, sb, 0, kind, s))
    o = O.bucketing(data, k, m, b, sb, threads=2)
    assert record_tuples(ctx.records()) == record_tuples(o.records())
    ctx.kmers_merge()
    o.merge(k, b, sb, kind, s, 2)
    assert np.array_equal(solid_array(ctx.solid()), solid_array(o.solid()))


@pytest.mark.parametrize("k,m,kind,tune", [(31, 15, HASH_SEQ, (1 << 20, 1, 2, 1 << 20)),
                                           (63, 21, HASH_RK128, (0, 1, 2, 0))])
def test_partitioned_passes_emulated(emu_ctx, k, m, kind, tune):
    g = synth.genome(5, 20_000)
    data = synth.reads(g, 6, 400, 150, 0.01)
    emu_ctx.set_tuning(*tune)
    try:
        # few sub-buckets: thousands of distinct k-mers per item
        _check(emu_ctx, data, k, m, 1, 1, 1, kind)
        c = emu_ctx.counters()
        assert c[4] > 0 and c[7] >= 2
        if tune[3]:
            assert c[7] >= 3  # the first partitions overflowed and were refined
    finally:
        emu_ctx.set_tuning()


def test_long_records_emulated(emu_ctx):
    rng = np.random.default_rng(1031)
    data = joined(adversarial_records(rng)[:40])
    emu_ctx.set_tuning(merge_wgs_per_cu=2)
    try:
        _check(emu_ctx, data, 31, 15, 4, 2, 1)
    finally:
        emu_ctx.set_tuning()


def test_colored_merge_emulated(emu_ctx):
    gs = synth.config3_genomes(0.002, 12)
    data, starts = synth.contigs_joined(gs, 5, 3000, 9000)
    colors = (np.arange(12, dtype=np.uint32) * 29) % 60  # ids need not follow file order; two bitmap words
    emu_ctx.set_tuning(merge_wgs_per_cu=2)
    try:
        emu_ctx.bucketing(data, Params(31, 15, 3, 2, 0, HASH_SEQ, 1), colors=(starts, colors))
        o = O.bucketing_colored(data, 31, 15, 3, 2, starts, colors, threads=2)
        assert record_tuples(emu_ctx.records()) == record_tuples(o.records())
        emu_ctx.kmers_merge()
        o.merge(31, 3, 2, HASH_SEQ, 1, 2)
        got = colored_solid_array(emu_ctx.solid(), emu_ctx.solid_colors())
        off, cols = o.solid_colors()
        want = colored_solid_array(o.solid(), bitmaps_from_sets(off, cols, 2))
        assert np.array_equal(got, want)
    finally:
        emu_ctx.set_tuning()


@pytest.mark.parametrize("unit_cap", [1, 7])
def test_unit_region_spill_emulated(emu_ctx, unit_cap):
    # nearly every unit spills records past its region
    rng = np.random.default_rng(57)
    data = np.concatenate([joined(adversarial_records(rng)[:12]), synth.reads(synth.genome(3, 4000), 8, 60, 120, 0.01)])
    emu_ctx.set_tuning(merge_wgs_per_cu=2, unit_cap=unit_cap)
    try:
        _check(emu_ctx, data, 31, 15, 4, 2, 1)
        seps = np.flatnonzero(data == 10) + 1  # color ranges start at record boundaries
        starts = np.array([0, seps[seps.size // 3], seps[seps.size // 2]], np.uint64)
        colors = np.array([2, 0, 1], np.uint32)
        emu_ctx.bucketing(data, Params(31, 15, 4, 2, 0, HASH_SEQ, 1), colors=(starts, colors))
        o = O.bucketing_colored(data, 31, 15, 4, 2, starts, colors, threads=2)
        assert record_tuples(emu_ctx.records()) == record_tuples(o.records())
    finally:
        emu_ctx.set_tuning()


# ---------------------------------------------------------------------------
# Input files read into device memory and parsed there (parse.hip) against the
# host reader (fasta.cpp, the restatement of sequences_reader.rs): the same
# records and solid k-mers for the corner cases of the reference's line
# reader, at read blocks small enough to put many block ends in small files.

@pytest.mark.parametrize("block,k,m,kind", [(512, 31, 15, HASH_SEQ), (1000, 63, 21, HASH_RK128)])
@pytest.mark.parametrize("colors", [0, 1])
def test_device_parse_matches_host_reader_emulated(emu_ctx, tmp_path, block, k, m, kind, colors):
    from ggcat_amd.native import BuildOptions
    paths = parse_case_files(tmp_path, block)
    opts = BuildOptions(k=k, m=m, min_multiplicity=1, hash_kind=kind, colors=colors)
    try:
        dev, host = build_both_readers(emu_ctx, paths, opts, block, merge_wgs_per_cu=2)
        (path_d, cuts, rec_d, sol_d), (path_h, _, rec_h, sol_h) = dev, host
    finally:
        emu_ctx.set_tuning()
    assert (path_d, path_h) == (1, 2)
    assert cuts > 0  # records longer than the flush size were cut on the device
    assert rec_d == rec_h
    assert np.array_equal(sol_d, sol_h)


@pytest.mark.parametrize("case", ["comment", "cr_line_at_block_end"])
def test_device_parse_falls_back_to_host_reader_emulated(emu_ctx, tmp_path, case):
    from ggcat_amd.native import BuildOptions
    block = 512
    rng = np.random.default_rng(9)
    data = bytearray(fasta_text(rng, 6, b"\n", block, long_header=False))
    if case == "comment":
        data[data.index(b">rec3"):data.index(b">rec3") + 1] = b";"
    else:
        # a line that is a lone '\r' at a block end, then a header: the
        # reference sees no piece for it and takes the header as a new line
        data = bytearray(b">a\n" + random_bases(rng, block - 5) + b"\n\r>late\n" + wrap_lines(random_bases(rng, 500), 60, b"\n"))
        assert data[block - 1] == ord("\r") and data[block - 2] == ord("\n")
    p = tmp_path / "x.fa"
    p.write_bytes(bytes(data))
    opts = BuildOptions(k=31, m=15, min_multiplicity=1)
    try:
        dev, host = build_both_readers(emu_ctx, [str(p)], opts, block, merge_wgs_per_cu=2)
        (path_d, _, rec_d, sol_d), (_, _, rec_h, sol_h) = dev, host
    finally:
        emu_ctx.set_tuning()
    assert path_d == 2
    assert rec_d == rec_h and np.array_equal(sol_d, sol_h)


def test_compaction_counts_identical_records_emulated(emu_ctx):
    # identical records of a unit are stored once with their multiplicity
    # (the reference's compactor); the decoded records, the solid table and
    # the colored solid table are unchanged, and copies of different colors
    # stay apart
    g = synth.genome(3, 8000)
    data = synth.reads(g, 4, 400, 120, 0.005)
    emu_ctx.set_tuning(merge_wgs_per_cu=2)
    try:
        _check(emu_ctx, data, 31, 15, 3, 2, 2)
        c = emu_ctx.counters()
        assert 0 < int(c[10]) < int(c[1]) // 1.5  # copies were folded
        # the same reads under two colors: nothing folds across them
        both = np.concatenate([data, data])
        starts = np.array([0, data.size], np.uint64)
        colors = np.array([1, 0], np.uint32)
        emu_ctx.bucketing(both, Params(31, 15, 3, 2, 0, HASH_SEQ, 2), colors=(starts, colors))
        c2 = emu_ctx.counters()
        assert int(c2[10]) == 2 * int(c[10])
        o = O.bucketing_colored(both, 31, 15, 3, 2, starts, colors, threads=2)
        assert record_tuples(emu_ctx.records()) == record_tuples(o.records())
        emu_ctx.kmers_merge()
        o.merge(31, 3, 2, HASH_SEQ, 2, 2)
        got = colored_solid_array(emu_ctx.solid(), emu_ctx.solid_colors())
        off, cols = o.solid_colors()
        assert np.array_equal(got, colored_solid_array(o.solid(), bitmaps_from_sets(off, cols, 1)))
    finally:
        emu_ctx.set_tuning()


@pytest.mark.parametrize("k,m,kind,big", [(31, 15, HASH_SEQ, False), (31, 15, HASH_SEQ, True),
                                          (63, 21, HASH_RK128, False)])
def test_direct_record_placement_emulated(emu_ctx, k, m, kind, big):
    # work items whose units the compaction deduplicated place their records
    # in the record table by position; items with a unit above the
    # deduplication's limit (big: two sub-buckets of one unit each, over a
    # thousand records) still hash them; the solid table is the oracle's
    # either way
    g = synth.genome(3, 8000)
    data = synth.reads(g, 4, 400, 120, 0.005)
    emu_ctx.set_tuning(merge_wgs_per_cu=2, unit_bytes=(1 << 20) if big else 0)
    try:
        b, sb = (1, 0) if big else (3, 2)
        _check(emu_ctx, data, k, m, b, sb, 2, kind)
        if big:
            assert int(emu_ctx.counters()[10]) > 2 * 1024  # a unit stores more than the limit
    finally:
        emu_ctx.set_tuning()


@pytest.mark.parametrize("k,m,kind", [(31, 15, HASH_SEQ), (63, 21, HASH_RK128)])
def test_resplit_emulated(emu_ctx, k, m, kind):
    # the three largest sub-buckets and the largest duplicates-bucket
    # sub-bucket of a repeat-rich input (with tandem repeats: long pool
    # records) resplit on the device as the oracle resplits them; then the
    # records and the solid table, every resplit bucket a sub-bucket of its own
    rng = np.random.default_rng(83)
    g = synth.genome(81, 30_000, repeat_frac=0.3, rep_min=500, rep_max=2000)
    data = np.concatenate([synth.reads(g, 82, 250, 150, 0.005), joined(adversarial_records(rng)[:30])])
    b, sb = 3, 1
    n_sub = ((1 << b) + 1) << sb
    emu_ctx.set_tuning(merge_wgs_per_cu=2)
    try:
        emu_ctx.bucketing(data, Params(k, m, b, sb, 0, kind, 2))
        o = O.bucketing(data, k, m, b, sb, threads=2)
        r = o.records()
        cnt = np.bincount(r["sub"].astype(np.int64), minlength=n_sub)
        normal = [int(s) for s in np.argsort(-cnt[:(1 << b) << sb], kind="stable")[:3]]
        dup = int(((1 << b) << sb) + np.argmax(cnt[(1 << b) << sb:]))
        assert cnt[dup] > 0
        lst = [(normal[0], 4), (dup, 8), (normal[1], 1), (normal[2], 16)]
        first = emu_ctx.resplit(lst)
        assert first == [n_sub, n_sub + 5, n_sub + 14, n_sub + 16]
        o.resplit(k, m, lst, n_sub)
        got_r, want_r = record_tuples(emu_ctx.records()), record_tuples(o.records())
        assert len(got_r) == len(want_r) and got_r == want_r
        assert any(t[0] >= n_sub for t in got_r)
        emu_ctx.kmers_merge()
        o.merge(k, b, sb, kind, 2, 2)
        assert np.array_equal(solid_array(emu_ctx.solid()), solid_array(o.solid()))
        with pytest.raises(Exception):
            emu_ctx.resplit(lst)  # once per bucketing
    finally:
        emu_ctx.set_tuning()


def test_resplit_colored_emulated(emu_ctx):
    # pieces keep their record's color: colored solid table after a resplit
    gs = synth.config3_genomes(0.002, 12)
    data, starts = synth.contigs_joined(gs, 5, 3000, 9000)
    colors = (np.arange(12, dtype=np.uint32) * 29) % 60
    b, sb = 3, 2
    n_sub = ((1 << b) + 1) << sb
    emu_ctx.set_tuning(merge_wgs_per_cu=2)
    try:
        emu_ctx.bucketing(data, Params(31, 15, b, sb, 0, HASH_SEQ, 1), colors=(starts, colors))
        o = O.bucketing_colored(data, 31, 15, b, sb, starts, colors, threads=2)
        cnt = np.bincount(o.records()["sub"].astype(np.int64), minlength=n_sub)
        lst = [(int(s), 8) for s in np.argsort(-cnt, kind="stable")[:2]]
        assert emu_ctx.resplit(lst) == [n_sub, n_sub + 9]
        o.resplit(31, 15, lst, n_sub)
        emu_ctx.kmers_merge()
        o.merge(31, b, sb, HASH_SEQ, 1, 2)
        got = colored_solid_array(emu_ctx.solid(), emu_ctx.solid_colors())
        off, cols = o.solid_colors()
        want = colored_solid_array(o.solid(), bitmaps_from_sets(off, cols, 2))
        assert np.array_equal(got, want)
        assert (emu_ctx.solid().sub >= n_sub).sum() > 100
    finally:
        emu_ctx.set_tuning()


@pytest.mark.parametrize("k,m,kind,chunk", [(31, 15, HASH_SEQ, 8192), (63, 21, HASH_RK128, 10000)])
def test_chunked_host_upload_emulated(emu_ctx, k, m, kind, chunk):
    # host input uploaded in chunks while the staging pass runs over the
    # tiles already resident: the same records and solid table as the oracle
    g = synth.genome(87, 20_000, repeat_frac=0.2, rep_min=300, rep_max=1000)
    data = synth.reads(g, 88, 600, 150, 0.005)
    assert data.size > 8 * chunk
    emu_ctx.set_tuning(merge_wgs_per_cu=2, upload_chunk=chunk)
    try:
        _check(emu_ctx, data, k, m, 3, 1, 2, kind)
    finally:
        emu_ctx.set_tuning()
