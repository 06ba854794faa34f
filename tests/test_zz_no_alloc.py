"""The query entry point that allocates nothing (ggcat_gpu_query_device_noalloc)
against ggcat_gpu_query on the same index: queries and results in device
memory owned by the caller, two calls in a row (the second smaller, so stale
scratch contents of the first call would show), colored and uncolored."""
import sys

import numpy as np
import pytest

from ggcat_amd import synth
from ggcat_amd.native import HASH_SEQ, GGCATGpuError, Params

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _noalloc(ctx, q: np.ndarray, n_colors: int):
    nq = int(np.count_nonzero(q == 10))
    d_q, _ = ctx.upload(q)
    d_k = ctx.alloc(max(nq, 1) * 8)
    d_m = ctx.alloc(max(nq, 1) * 8)
    d_c = ctx.alloc(max(nq * n_colors, 1) * 4) if n_colors else None
    try:
        ctx.query_device_noalloc(d_q, q.size, nq, d_k, d_m, d_c)
        qk = ctx.download(d_k, np.zeros(nq, np.uint64))
        mt = ctx.download(d_m, np.zeros(nq, np.uint64))
        cc = ctx.download(d_c, np.zeros((nq, n_colors), np.uint32)) if n_colors else None
    finally:
        for p in (d_q, d_k, d_m, d_c):
            if p is not None:
                ctx.free(p)
    return qk, mt, cc


def _same(a, b):
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], b[1])
    if a[2] is None:
        assert b[2] is None
    else:
        assert np.array_equal(a[2], b[2])


def test_noalloc_query_colored_two_calls(gpu_ctx):
    gs = synth.config3_genomes(0.01, 48)
    data, starts = synth.contigs_joined(gs, 404)
    colors = np.arange(48, dtype=np.uint32)
    _say(f"case no-alloc colored query: genomes=48 bytes={data.size}")
    gpu_ctx.bucketing(data, Params(31, 15, 5, 3, 0, HASH_SEQ, 1), colors=(starts, colors))
    gpu_ctx.kmers_merge()
    gpu_ctx.index_build()
    nc = gpu_ctx.index_colors()
    assert nc == 48
    big = synth.queries(gs, 71, 900, 100, 5000, 0.02)
    small = synth.queries(gs, 72, 37, 100, 2000, 0.02)
    for q in (big, small):
        _say(f"  call: {int(np.count_nonzero(q == 10))} queries, {q.size} bytes")
        _same(_noalloc(gpu_ctx, q, nc), gpu_ctx.query(q))


def test_noalloc_query_uncolored_two_calls(gpu_ctx):
    g = synth.genome(81, 200_000)
    data = synth.reads(g, 82, 30_000, 150, 0.005)
    _say(f"case no-alloc uncolored query: bytes={data.size}")
    gpu_ctx.bucketing(data, Params(31, 15, 6, 3, 0, HASH_SEQ, 2))
    gpu_ctx.kmers_merge()
    gpu_ctx.index_build()
    big = synth.queries([g], 83, 600, 100, 5000, 0.02)
    small = synth.queries([g], 84, 5, 100, 400, 0.02)
    for q in (big, small):
        _say(f"  call: {int(np.count_nonzero(q == 10))} queries, {q.size} bytes")
        _same(_noalloc(gpu_ctx, q, 0), gpu_ctx.query(q))
    # inputs the entry point rejects before any query kernel runs
    q = small.copy()
    d_q, _ = gpu_ctx.upload(q)
    d_m = gpu_ctx.alloc(64)
    try:
        with pytest.raises(GGCATGpuError):  # one record fewer than announced
            gpu_ctx.query_device_noalloc(d_q, q.size, int(np.count_nonzero(q == 10)) + 1, None, d_m, None)
        with pytest.raises(GGCATGpuError):  # last record not terminated
            gpu_ctx.query_device_noalloc(d_q, q.size - 1, int(np.count_nonzero(q == 10)), None, d_m, None)
        with pytest.raises(GGCATGpuError):  # above the stated maximum
            gpu_ctx.query_device_noalloc(d_q, q.size, 500_001, None, d_m, None)
    finally:
        gpu_ctx.free(d_q)
        gpu_ctx.free(d_m)
    # a valid call after the rejected ones still matches
    _same(_noalloc(gpu_ctx, small, 0), gpu_ctx.query(small))
