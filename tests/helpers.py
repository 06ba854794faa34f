"""Shared helpers for parity tests: canonical forms of record multisets and
solid tables, and adversarial inputs."""
from __future__ import annotations

import numpy as np

SEP = ord("\n")


def record_tuples(r) -> list:
    """Sorted list of (sub, len, flags, minpos, bases) for small record sets."""
    get = (lambda x, n: x[n]) if isinstance(r, dict) else (lambda x, n: getattr(x, n))
    sub, ln, fl, mp, off, bases = (get(r, n) for n in ("sub", "length", "flags", "minpos", "off", "bases"))
    out = []
    for i in range(sub.size):
        o = int(off[i])
        out.append((int(sub[i]), int(ln[i]), int(fl[i]), int(mp[i]), bytes(bases[o:o + int(ln[i])])))
    out.sort()
    return out


def record_digests(r, with_color: bool = False) -> np.ndarray:
    """Sorted 2x64-bit digests of every record (sub, len, flags, minpos, bases,
    and the record color when with_color); used to compare multisets of
    millions of records."""
    get = (lambda x, n: x[n]) if isinstance(r, dict) else (lambda x, n: getattr(x, n))
    sub, ln, fl, mp, off, bases = (get(r, n) for n in ("sub", "length", "flags", "minpos", "off", "bases"))
    n = sub.size
    if n == 0:
        return np.zeros((0, 2), np.uint64)
    ln64 = ln.astype(np.uint64)
    total = int(ln64.sum())
    # records are contiguous in `bases` in record order with off == running sum
    assert np.array_equal(off.astype(np.uint64), np.concatenate(([0], np.cumsum(ln64)[:-1])).astype(np.uint64))
    pos = np.arange(total, dtype=np.uint64) - np.repeat(off.astype(np.uint64), ln64.astype(np.int64))
    b = bases[:total].astype(np.uint64)
    out = np.empty((n, 2), np.uint64)
    with np.errstate(over="ignore"):
        for j, (p1, p2) in enumerate(((0x9E3779B97F4A7C15, 0xC2B2AE3D27D4EB4F), (0x165667B19E3779F9, 0xD6E8FEB86659FD93))):
            w = (pos + np.uint64(p1)) * np.uint64(p2)
            w ^= w >> np.uint64(29)
            w *= np.uint64(0xBF58476D1CE4E5B9)
            h = np.add.reduceat(b * w, np.concatenate(([0], np.cumsum(ln64)[:-1])).astype(np.int64)) if total else np.zeros(n, np.uint64)
            h = h.astype(np.uint64)
            h ^= (sub.astype(np.uint64) * np.uint64(p1)) ^ (ln64 << np.uint64(20)) ^ (fl.astype(np.uint64) << np.uint64(40)) ^ (mp.astype(np.uint64) << np.uint64(48))
            if with_color:
                h ^= (get(r, "color").astype(np.uint64) + np.uint64(1)) * np.uint64(0x9FB21C651E98DF25)
            h *= np.uint64(p2)
            h ^= h >> np.uint64(31)
            out[:, j] = h
    idx = np.lexsort((out[:, 1], out[:, 0]))
    return out[idx]


def solid_tuples(s) -> list:
    if isinstance(s, dict):
        lo, hi, mult, fl, sub = s["key_lo"], s["key_hi"], s["mult"], s["flags"], s["sub"]
    else:
        lo = s.keys[:, 0]
        hi = s.keys[:, 1] if s.keys.shape[1] > 1 else np.zeros_like(lo)
        mult, fl, sub = s.mult, s.flags, s.sub
    return sorted(zip(sub.tolist(), lo.tolist(), hi.tolist(), mult.tolist(), fl.tolist()))


def solid_array(s) -> np.ndarray:
    """(n, 5) uint64 array sorted lexicographically: sub, key_lo, key_hi, mult, flags."""
    if isinstance(s, dict):
        cols = [s["sub"], s["key_lo"], s["key_hi"], s["mult"], s["flags"]]
    else:
        lo = s.keys[:, 0]
        hi = s.keys[:, 1] if s.keys.shape[1] > 1 else np.zeros_like(lo)
        cols = [s.sub, lo, hi, s.mult, s.flags]
    a = np.stack([c.astype(np.uint64) for c in cols], axis=1)
    idx = np.lexsort(tuple(a[:, j] for j in range(4, -1, -1)))
    return a[idx]


def bitmaps_from_sets(off: np.ndarray, cols: np.ndarray, words: int) -> np.ndarray:
    """Color lists (offsets [n+1] into sorted ids) -> (n, words) uint32 bitmaps."""
    n = off.size - 1
    out = np.zeros((n, words), np.uint32)
    row = np.repeat(np.arange(n), np.diff(off.astype(np.int64)))
    c = cols.astype(np.int64)
    np.bitwise_or.at(out, (row, c >> 5), (np.uint32(1) << (c & 31).astype(np.uint32)))
    return out


def colored_solid_array(s, bitmaps: np.ndarray) -> np.ndarray:
    """solid_array rows extended with the color bitmap words, same sort."""
    if isinstance(s, dict):
        cols = [s["sub"], s["key_lo"], s["key_hi"], s["mult"], s["flags"]]
    else:
        lo = s.keys[:, 0]
        hi = s.keys[:, 1] if s.keys.shape[1] > 1 else np.zeros_like(lo)
        cols = [s.sub, lo, hi, s.mult, s.flags]
    a = np.stack([c.astype(np.uint64) for c in cols] + [bitmaps[:, j].astype(np.uint64)
                                                         for j in range(bitmaps.shape[1])], axis=1)
    idx = np.lexsort(tuple(a[:, j] for j in range(4, -1, -1)))
    return a[idx]


def joined(records: list[bytes]) -> np.ndarray:
    return np.frombuffer(b"".join(r + b"\n" for r in records), np.uint8).copy()


def adversarial_records(rng: np.random.Generator) -> list[bytes]:
    """Edge cases: N runs, lowercase, short/exact-k records, homopolymers,
    tandem repeats, a long random contig and a long poly-A run."""
    def rnd(n):
        return bytes(rng.choice(np.frombuffer(b"ACGT", np.uint8), n))

    recs = [
        b"",
        b"ACGT",
        rnd(30),  # k-1
        rnd(31),  # exactly k
        rnd(32),
        rnd(60) + b"N" + rnd(40),
        rnd(10) + b"NNNN" + rnd(31) + b"N" + rnd(29) + b"NN" + rnd(80),
        rnd(70).lower(),
        rnd(20) + rnd(40).lower() + rnd(20),
        b"A" * 200,
        b"C" * 31,
        b"AC" * 120,
        b"ACG" * 90,
        b"AAAAAAAAAAAAAAAACAAAAAAAAAAAAAAAAAAAAA" * 4,
        rnd(40) + b"A" * 90 + rnd(40),
        rnd(100) + b"X" + rnd(100) + b"-" + rnd(50),
        rnd(5000),
        b"A" * 5000,
        rnd(300) + (b"GATTACA" * 200) + rnd(300),
    ]
    for _ in range(40):
        recs.append(rnd(int(rng.integers(20, 400))))
    return recs


def inline_capacity(k: int, m: int) -> int:
    """Bases a device record slot holds inline: the slot has 1 + ceil((2k - m)
    / 16) words rounded up to a power of two (at least 4), one of them the meta
    word, 16 bases per word.  Longer records go to the long-record pool."""
    sw = 1 + (2 * k - m + 15) // 16
    w = 4
    while w < sw:
        w <<= 1
    return (w - 1) * 16


def edge_records_inline(rng: np.random.Generator, k: int) -> list[bytes]:
    """The reference's edge cases that stay within an inline slot: empty and
    short records (k-1, k, k+1), N runs, lowercase and IUPAC / other non-ACGT
    bytes (all of them split records, crates/minimizer_bucketing/src/sequences_splitter.rs:15-40),
    short homopolymers and tandem repeats (windows whose minimum occurs twice:
    duplicates bucket, crates/assembler_minimizer_bucketing/src/lib.rs:221-227)
    embedded in random sequence, and long random contigs spanning many tiles."""
    def rnd(n):
        return bytes(rng.choice(np.frombuffer(b"ACGT", np.uint8), n))

    recs = [b"", b"ACGT", rnd(k - 1), rnd(k), rnd(k + 1), rnd(k - 1) + b"N" + rnd(k), b"N" * 40, rnd(k).lower(),
            rnd(60) + b"N" + rnd(40), rnd(10) + b"NNNN" + rnd(k) + b"N" + rnd(k - 2) + b"NN" + rnd(80),
            rnd(70).lower(), rnd(20) + rnd(40).lower() + rnd(20)]
    for c in b"RYKMSWBDHVUX-.*n":
        recs.append(rnd(int(rng.integers(k, 3 * k))) + bytes([c]) + rnd(int(rng.integers(k, 3 * k))))
    for unit in (b"A", b"C", b"G", b"T", b"AC", b"GT", b"ACG", b"GATTACA"):
        for reps in (k // len(unit) + 1, k // len(unit) + 6):
            recs.append(rnd(int(rng.integers(0, 40))) + unit * reps + rnd(int(rng.integers(0, 40))))
    recs += [rnd(20_000), rnd(7_777), rnd(3000).lower() + rnd(3000)]
    for _ in range(120):
        recs.append(rnd(int(rng.integers(k - 2, 400))))
    return recs


# Input files for the device parse path: the corner cases of the reference's
# line reader (sequences_reader.rs / lines_reader.rs) at a given read block.

def random_bases(rng, n: int) -> bytes:
    s = rng.choice(np.frombuffer(b"ACGTacgtN", np.uint8), n, p=[.23, .23, .23, .23, .02, .02, .02, .01, .01])
    return s.tobytes()


def wrap_lines(seq: bytes, width: int, eol: bytes) -> bytes:
    return b"".join(seq[i:i + width] + eol for i in range(0, len(seq), width))


def fasta_text(rng, n_rec: int, eol: bytes, block: int, long_header: bool = True) -> bytes:
    out = bytearray()
    out += wrap_lines(random_bases(rng, 150), 60, eol)  # sequence before the first header
    out += eol  # an empty line
    for r in range(n_rec):
        hdr = b">rec%d" % r + (b" x" * (block // 2 + 50) if long_header and r == 2 else b"")
        out += hdr + eol
        width = int(rng.choice([50, 60, 61, 80]))
        out += wrap_lines(random_bases(rng, int(rng.integers(0, 3200))), width, eol)
        if r % 4 == 3:
            out += eol
    return bytes(out)


def lone_cr_at_block_ends(data: bytes, block: int) -> bytes:
    """A '\\r' in the middle of sequence lines at a few block ends."""
    b = bytearray(data)
    for e in range(block - 1, len(b) - 2, 3 * block):
        line0 = b.rfind(b"\n", 0, e) + 1
        if e > line0 and b[line0] not in b">;" and b[e] not in b"\r\n" and b[e + 1] not in b"\r\n":
            b[e] = ord("\r")
    return bytes(b)


def fastq_text(rng, n: int, eol: bytes) -> bytes:
    out = bytearray()
    for i in range(n):
        s = random_bases(rng, int(rng.integers(60, 300)))
        out += b"@r%d" % i + eol + s + eol + b"+" + eol + b"I" * len(s) + (eol if i + 1 < n else b"")
    return bytes(out)


def parse_case_files(tmp_path, block: int) -> list[str]:
    rng = np.random.default_rng(block)
    files = {
        "multi.fa": fasta_text(rng, 12, b"\n", block),
        "crlf.fasta": lone_cr_at_block_ends(fasta_text(rng, 10, b"\r\n", block), block),
        "oneline.fa": b">one\n" + random_bases(rng, 5000) + b"\r",  # one line, no '\n' at the end
        "reads.fq": lone_cr_at_block_ends(fastq_text(rng, 40, b"\r\n"), block),
        "empty.fa": b"",
        "tail.fna": fasta_text(rng, 3, b"\n", block, long_header=False)[:-1],
    }
    paths = []
    for name, data in files.items():
        p = tmp_path / name
        p.write_bytes(data)
        paths.append(str(p))
    return paths


def build_both_readers(ctx, paths, opts, block: int, **tuning) -> list:
    """build_from_files with the device parse path, then with the host
    reader: (input path, flush cuts, records, solid table) of each."""
    out = []
    for host in (0, 1):
        ctx.set_tuning(read_block=block, host_reader=host, **tuning)
        ctx.build_from_files(paths, opts)
        c = ctx.counters()
        s = (colored_solid_array(ctx.solid(), ctx.solid_colors()) if opts.colors else solid_array(ctx.solid()))
        out.append((int(c[8]), int(c[9]), record_tuples(ctx.records()), s))
    return out


def query_by_lookup(queries: np.ndarray, keys: np.ndarray, bitmaps: np.ndarray | None, k: int):
    """Query counts by direct lookup (seq-hash keys, k <= 32): for each query
    record of the joined layout, query_kmers = len - (k - 1), matched = its
    all-ACGT k-mers whose canonical 2-bit key is in `keys`, and per color the
    matched k-mers whose graph k-mer's color bitmap holds that color.  keys
    must be sorted, bitmaps (n, W) uint32 in the same order (or None)."""
    q = np.ascontiguousarray(queries, np.uint8)
    nl = np.flatnonzero(q == 10)
    nq = nl.size
    starts = np.concatenate(([0], nl[:-1] + 1))
    lens = nl - starts
    qk = np.maximum(lens.astype(np.int64) - (k - 1), 0).astype(np.uint64)
    up = q & 0xDF
    valid = (up == ord("A")) | (up == ord("C")) | (up == ord("G")) | (up == ord("T"))
    code = ((q >> 1) & 3).astype(np.uint64)
    n = q.size - k + 1
    if n <= 0:
        return qk, np.zeros(nq, np.uint64), None
    fw = np.zeros(n, np.uint64)
    rc = np.zeros(n, np.uint64)
    for i in range(k):
        fw |= code[i:i + n] << np.uint64(2 * i)
        rc |= (code[k - 1 - i:k - 1 - i + n] ^ np.uint64(2)) << np.uint64(2 * i)
    key = np.minimum(fw, rc)
    bad = np.concatenate(([0], np.cumsum(~valid, dtype=np.int64)))
    ok = (bad[k:k + n] - bad[:n]) == 0
    pos = np.searchsorted(keys, key)
    pos_c = np.minimum(pos, max(keys.size - 1, 0))
    hit = ok & (keys.size > 0) & (keys[pos_c] == key) if keys.size else np.zeros(n, bool)
    qid = np.concatenate(([0], np.cumsum(q[:n - 1] == 10, dtype=np.int64)))  # record holding position p
    mt = np.bincount(qid[hit], minlength=nq).astype(np.uint64)
    cc = None
    if bitmaps is not None:
        W = bitmaps.shape[1]
        rows = bitmaps[pos_c[hit]]
        hq = qid[hit]
        cc = np.zeros((nq, 32 * W), np.uint64)
        for w in range(W):
            col = rows[:, w]
            for b in range(32):
                m = ((col >> np.uint32(b)) & 1).astype(bool)
                if m.any():
                    cc[:, 32 * w + b] = np.bincount(hq[m], minlength=nq)
    return qk, mt, cc


def graph_index(keys: np.ndarray, bitmaps: np.ndarray | None):
    """Solid entries (a boundary k-mer can be held by two sub-buckets) ->
    sorted unique keys and, per key, the OR of its entries' color bitmaps."""
    u, first, inv, cnt = np.unique(keys, return_index=True, return_inverse=True, return_counts=True)
    if bitmaps is None:
        return u, None
    bm = bitmaps[first].copy()
    dup = cnt[inv] > 1
    if dup.any():
        np.bitwise_or.at(bm, inv[dup], bitmaps[dup])
    return u, bm
