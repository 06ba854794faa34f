"""Input reading (FASTA/FASTQ; plain, gzip, lz4, bz2, xz, zstd) and build-parameter planning.

Host code only: ggcat_gpu_read_files and ggcat_gpu_plan_build need no GPU.
Expected outputs are known answers worked out from the reference's reader
(crates/io/src/sequences_reader.rs:106-240, crates/io/src/lines_reader.rs:
147-219) and parameter derivation (crates/io/src/lib.rs:67-140,
crates/utils/src/lib.rs:29-40).
"""
from __future__ import annotations

import gzip
import os

import numpy as np
import pytest

from ggcat_amd import native

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "example-inputs")
BLOCK = 4 << 20  # the reference's LinesReader buffer (DEFAULT_OUTPUT_BUFFER_SIZE)


def records(paths, k=31, colors=False):
    r = native.read_files([str(p) for p in paths], k, colors)
    text = r.joined.tobytes()
    assert text == b"" or text.endswith(b"\n")
    return (text.split(b"\n")[:-1] if text else []), r


def write(tmp_path, name, data: bytes):
    p = tmp_path / name
    if name.endswith(".gz"):
        with gzip.open(p, "wb") as f:
            f.write(data)
    else:
        p.write_bytes(data)
    return p


def test_fasta_multiline_records(tmp_path):
    p = write(tmp_path, "a.fa", b">r1 desc\nACGT\nacgtN\n\n>r2\nTTTT\n>empty\n>r3\nGG")
    recs, r = records([p])
    # an empty record is not emitted; the last line needs no newline
    assert recs == [b"ACGTacgtN", b"TTTT", b"GG"]
    assert r.estimated_bases == os.path.getsize(p)


def test_fasta_crlf(tmp_path):
    p = write(tmp_path, "a.fasta", b">r1\r\nACG\r\nTA\r\n>r2\r\nCC\r\n")
    recs, _ = records([p])
    assert recs == [b"ACGTA", b"CC"]


def test_fasta_comment_swallows_complete_lines(tmp_path):
    # After a ';' line, on_comment = !partial keeps consuming every complete
    # line -- the final call included -- until a piece cut by a block end
    # arrives.  In a small file that is the rest of the file: the pending
    # record (AAAA) and everything after the comment are lost.
    p = write(tmp_path, "c.fa", b">r1\nCCCC\n>r2\nAAAA\n;comment\nGGGG\n>r3\nTTTT\n")
    recs, _ = records([p])
    assert recs == [b"CCCC"]


def test_fasta_comment_ends_at_block_boundary(tmp_path):
    # A line that straddles the 4 MiB block end arrives as a partial piece,
    # which ends the comment state; that piece is dropped and the rest of the
    # line (no longer at a line start) is appended to the current record.
    head = b">r1\nAC\n;c\n"
    filler_line = b"G" * 99 + b"\n"
    body = bytearray(head)
    while len(body) + len(filler_line) <= BLOCK - 10:
        body += filler_line
    pad = BLOCK - len(body)
    body += b"T" * (pad + 5) + b"\n"  # this line straddles the block boundary
    body += b">r2\nCCC\n"
    p = write(tmp_path, "b.fa", bytes(body))
    recs, _ = records([p])
    # r1 = "AC" + the 5 T's after the boundary; r2 follows normally
    assert recs == [b"AC" + b"T" * 5, b"CCC"]


def test_fasta_gzip_equals_plain(tmp_path):
    rng = np.random.default_rng(5)
    seqs = [bytes(rng.choice(list(b"ACGT"), size=int(n)).astype(np.uint8)) for n in rng.integers(1, 500, 40)]
    text = b"".join(b">s%d\n" % i + s[:60] + b"\n" + s[60:] + b"\n" for i, s in enumerate(seqs))
    plain = write(tmp_path, "x.fa", text)
    gz = write(tmp_path, "x.fa.gz", text)
    rp, r1 = records([plain])
    rg, r2 = records([gz])
    assert rp == rg == seqs
    assert r2.estimated_bases == os.path.getsize(gz) // 2


def test_fastq_records(tmp_path):
    fq = b"@r1\nACGT\n+\nIIII\n@r2\nGGCC\n+r2\nIIII\n@r3\nTTA"
    recs, _ = records([write(tmp_path, "a.fq", fq)])
    # r3 has no quality line: emitted by the final call
    assert recs == [b"ACGT", b"GGCC", b"TTA"]
    recs, _ = records([write(tmp_path, "b.fastq.gz", fq.replace(b"\n", b"\r\n"))])
    assert recs == [b"ACGT", b"GGCC", b"TTA"]


def test_long_record_chunking(tmp_path):
    k = 31
    rng = np.random.default_rng(7)
    seq = rng.choice(list(b"ACGT"), size=9 * (1 << 20) + 12345).astype(np.uint8).tobytes()
    lines = [seq[i:i + 80] for i in range(0, len(seq), 80)]
    p = write(tmp_path, "long.fa", b">chr\n" + b"\n".join(lines) + b"\n")
    recs, _ = records([p], k=k)
    assert len(recs) == 3
    flush = 4 << 20
    # each chunk is emitted at the first line end where it reaches 4 MiB and
    # the next one starts with its last k-1 bases
    for a, b in zip(recs, recs[1:]):
        assert len(a) >= flush and len(a) < flush + 80
        assert b[:k - 1] == a[-(k - 1):]
    joined = recs[0] + b"".join(r[k - 1:] for r in recs[1:])
    assert joined == seq


def test_single_line_record_chunks_at_block_pieces(tmp_path):
    # A 9 MiB single-line record arrives as 4 MiB pieces; the chunk check runs
    # after every piece: the first piece (the block minus ">x\n") stays just
    # under 4 MiB, so the first chunk is the first two pieces.
    k = 21
    seq = (b"ACGTTGCA" * ((9 << 20) // 8))[: (9 << 20)]
    p = write(tmp_path, "one.fa", b">x\n" + seq + b"\n")
    recs, _ = records([p], k=k)
    assert len(recs) == 2 and len(recs[0]) == 2 * BLOCK - 3
    for a, b in zip(recs, recs[1:]):
        assert b[:k - 1] == a[-(k - 1):]
    assert recs[0] + b"".join(r[k - 1:] for r in recs[1:]) == seq


def test_colors_per_file(tmp_path):
    a = write(tmp_path, "a.fa", b">1\nAAAA\n>2\nCCCC\n")
    b = write(tmp_path, "b.fa", b">1\nGGGG\n")
    recs, r = records([a, b], colors=True)
    assert recs == [b"AAAA", b"CCCC", b"GGGG"]
    assert r.color_starts.tolist() == [0, 10]
    assert r.colors.tolist() == [0, 1]


def test_file_type_errors(tmp_path):
    bad = write(tmp_path, "a.txt", b">1\nA\n")
    with pytest.raises(native.GGCATGpuError, match="recognize"):
        native.read_files([str(bad)], 31)
    xz = write(tmp_path, "a.fa.xz", b"\x00")  # not an xz stream
    with pytest.raises(native.GGCATGpuError, match="error while reading"):
        native.read_files([str(xz)], 31)
    with pytest.raises(native.GGCATGpuError, match="cannot open"):
        native.read_files([str(tmp_path / "missing.fa")], 31)
    # extensions after the format are skipped one at a time
    ok = write(tmp_path, "a.fna.backup", b">1\nAC\n")
    assert records([ok])[0] == [b"AC"]


def test_example_inputs():
    names = ["sal1.fa", "sal2.fa", "sal3.fa", "query.fa"]
    for n in names:
        path = os.path.join(GOLDEN, n)
        recs, r = records([path])
        raw = open(path, "rb").read().split(b">")[1:]
        expect = [b"".join(x.split(b"\n")[1:]).replace(b"\r", b"") for x in raw]
        assert recs == expect
        assert r.estimated_bases == os.path.getsize(path)


@pytest.mark.parametrize("k,m", [(4, 2), (9, 5), (13, 9), (14, 9), (15, 9), (16, 10), (21, 10),
                                 (22, 11), (30, 11), (31, 12), (37, 12), (38, 13), (42, 13), (43, 14), (63, 14),
                                 (64, 14)])
def test_plan_best_m(k, m):
    # compute_best_m: max(k/2, k-4) up to 13, then a table
    p = native.plan_build(native.BuildOptions(k=k), 1 << 30)
    assert p.m == m


def test_plan_rejects_small_k_without_m():
    # below k = 4 the reference's k - 4 underflows; an explicit m is required
    with pytest.raises(native.GGCATGpuError):
        native.plan_build(native.BuildOptions(k=3), 1000)
    assert native.plan_build(native.BuildOptions(k=3, m=1), 1000).m == 1


@pytest.mark.parametrize("bases,b,sb", [
    (0, 2, 1),                      # clamps at 4 buckets, 2 second buckets
    (150_000, 2, 5),                # 37500/bucket -> 18 -> 32
    (1 << 30, 10, 6),               # 2048 by min size -> default 1024; 1 MiB/bucket -> default 64
    (4_000_000_000, 10, 6),
    (5 * (1 << 40), 13, 8),         # 5120 buckets by the max size -> 8192
    (3 << 19, 2, 6),                # 3 buckets by min size -> 4; 393216/bucket -> min(64, 192) = 64
])
def test_plan_bucket_logs(bases, b, sb):
    p = native.plan_build(native.BuildOptions(k=31, m=15, min_multiplicity=2), bases)
    assert (p.buckets_log, p.second_buckets_log) == (b, sb)
    assert (p.k, p.m, p.min_multiplicity) == (31, 15, 2)
    # an explicit bucket count wins; the second-bucket count stays automatic
    q = native.plan_build(native.BuildOptions(k=31, m=15, buckets_log=9), bases)
    assert (q.buckets_log, q.second_buckets_log) == (9, sb)


def test_many_files_keep_their_order(tmp_path):
    # files are read concurrently and concatenated in the order given
    rng = np.random.default_rng(3)
    paths, want = [], []
    for i in range(40):
        recs = [bytes(rng.choice(np.frombuffer(b"ACGT", np.uint8), int(rng.integers(1, 300))))
                for _ in range(int(rng.integers(1, 6)))]
        body = b"".join(b">r\n" + r + b"\n" for r in recs)
        paths.append(write(tmp_path, f"f{i}.fa", body))
        want.append(recs)
    recs, r = records(paths, colors=True)
    assert recs == [x for f in want for x in f]
    starts = np.cumsum([0] + [sum(len(x) + 1 for x in f) for f in want[:-1]])
    assert r.color_starts.tolist() == starts.tolist()
    assert r.colors.tolist() == list(range(40))
    bad = paths[:5] + [str(tmp_path / "missing.fa")] + paths[5:]
    with pytest.raises(native.GGCATGpuError, match="cannot open"):
        native.read_files(bad, 31)


def _compress(fmt: str, data: bytes) -> bytes:
    """One stream / frame of `data` in `fmt`, from Python's own modules (gzip,
    bz2, xz) or the system's liblz4 / libzstd through ctypes."""
    import bz2
    import ctypes
    import lzma

    if fmt == "gz":
        return gzip.compress(data)
    if fmt == "bz2":
        return bz2.compress(data)
    if fmt == "xz":
        return lzma.compress(data, format=lzma.FORMAT_XZ)
    if fmt == "zst":
        z = ctypes.CDLL("libzstd.so.1")
        z.ZSTD_compressBound.restype = ctypes.c_size_t
        z.ZSTD_compressBound.argtypes = [ctypes.c_size_t]
        z.ZSTD_compress.restype = ctypes.c_size_t
        z.ZSTD_compress.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        cap = z.ZSTD_compressBound(len(data))
        out = ctypes.create_string_buffer(cap)
        n = z.ZSTD_compress(out, cap, data, len(data), 3)
        return out.raw[:n]
    if fmt == "lz4":
        z = ctypes.CDLL("liblz4.so.1")
        z.LZ4F_compressFrameBound.restype = ctypes.c_size_t
        z.LZ4F_compressFrameBound.argtypes = [ctypes.c_size_t, ctypes.c_void_p]
        z.LZ4F_compressFrame.restype = ctypes.c_size_t
        z.LZ4F_compressFrame.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                         ctypes.c_void_p]
        cap = z.LZ4F_compressFrameBound(len(data), None)
        out = ctypes.create_string_buffer(cap)
        n = z.LZ4F_compressFrame(out, cap, data, len(data), None)
        return out.raw[:n]
    raise ValueError(fmt)


@pytest.mark.parametrize("fmt", ["gz", "bz2", "xz", "zst", "zstd", "lz4"])
def test_compressed_inputs_read_like_plain(tmp_path, fmt):
    # crates/io/src/lines_reader.rs:58-130 picks the decoder by the last
    # extension.  Content spanning several 4 MiB blocks, written as two
    # concatenated streams / frames, reads to the same records as the plain
    # file; so does a FASTQ file.
    rng = np.random.default_rng(99)
    parts = []
    for i in range(5000):
        seq = bytes(rng.choice(np.frombuffer(b"ACGTN", np.uint8), int(rng.integers(50, 4000))))
        parts.append(b">r%d\n" % i + b"\n".join(seq[j:j + 80] for j in range(0, len(seq), 80)) + b"\n")
    data = b"".join(parts)
    assert len(data) > 2 * BLOCK
    half = len(data) // 2
    cut = data.index(b">", half)
    plain = tmp_path / "x.fa"
    plain.write_bytes(data)
    comp = tmp_path / f"x.fa.{fmt}"
    comp.write_bytes(_compress("zst" if fmt == "zstd" else fmt, data[:cut]) +
                     _compress("zst" if fmt == "zstd" else fmt, data[cut:]))
    want, _ = records([plain])
    got, r = records([comp])
    assert got == want
    assert r.estimated_bases == os.path.getsize(comp) // 2
    fq = b"".join(b"@q%d\n%s\n+\n%s\n" % (i, s, b"I" * len(s)) for i, s in enumerate(want[:200]))
    pq, cq = tmp_path / "y.fq", tmp_path / f"y.fq.{fmt}"
    pq.write_bytes(fq)
    cq.write_bytes(_compress("zst" if fmt == "zstd" else fmt, fq))
    assert records([cq])[0] == records([pq])[0] == want[:200]


@pytest.mark.parametrize("fmt", ["bz2", "xz", "zst", "lz4"])
def test_truncated_compressed_input_is_an_error(tmp_path, fmt):
    data = b">r1\n" + b"ACGT" * 50000 + b"\n"
    c = _compress(fmt, data)
    p = tmp_path / f"t.fa.{fmt}"
    p.write_bytes(c[: len(c) // 2])
    with pytest.raises(native.GGCATGpuError):
        records([p])
