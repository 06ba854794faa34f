"""Bucket-file record layout (ggcat_gpu_records_serialize): a reader written
from the reference's decoding side (read_from,
crates/io/src/concurrent/temp_reads/creads_utils.rs:434-470;
decode_varint / decode_varint_flags, crates/io/src/varint.rs:59-99), pinned
by the reference's own known answers for the varints and the 2-bit packing,
then used to decode the device's streams and compare them with the oracle's
records."""
from __future__ import annotations

import numpy as np
import pytest

from ggcat_amd import synth
from oracle import binding as O

HASH_SEQ = 0
CODE = {ord("A"): 0, ord("C"): 1, ord("T"): 2, ord("G"): 3}
LETTER = b"ACTG"


# ---- restatements of the reference's encoders / decoders -----------------
def enc_varint(v: int) -> bytes:
    """crates/io/src/varint.rs:8-22."""
    out = bytearray()
    while True:
        out.append((v & 0x7F) | (0x80 if v > 0x7F else 0))
        v >>= 7
        if v == 0:
            return bytes(out)


def dec_varint(buf: bytes, i: int) -> tuple[int, int]:
    v, sh = 0, 0
    while True:
        b = buf[i]
        i += 1
        v |= (b & 0x7F) << sh
        sh += 7
        if not b & 0x80:
            return v, i


def enc_varint_flags2(v: int, flags: int) -> bytes:
    """crates/io/src/varint.rs:26-57 with FlagsCount = U2."""
    out = bytearray([((flags << 6) & 0xFF) | (v & 0x1F) | (0x20 if v > 0x1F else 0)])
    v >>= 5
    while v:
        out.append((v & 0x7F) | (0x80 if v > 0x7F else 0))
        v >>= 7
    return bytes(out)


def dec_varint_flags2(buf: bytes, i: int) -> tuple[int, int, int]:
    """crates/io/src/varint.rs:59-82: (value, flags, next index)."""
    first = buf[i]
    i += 1
    flags, v, sh, more = first >> 6, first & 0x1F, 5, bool(first & 0x20)
    while more:
        b = buf[i]
        i += 1
        more = bool(b & 0x80)
        v |= (b & 0x7F) << sh
        sh += 7
    return v, flags, i


def pack_plain(seq: bytes) -> bytes:
    """crates/io/src/compressed_read.rs:610-618: chunks of 16 bases, first
    base lowest, little-endian, ceil(len / 4) bytes per chunk."""
    out = bytearray()
    for c in range(0, len(seq), 16):
        chunk = seq[c:c + 16]
        v = 0
        for ch in reversed(chunk):
            v = (v << 2) | CODE[ch]
        out += v.to_bytes(4, "little")[:(len(chunk) + 3) // 4]
    return bytes(out)


def pack_plain_rc(seq: bytes) -> bytes:
    """crates/io/src/compressed_read.rs:621-633: 16-base chunks from the end,
    complemented (code ^ 2), in the chunk's own order."""
    out = bytearray()
    cur = len(seq)
    while cur > 0:
        n = min(cur, 16)
        v = 0
        for ch in seq[cur - n:cur]:
            v = (v << 2) | (CODE[ch] ^ 2)
        out += v.to_bytes(4, "little")[:(n + 3) // 4]
        cur -= n
    return bytes(out)


def unpack(buf: bytes, n: int) -> bytes:
    return bytes(LETTER[(buf[i >> 2] >> (2 * (i & 3))) & 3] for i in range(n))


def decode_stream(buf: bytes, k: int, colored: bool) -> list[tuple]:
    """Records of one bucket stream: (second bucket, length, flags, minimizer
    position, bases, color, tail bits of the last byte)."""
    log = (k - 1).bit_length()  # k.next_power_of_two().ilog2()
    out, i, last = [], 0, 0xFFFFFFFF
    while i < len(buf):
        second = buf[i]
        v, flags, i = dec_varint_flags2(buf, i + 1)
        length = (v >> log) + k
        minpos = v & ((1 << log) - 1)
        nb = (length + 3) // 4
        packed = buf[i:i + nb]
        i += nb
        color = 0
        if colored:
            c, i = dec_varint(buf, i)
            color = last if c == 0 else c - 1
            last = color
        tail = packed[-1] >> (2 * (length & 3)) if length & 3 else 0
        out.append((second, length, flags, minpos, unpack(packed, length), color, tail))
    return out


# ---- the restatements against the reference's known answers ---------------
def test_varint_round_trips_and_known_bytes():
    # crates/io/src/varint.rs:107-136 (varints, varints_flags)
    for v in range(100000):
        assert dec_varint(enc_varint(v), 0) == (v, len(enc_varint(v)))
        b = enc_varint_flags2(v, v % 4)
        assert dec_varint_flags2(b, 0) == (v, v % 4, len(b))
    # worked from the definition: 5 bits + continuation bit under the flags
    assert enc_varint_flags2(5, 2) == bytes([0x85])
    assert enc_varint_flags2(100, 1) == bytes([0x64, 0x03])
    assert enc_varint(300) == bytes([0xAC, 0x02])


def test_packing_known_answers():
    # crates/io/src/compressed_read.rs:1003-1023 (test_compression,
    # test_rc_compression)
    bases = b"ACGTACGCGGTAGCTAAGCATCGATGCCGATCGTGTTTAACCATG"
    expected_rc = b"CATGGTTAAACACGATCGGCATCGATGCTTAGCTACCGCGTACGT"
    assert unpack(pack_plain(bases), len(bases)) == bases
    assert unpack(pack_plain_rc(bases), len(bases)) == expected_rc
    assert pack_plain(b"ACTG") == bytes([0b11100100])


# ---- the device's streams ---------------------------------------------------
def _oracle_streams(o, k: int, sb: int, colored: bool) -> dict:
    r = o.records()
    want = {}
    for i in range(r["sub"].size):
        off, n = int(r["off"][i]), int(r["length"][i])
        bucket = int(r["sub"][i]) >> sb
        want.setdefault(bucket, []).append((int(r["sub"][i]) & ((1 << sb) - 1), n, int(r["flags"][i]),
                                            int(r["minpos"][i]), r["bases"][off:off + n].tobytes(),
                                            int(r["color"][i]) if colored else 0, 0))
    return {b: sorted(v) for b, v in want.items()}


def _device_streams(ctx, k: int, colored: bool, b: int) -> dict:
    buf, off = ctx.records_serialize()
    assert off.size == (1 << b) + 2  # 2^b buckets and the duplicates bucket
    got = {}
    for b in range(off.size - 1):
        s = buf[int(off[b]):int(off[b + 1])]
        if s:
            got[b] = sorted(decode_stream(s, k, colored))
    assert int(off[-1]) == len(buf)
    return got


@pytest.fixture(scope="module")
def emu_ctx():
    from ggcat_amd.native import GpuContext
    from tests.emu import build as emu_build
    ctx = GpuContext(0, lib_path=emu_build.build())
    yield ctx
    ctx.close()


@pytest.mark.parametrize("k,m", [(31, 15), (45, 21)])
def test_serialized_streams_decode_to_oracle_records(emu_ctx, k, m):
    from ggcat_amd.native import Params
    g = synth.genome(91, 20_000, repeat_frac=0.2, rep_min=300, rep_max=1000)
    data = synth.reads(g, 92, 800, 150, 0.01)
    emu_ctx.bucketing(data, Params(k, m, 3, 2, 0, HASH_SEQ, 2))
    o = O.bucketing(data, k, m, 3, 2, threads=2)
    assert _device_streams(emu_ctx, k, False, 3) == _oracle_streams(o, k, 2, False)


def test_serialized_colored_streams(emu_ctx):
    from ggcat_amd.native import Params
    gs = synth.config3_genomes(0.002, 6)
    data, starts = synth.contigs_joined(gs, 5, 3000, 9000)
    colors = (np.arange(6, dtype=np.uint32) * 7) % 40
    emu_ctx.bucketing(data, Params(31, 15, 3, 2, 0, HASH_SEQ, 1), colors=(starts, colors))
    o = O.bucketing_colored(data, 31, 15, 3, 2, starts, colors, threads=2)
    assert _device_streams(emu_ctx, 31, True, 3) == _oracle_streams(o, 31, 2, True)


def test_empty_inputs_emulated(emu_ctx, tmp_path):
    """Empty and separator-only inputs through every output of a build."""
    from ggcat_amd.native import Params
    for data in (np.zeros(0, np.uint8), np.frombuffer(b"\n\nNNNN\nACGT\n", np.uint8).copy()):
        emu_ctx.bucketing(data, Params(31, 15, 4, 2, 0, HASH_SEQ, 2))
        buf, off = emu_ctx.records_serialize()
        assert buf == b"" and off.size == (1 << 4) + 2 and not off.any()
        emu_ctx.kmers_merge()
        assert emu_ctx.unitigs()[0] == [] and emu_ctx.unitigs(joined=True)[0] == []
        path = str(tmp_path / "empty.fa")
        assert emu_ctx.write_unitigs_fasta(path) == 0 and open(path, "rb").read() == b""
    from ggcat_amd.native import GGCATGpuError
    with pytest.raises(GGCATGpuError, match="cannot create"):
        emu_ctx.write_unitigs_fasta(str(tmp_path / "missing-dir" / "u.fa"))
