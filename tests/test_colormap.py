"""The color map file (<graph>.colors.dat) in the reference's run-length
format, host code only (no GPU): the library's writer and reader against an
independent Python restatement of the format
(crates/colors/src/storage/serializer.rs:17-241, run_length.rs:26-334,
deserializer.rs:1-181), and the subset encoding against known answers worked
out from ColorIndexSerializer::serialize_colors on the reference's own test
vectors (run_length.rs:300-333)."""
from __future__ import annotations

import ctypes
import struct

import numpy as np
import pytest

from ggcat_amd import native

MAGIC = b"GGCAT_CMAP_RNLEN"
BATCH = 20000  # COLORS_SINGLE_BATCH_SIZE, crates/config/src/lib.rs:97


def varint(v: int) -> bytes:
    out = bytearray()
    while True:
        out.append((v & 0x7F) | (0x80 if v > 0x7F else 0))
        v >>= 7
        if v == 0:
            return bytes(out)


def get_varint(b: bytes, p: int) -> tuple[int, int]:
    v, s = 0, 0
    while True:
        x = b[p]
        p += 1
        v |= (x & 0x7F) << s
        s += 7
        if not x & 0x80:
            return v, p


def encode_subset(c: list[int]) -> bytes:
    """run_length.rs:26-80, restated."""
    out = bytearray(varint(c[0] + 2))
    last, run, run_value, run_count = c[0], False, 0, 0
    for i in range(1, len(c)):
        diff = c[i] - last
        flush = False
        if i + 1 < len(c):
            if diff == c[i + 1] - c[i]:
                if not run:
                    if i + 2 < len(c) and c[i + 2] - c[i + 1] == diff:
                        out += varint(1)
                        run, run_count, run_value = True, 2, diff
                else:
                    run_count += 1
            elif run:
                flush = True
        elif run:
            flush = True
        if flush:
            out += varint(run_value) + varint(run_count)
            run = False
        elif not run:
            out += varint(diff + 1)
        last = c[i]
    return bytes(out + varint(0))


def decode_subset(b: bytes, p: int) -> tuple[list[int], int]:
    """run_length.rs:82-120, restated."""
    v, p = get_varint(b, p)
    diffs = [v - 2]
    while True:
        v, p = get_varint(b, p)
        if v == 0:
            break
        if v == 1:
            value, p = get_varint(b, p)
            count, p = get_varint(b, p)
            diffs += [value] * count
        else:
            diffs.append(v - 1)
    return np.cumsum(diffs).tolist(), p


def _lz4():
    z = ctypes.CDLL("liblz4.so.1")
    z.LZ4F_compressFrameBound.restype = ctypes.c_size_t
    z.LZ4F_compressFrameBound.argtypes = [ctypes.c_size_t, ctypes.c_void_p]
    z.LZ4F_compressFrame.restype = ctypes.c_size_t
    z.LZ4F_compressFrame.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_void_p]
    z.LZ4F_createDecompressionContext.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_uint]
    z.LZ4F_decompress.restype = ctypes.c_size_t
    z.LZ4F_decompress.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t),
                                  ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t), ctypes.c_void_p]
    z.LZ4F_isError.argtypes = [ctypes.c_size_t]
    z.LZ4F_freeDecompressionContext.argtypes = [ctypes.c_void_p]
    return z


def lz4_frame(data: bytes) -> bytes:
    z = _lz4()
    cap = z.LZ4F_compressFrameBound(len(data), None)
    out = ctypes.create_string_buffer(cap)
    n = z.LZ4F_compressFrame(out, cap, data, len(data), None)
    return out.raw[:n]


def lz4_unframe(data: bytes, p: int) -> tuple[bytes, int]:
    """One LZ4 frame starting at data[p] -> (content, offset after it)."""
    z = _lz4()
    ctx = ctypes.c_void_p()
    z.LZ4F_createDecompressionContext(ctypes.byref(ctx), 100)
    out = bytearray()
    buf = ctypes.create_string_buffer(1 << 16)
    src = ctypes.create_string_buffer(bytes(data[p:]), len(data) - p)
    pos = 0
    try:
        while True:
            dst_n = ctypes.c_size_t(len(buf))
            src_n = ctypes.c_size_t(len(data) - p - pos)
            r = z.LZ4F_decompress(ctx, buf, ctypes.byref(dst_n), ctypes.addressof(src) + pos, ctypes.byref(src_n),
                                  None)
            assert not z.LZ4F_isError(r)
            pos += src_n.value
            out += buf.raw[:dst_n.value]
            if r == 0:
                return bytes(out), p + pos
            assert src_n.value or dst_n.value
    finally:
        z.LZ4F_freeDecompressionContext(ctx)


def parse_colormap(raw: bytes) -> tuple[list[str], list[list[int]], dict]:
    """Independent reader: header, names frame, index, chunk frames."""
    assert raw[:16] == MAGIC
    version, index_offset, n_colors, n_subsets, total, uncompressed = struct.unpack_from("<6Q", raw, 16)
    names_raw, _ = lz4_unframe(raw, 64)
    (cnt,) = struct.unpack_from("<Q", names_raw, 0)
    names, q = [], 8
    for _ in range(cnt):
        (ln,) = struct.unpack_from("<Q", names_raw, q)
        names.append(names_raw[q + 8:q + 8 + ln].decode())
        q += 8 + ln
    (n_pairs,) = struct.unpack_from("<Q", raw, index_offset)
    pairs = [struct.unpack_from("<IQ", raw, index_offset + 8 + 12 * i) for i in range(n_pairs)]
    (idx_subsets,) = struct.unpack_from("<Q", raw, index_offset + 8 + 12 * n_pairs)
    subsets, plain_total = [], 0
    for i, (start, off) in enumerate(pairs):
        end = pairs[i + 1][0] if i + 1 < n_pairs else idx_subsets
        plain, _ = lz4_unframe(raw, off)
        plain_total += len(plain)
        p = 0
        for _ in range(start, end):
            s, p = decode_subset(plain, p)
            subsets.append(s)
        assert p == len(plain)
    hdr = dict(version=version, index_offset=index_offset, colors=n_colors, subsets=n_subsets, total=total,
               uncompressed=uncompressed, pairs=pairs, plain_total=plain_total, idx_subsets=idx_subsets)
    return names, subsets, hdr


def write_colormap_py(path, names: list[str], subsets: list[list[int]]) -> None:
    """Independent writer (serializer.rs:57-120 + run_length.rs:240-296)."""
    body = bytearray(64)
    nb = struct.pack("<Q", len(names)) + b"".join(struct.pack("<Q", len(n.encode())) + n.encode() for n in names)
    body += lz4_frame(nb)
    pairs, unc = [], 0
    for s0 in range(0, len(subsets), BATCH):
        chunk = b"".join(encode_subset(s) for s in subsets[s0:s0 + BATCH])
        unc += len(chunk)
        pairs.append((s0, len(body)))
        body += lz4_frame(chunk)
    index_offset = len(body)
    body += struct.pack("<Q", len(pairs)) + b"".join(struct.pack("<IQ", a, b) for a, b in pairs)
    body += struct.pack("<Q", len(subsets))
    body[:64] = MAGIC + struct.pack("<6Q", 1, index_offset, len(names), len(subsets), len(body), unc)
    open(path, "wb").write(bytes(body))


KNOWN = [  # the reference's test vectors (run_length.rs:318-332), encoded by hand from serialize_colors
    ([0, 1, 2, 3, 4, 5, 6, 7], bytes([2, 1, 1, 7, 0])),
    ([0], bytes([2, 0])),
    ([1], bytes([3, 0])),
    ([1, 2, 5, 10, 15, 30, 45], bytes([3, 2, 4, 6, 6, 16, 16, 0])),
    ([1, 100, 200, 300, 400, 800], bytes([3, 100, 1, 100, 3, 0x91, 3, 0])),
]


@pytest.mark.parametrize("colors,enc", KNOWN)
def test_subset_encoding_known_answers(colors, enc):
    assert encode_subset(colors) == enc
    assert decode_subset(enc, 0) == (colors, len(enc))


def test_subset_encoding_round_trip_reference_vectors():
    v = [3, 6, 9, 12, 70, 71, 72, 73, 74, 88, 95, 100, 105, 110, 198, 384]
    assert decode_subset(encode_subset(v), 0)[0] == v


def _random_subsets(rng, n, n_colors):
    out = []
    for _ in range(n):
        kind = rng.integers(0, 3)
        if kind == 0:
            s = sorted(set(rng.integers(0, n_colors, int(rng.integers(1, 12))).tolist()))
        elif kind == 1:  # arithmetic runs (second-order encoding)
            a, d, c = int(rng.integers(0, 50)), int(rng.integers(1, 5)), int(rng.integers(1, 40))
            s = [a + d * i for i in range(c) if a + d * i < n_colors] or [a]
        else:
            s = list(range(int(rng.integers(0, n_colors - 1)), n_colors))
        out.append(s)
    return out


def test_library_writer_matches_the_format(tmp_path):
    rng = np.random.default_rng(5)
    names = [f"genome_{i}.fa" for i in range(300)] + ['quo"te\\slash']
    subsets = _random_subsets(rng, 45_000, len(names))
    path = tmp_path / "g.colors.dat"
    native.colormap_write_sets(str(path), names, subsets)
    raw = path.read_bytes()
    got_names, got_subsets, hdr = parse_colormap(raw)
    assert got_names == names and got_subsets == subsets
    assert hdr["version"] == 1 and hdr["colors"] == len(names) and hdr["subsets"] == len(subsets)
    assert hdr["idx_subsets"] == len(subsets) and hdr["total"] == len(raw)
    assert [a for a, _ in hdr["pairs"]] == [0, 20000, 40000]  # one frame per 20000 subsets
    assert hdr["uncompressed"] == hdr["plain_total"] == sum(len(encode_subset(s)) for s in subsets)


def test_library_reader_reads_an_independent_file(tmp_path):
    rng = np.random.default_rng(6)
    names = [f"c{i}" for i in range(64)]
    subsets = _random_subsets(rng, 21_003, 64)
    path = tmp_path / "x.colors.dat"
    write_colormap_py(path, names, subsets)
    assert native.colormap_read(str(path)) == (names, subsets)


def test_reader_rejects_damaged_files(tmp_path):
    names, subsets = ["a", "b"], [[0], [0, 1], [1]]
    good = tmp_path / "g.colors.dat"
    native.colormap_write_sets(str(good), names, subsets)
    raw = good.read_bytes()
    bad_magic = tmp_path / "m.colors.dat"
    bad_magic.write_bytes(b"X" + raw[1:])
    trunc = tmp_path / "t.colors.dat"
    trunc.write_bytes(raw[: len(raw) - 10])
    for p in (bad_magic, trunc, tmp_path / "missing.colors.dat"):
        with pytest.raises(native.GGCATGpuError):
            native.colormap_read(str(p))
    with pytest.raises(native.GGCATGpuError):  # colors must be increasing and named
        native.colormap_write_sets(str(tmp_path / "e.colors.dat"), names, [[1, 0]])
    with pytest.raises(native.GGCATGpuError):
        native.colormap_write_sets(str(tmp_path / "e.colors.dat"), names, [[2]])
