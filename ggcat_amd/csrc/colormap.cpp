// colormap.cpp -- the reference's color map file (`<graph>.colors.dat`,
// "GGCAT_CMAP_RNLEN" run-length storage) written and read on the host.
//
// File layout (crates/colors/src/storage/serializer.rs:17-241,
// run_length.rs:26-334, deserializer.rs:1-181):
//   [0, 64)   header: magic[16] "GGCAT_CMAP_RNLEN", then little-endian u64
//             version (1), index_offset, colors_count, subsets_count,
//             total_size (file bytes), total_uncompressed_size (bytes of the
//             encoded subsets before compression)
//   64..      one LZ4 frame holding bincode(Vec<String>) of the color names
//             (u64 count, then per name u64 length + UTF-8 bytes)
//   ...       the subsets in id order, COLORS_SINGLE_BATCH_SIZE (20000) per
//             LZ4 frame (crates/config/src/lib.rs:97)
//   index_offset: bincode(ColorsIndexMap) = u64 pair count, per chunk
//             (u32 first subset id, u64 file offset of its frame), u64
//             subsets_count
// A subset (ColorIndexSerializer::serialize_colors, run_length.rs:26-80):
// varint(first color + 2), then per following color varint(diff + 1), except
// that a run of at least three equal diffs is written varint(1), varint(diff),
// varint(run length); varint(0) ends the subset.  Varints: 7 bits per byte,
// low first (crates/io/src/varint.rs:8-22).
//
// LZ4 frames are produced and decoded through liblz4.so.1, loaded at run time
// like the input decoders (decompress.cpp).
//
// Included by api.hip.

namespace gg_host {

struct Lz4FrameApi {
    size_t (*bound)(size_t, const void*) = nullptr;
    size_t (*compress)(void*, size_t, const void*, size_t, const void*) = nullptr;
    bool ok = false;
};
inline const Lz4FrameApi& lz4_frame_api() {
    static Lz4FrameApi a = [] {
        Lz4FrameApi x;
        void* h = dl_open("liblz4.so.1");
        x.ok = h && dl_sym(h, "LZ4F_compressFrameBound", x.bound) && dl_sym(h, "LZ4F_compressFrame", x.compress) &&
               lz4_api().ok;
        return x;
    }();
    return a;
}

// Appends one LZ4 frame of src to out.
inline bool lz4_frame_append(const uint8_t* src, size_t n, std::vector<uint8_t>& out) {
    const Lz4FrameApi& z = lz4_frame_api();
    if (!z.ok) return false;
    const size_t cap = z.bound(n, nullptr);
    const size_t at = out.size();
    out.resize(at + cap);
    const size_t w = z.compress(out.data() + at, cap, src, n, nullptr);
    if (lz4_api().is_error(w)) {
        out.resize(at);
        return false;
    }
    out.resize(at + w);
    return true;
}

// Decodes the LZ4 frame starting at data[0]; *used = its compressed size.
inline bool lz4_frame_decode(const uint8_t* data, size_t n, std::vector<uint8_t>& out, size_t* used) {
    const Lz4Api& z = lz4_api();
    if (!z.ok) return false;
    void* dctx = nullptr;
    if (z.is_error(z.create(&dctx, 100 /* LZ4F_VERSION */))) return false;
    out.clear();
    std::vector<uint8_t> buf(1u << 16);
    size_t pos = 0;
    bool ok = false;
    while (true) {
        size_t dst = buf.size(), src = n - pos;
        const size_t r = z.decompress(dctx, buf.data(), &dst, data + pos, &src, nullptr);
        if (z.is_error(r)) break;
        pos += src;
        out.insert(out.end(), buf.begin(), buf.begin() + dst);
        if (r == 0) {  // the frame is complete
            ok = true;
            break;
        }
        if (src == 0 && dst == 0) break;  // truncated
    }
    z.free_(dctx);
    *used = pos;
    return ok;
}

inline void cm_varint(std::vector<uint8_t>& o, uint64_t v) {
    do {
        o.push_back((uint8_t)((v & 0x7F) | (v > 0x7F ? 0x80 : 0)));
        v >>= 7;
    } while (v);
}

inline bool cm_get_varint(const uint8_t* d, size_t n, size_t& p, uint64_t& v) {
    v = 0;
    for (uint32_t s = 0; s < 63; s += 7) {
        if (p >= n) return false;
        const uint8_t b = d[p++];
        v |= (uint64_t)(b & 0x7F) << s;
        if (!(b & 0x80)) return true;
    }
    return false;
}

// ColorIndexSerializer::serialize_colors (run_length.rs:26-80); colors
// strictly increasing, at least one.
inline void encode_subset(const uint32_t* c, size_t n, std::vector<uint8_t>& o) {
    cm_varint(o, (uint64_t)c[0] + 2);
    uint32_t last = c[0];
    bool run = false;
    uint32_t run_value = 0;
    uint64_t run_count = 0;
    for (size_t i = 1; i < n; i++) {
        const uint32_t diff = c[i] - last;
        bool flush = false;
        if (i + 1 < n) {
            const uint32_t next = c[i + 1] - c[i];
            if (diff == next) {
                if (!run) {
                    // a run pays off only with at least three equal diffs
                    if (i + 2 < n && c[i + 2] - c[i + 1] == diff) {
                        cm_varint(o, 1);
                        run_count = 2;
                        run = true;
                        run_value = diff;
                    }
                } else {
                    run_count++;
                }
            } else if (run) {
                flush = true;
            }
        } else if (run) {
            flush = true;
        }
        if (flush) {
            cm_varint(o, run_value);
            cm_varint(o, run_count);
            run = false;
        } else if (!run) {
            cm_varint(o, (uint64_t)diff + 1);
        }
        last = c[i];
    }
    cm_varint(o, 0);
}

// ColorIndexSerializer::deserialize_colors (run_length.rs:82-120).
inline bool decode_subset(const uint8_t* d, size_t n, size_t& p, std::vector<uint32_t>& out) {
    out.clear();
    uint64_t v;
    if (!cm_get_varint(d, n, p, v) || v < 2) return false;
    out.push_back((uint32_t)(v - 2));
    while (true) {
        if (!cm_get_varint(d, n, p, v)) return false;
        if (v == 0) break;
        if (v == 1) {
            uint64_t value, count;
            if (!cm_get_varint(d, n, p, value) || !cm_get_varint(d, n, p, count)) return false;
            if (count > (1u << 24)) return false;
            for (uint64_t i = 0; i < count; i++) out.push_back((uint32_t)value);
        } else {
            out.push_back((uint32_t)(v - 1));
        }
    }
    for (size_t i = 1; i < out.size(); i++) out[i] += out[i - 1];
    return true;
}

inline void put_u64(std::vector<uint8_t>& o, uint64_t v) {
    for (int i = 0; i < 8; i++) o.push_back((uint8_t)(v >> (8 * i)));
}
inline void put_u32(std::vector<uint8_t>& o, uint32_t v) {
    for (int i = 0; i < 4; i++) o.push_back((uint8_t)(v >> (8 * i)));
}
inline uint64_t get_u64(const uint8_t* p) {
    uint64_t v = 0;
    for (int i = 0; i < 8; i++) v |= (uint64_t)p[i] << (8 * i);
    return v;
}
inline uint32_t get_u32(const uint8_t* p) {
    uint32_t v = 0;
    for (int i = 0; i < 4; i++) v |= (uint32_t)p[i] << (8 * i);
    return v;
}

constexpr char CMAP_MAGIC[17] = "GGCAT_CMAP_RNLEN";
constexpr uint64_t CMAP_VERSION = 1;
constexpr uint64_t CMAP_BATCH = 20000;  // COLORS_SINGLE_BATCH_SIZE

// Writes a color map of `subsets` (each a strictly increasing color list, in
// id order) and `names`.
inline bool colormap_write(const std::string& path, const std::vector<std::string>& names,
                           const std::vector<std::vector<uint32_t>>& subsets, std::string* err) {
    std::vector<uint8_t> file(64, 0);
    {
        std::vector<uint8_t> b;
        put_u64(b, names.size());
        for (const std::string& s : names) {
            put_u64(b, s.size());
            b.insert(b.end(), s.begin(), s.end());
        }
        if (!lz4_frame_append(b.data(), b.size(), file)) {
            *err = "LZ4 frame compression unavailable (liblz4.so.1)";
            return false;
        }
    }
    std::vector<std::pair<uint32_t, uint64_t>> index;
    uint64_t uncompressed = 0;
    std::vector<uint8_t> chunk;
    for (uint64_t s0 = 0; s0 < subsets.size(); s0 += CMAP_BATCH) {
        chunk.clear();
        const uint64_t s1 = std::min<uint64_t>(subsets.size(), s0 + CMAP_BATCH);
        for (uint64_t s = s0; s < s1; s++) {
            if (subsets[s].empty()) {
                *err = "empty color subset";
                return false;
            }
            encode_subset(subsets[s].data(), subsets[s].size(), chunk);
        }
        uncompressed += chunk.size();
        index.push_back({(uint32_t)s0, (uint64_t)file.size()});
        if (!lz4_frame_append(chunk.data(), chunk.size(), file)) {
            *err = "LZ4 frame compression failed";
            return false;
        }
    }
    const uint64_t index_offset = file.size();
    put_u64(file, index.size());
    for (const auto& e : index) {
        put_u32(file, e.first);
        put_u64(file, e.second);
    }
    put_u64(file, subsets.size());
    std::vector<uint8_t> head(CMAP_MAGIC, CMAP_MAGIC + 16);
    put_u64(head, CMAP_VERSION);
    put_u64(head, index_offset);
    put_u64(head, names.size());
    put_u64(head, subsets.size());
    put_u64(head, file.size());
    put_u64(head, uncompressed);
    memcpy(file.data(), head.data(), 64);
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) {
        *err = "cannot create '" + path + "'";
        return false;
    }
    const bool ok = fwrite(file.data(), 1, file.size(), f) == file.size();
    if (fclose(f) != 0 || !ok) {
        *err = "write failed: '" + path + "'";
        return false;
    }
    return true;
}

struct ColorMap {
    std::vector<std::string> names;
    std::vector<uint64_t> subset_off;  // subsets_count + 1 offsets into colors
    std::vector<uint32_t> colors;
    uint64_t total_uncompressed = 0;
};

// Reads a whole color map (ColorsDeserializer::new + get_color_mappings for
// every subset).
inline bool colormap_read(const std::string& path, ColorMap& M, std::string* err) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) {
        *err = "cannot open '" + path + "'";
        return false;
    }
    std::vector<uint8_t> d;
    {
        std::vector<uint8_t> buf(1u << 20);
        size_t r;
        while ((r = fread(buf.data(), 1, buf.size(), f)) > 0) d.insert(d.end(), buf.begin(), buf.begin() + r);
        fclose(f);
    }
    auto bad = [&](const char* what) {
        *err = std::string("colormap '") + path + "': " + what;
        return false;
    };
    if (d.size() < 64 || memcmp(d.data(), CMAP_MAGIC, 16) != 0) return bad("header mismatch");
    const uint64_t index_offset = get_u64(&d[24]), n_colors = get_u64(&d[32]), n_subsets = get_u64(&d[40]);
    M.total_uncompressed = get_u64(&d[56]);
    if (get_u64(&d[16]) != CMAP_VERSION) return bad("unsupported version");
    if (index_offset > d.size() || index_offset < 64) return bad("index offset outside the file");
    std::vector<uint8_t> plain;
    size_t used = 0;
    if (!lz4_frame_decode(d.data() + 64, index_offset - 64, plain, &used)) return bad("color names frame");
    {
        size_t p = 0;
        if (plain.size() < 8) return bad("color names");
        const uint64_t cnt = get_u64(&plain[0]);
        p = 8;
        M.names.clear();
        for (uint64_t i = 0; i < cnt; i++) {
            if (p + 8 > plain.size()) return bad("color names");
            const uint64_t len = get_u64(&plain[p]);
            p += 8;
            if (len > plain.size() - p) return bad("color names");
            M.names.emplace_back(reinterpret_cast<const char*>(&plain[p]), (size_t)len);
            p += len;
        }
        if (M.names.size() != n_colors) return bad("color count differs from the header");
    }
    size_t p = index_offset;
    if (p + 8 > d.size()) return bad("index");
    const uint64_t n_pairs = get_u64(&d[p]);
    p += 8;
    if (n_pairs > (d.size() - p) / 12) return bad("index");
    std::vector<std::pair<uint32_t, uint64_t>> pairs(n_pairs);
    for (uint64_t i = 0; i < n_pairs; i++) {
        pairs[i] = {get_u32(&d[p]), get_u64(&d[p + 4])};
        p += 12;
    }
    if (p + 8 > d.size() || get_u64(&d[p]) != n_subsets) return bad("index subset count");
    M.subset_off.assign(1, 0);
    M.colors.clear();
    std::vector<uint32_t> one;
    for (uint64_t i = 0; i < n_pairs; i++) {
        const uint64_t end_id = i + 1 < n_pairs ? pairs[i + 1].first : n_subsets;
        if (pairs[i].first != M.subset_off.size() - 1 || end_id < pairs[i].first) return bad("chunk order");
        if (pairs[i].second >= index_offset) return bad("chunk offset");
        if (!lz4_frame_decode(d.data() + pairs[i].second, index_offset - pairs[i].second, plain, &used))
            return bad("subset chunk frame");
        size_t q = 0;
        for (uint64_t s = pairs[i].first; s < end_id; s++) {
            if (!decode_subset(plain.data(), plain.size(), q, one)) return bad("subset encoding");
            M.colors.insert(M.colors.end(), one.begin(), one.end());
            M.subset_off.push_back(M.colors.size());
        }
    }
    if (M.subset_off.size() != n_subsets + 1) return bad("subset count");
    return true;
}

}  // namespace gg_host
