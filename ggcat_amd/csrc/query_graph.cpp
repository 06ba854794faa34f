// query_graph.cpp -- querying a graph given as files: the unitig FASTA of a
// build and, for colored graphs, its color map.  Mirrors ggcat_query_graph
// (crates/capi/src/lib.rs:635-651 -> crates/api/src/lib.rs:325-382 ->
// crates/querier/src/lib.rs:54-199): every k-mer of the graph's unitigs is a
// graph k-mer; with colors, the k-mer's color set is the subset named by the
// unitig header's " C:{id hex}:{count}" runs, read through the color map
// `<graph>.colors.dat` (Path::with_extension).  The counting is the device
// query (query.hip) over an index of those k-mers; the output is the
// reference's CSV (crates/querier/src/pipeline/counters_sorting.rs:235-270)
// or JSON lines (crates/querier/src/pipeline/colored_query_output.rs:95-320),
// compressed by the output file's extension like the reference's writer
// (lz4, gz, zst / zstd, bz2, xz; anything else plain).
//
// Included by api.hip (host code).
#include <cmath>

namespace gg_host {

// Path::with_extension: the last extension of the file name replaced (or added).
inline std::string with_extension(const std::string& path, const std::string& ext) {
    const size_t slash = path.find_last_of('/');
    const size_t name0 = slash == std::string::npos ? 0 : slash + 1;
    const size_t dot = path.find_last_of('.');
    std::string stem = path;
    if (dot != std::string::npos && dot > name0) stem = path.substr(0, dot);
    return stem + "." + ext;
}

inline bool has_extension(const std::string& path) {
    const size_t slash = path.find_last_of('/');
    const size_t name0 = slash == std::string::npos ? 0 : slash + 1;
    const size_t dot = path.find_last_of('.');
    return dot != std::string::npos && dot > name0 && dot + 1 < path.size();
}

struct GraphRecord {
    uint64_t seq_off, seq_len;
    std::string header;
};

// The records of a FASTA file: header line (without '>') and the
// concatenated sequence lines.
inline bool read_graph_fasta(const std::string& path, std::vector<uint8_t>& seq, std::vector<GraphRecord>& recs,
                             std::string* err) {
    std::unique_ptr<ByteSource> src = open_source(path, last_ext(path), err);
    if (!src) return false;
    std::vector<uint8_t> buf(4u << 20);
    std::string line;
    bool in_header = false;
    auto finish_line = [&]() {
        if (in_header) {
            recs.back().header = line;
            in_header = false;
        } else if (!line.empty() && !recs.empty()) {
            seq.insert(seq.end(), line.begin(), line.end());
            recs.back().seq_len += line.size();
        }
        line.clear();
    };
    bool line_start = true;
    while (true) {
        const long r = src->read(buf.data(), buf.size());
        if (r < 0) {
            *err = "error while reading '" + path + "'";
            return false;
        }
        if (r == 0) break;
        for (long i = 0; i < r; i++) {
            const uint8_t c = buf[i];
            if (line_start && c == '>') {
                recs.push_back(GraphRecord{seq.size(), 0, std::string()});
                in_header = true;
                line_start = false;
                continue;
            }
            line_start = false;
            if (c == '\n') {
                if (!line.empty() && line.back() == '\r') line.pop_back();
                finish_line();
                line_start = true;
            } else {
                line.push_back((char)c);
            }
        }
    }
    if (!line.empty()) finish_line();
    return true;
}

// " C:{hex}:{count}" tags of a unitig header, expanded to one subset id per k-mer.
inline bool header_subsets(const std::string& h, uint64_t n_kmers, std::vector<uint32_t>& out) {
    out.clear();
    size_t p = 0;
    while ((p = h.find(" C:", p)) != std::string::npos) {
        p += 3;
        const size_t colon = h.find(':', p);
        if (colon == std::string::npos) return false;
        char* end = nullptr;
        const unsigned long long id = strtoull(h.c_str() + p, &end, 16);
        if (end != h.c_str() + colon) return false;
        const unsigned long long cnt = strtoull(h.c_str() + colon + 1, &end, 10);
        if (end == h.c_str() + colon + 1 || cnt == 0 || out.size() + cnt > n_kmers) return false;
        out.insert(out.end(), (size_t)cnt, (uint32_t)id);
        p = (size_t)(end - h.c_str());
    }
    return out.size() == n_kmers;
}

// Rust's "{:.2}" of count / denominator as f64 (NaN for 0 / 0, inf for x / 0).
inline std::string ratio2(uint64_t num, uint64_t den) {
    const double r = (double)num / (double)den;
    if (r != r) return "NaN";
    if (r == INFINITY) return "inf";
    char b[64];
    snprintf(b, sizeof b, "%.2f", r);
    return b;
}

// ColorsDeserializer's json_escaped_color_names
// (crates/colors/src/storage/deserializer.rs:106-109): '"' -> '\"', then every
// '\' (including the ones just added) doubled.
inline std::string json_escaped_name(const std::string& s) {
    std::string a;
    for (char c : s) {
        if (c == '"') a += "\\\"";
        else a += c;
    }
    std::string b;
    for (char c : a) {
        if (c == '\\') b += "\\\\";
        else b += c;
    }
    return b;
}

// Compresses the whole output by the file's extension (one stream / frame).
inline bool write_output(const std::string& path, const std::string& text, std::string* err) {
    const std::string ext = last_ext(path);
    std::vector<uint8_t> out;
    const uint8_t* src = reinterpret_cast<const uint8_t*>(text.data());
    if (ext == "lz4") {
        if (!lz4_frame_append(src, text.size(), out)) {
            *err = "lz4 output needs liblz4.so.1";
            return false;
        }
    } else if (ext == "gz") {
        gzFile f = gzopen(path.c_str(), "wb");
        if (!f) {
            *err = "cannot create '" + path + "'";
            return false;
        }
        const bool ok = text.empty() || gzwrite(f, text.data(), (unsigned)text.size()) == (int)text.size();
        if (gzclose(f) != Z_OK || !ok) {
            *err = "write failed: '" + path + "'";
            return false;
        }
        return true;
    } else if (ext == "zst" || ext == "zstd") {
        void* h = dl_open("libzstd.so.1");
        size_t (*bound)(size_t) = nullptr;
        size_t (*comp)(void*, size_t, const void*, size_t, int) = nullptr;
        unsigned (*is_err)(size_t) = nullptr;
        if (!h || !dl_sym(h, "ZSTD_compressBound", bound) || !dl_sym(h, "ZSTD_compress", comp) ||
            !dl_sym(h, "ZSTD_isError", is_err)) {
            *err = "zstd output needs libzstd.so.1";
            return false;
        }
        out.resize(bound(text.size()));
        const size_t w = comp(out.data(), out.size(), src, text.size(), 3);
        if (is_err(w)) {
            *err = "zstd compression failed";
            return false;
        }
        out.resize(w);
    } else if (ext == "bz2") {
        void* h = dl_open("libbz2.so.1");
        int (*comp)(char*, unsigned*, char*, unsigned, int, int, int) = nullptr;
        if (!h || !dl_sym(h, "BZ2_bzBuffToBuffCompress", comp)) {
            *err = "bz2 output needs libbz2.so.1";
            return false;
        }
        unsigned n = (unsigned)(text.size() + text.size() / 100 + 600);
        out.resize(n);
        if (comp(reinterpret_cast<char*>(out.data()), &n, const_cast<char*>(text.data()), (unsigned)text.size(), 9, 0,
                 0) != 0) {
            *err = "bz2 compression failed";
            return false;
        }
        out.resize(n);
    } else if (ext == "xz") {
        void* h = dl_open("liblzma.so.5");
        int (*enc)(uint32_t, int, const void*, const uint8_t*, size_t, uint8_t*, size_t*, size_t) = nullptr;
        size_t (*bound)(size_t) = nullptr;
        if (!h || !dl_sym(h, "lzma_easy_buffer_encode", enc) || !dl_sym(h, "lzma_stream_buffer_bound", bound)) {
            *err = "xz output needs liblzma.so.5";
            return false;
        }
        out.resize(bound(text.size()));
        size_t pos = 0;
        if (enc(6, 4 /* LZMA_CHECK_CRC64 */, nullptr, src, text.size(), out.data(), &pos, out.size()) != 0) {
            *err = "xz compression failed";
            return false;
        }
        out.resize(pos);
    } else {
        out.assign(src, src + text.size());
    }
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) {
        *err = "cannot create '" + path + "'";
        return false;
    }
    const bool ok = out.empty() || fwrite(out.data(), 1, out.size(), f) == out.size();
    if (fclose(f) != 0 || !ok) {
        *err = "write failed: '" + path + "'";
        return false;
    }
    return true;
}

}  // namespace gg_host

namespace {

// Canonical key of the k-mer at s[0..k) on the host, in the device's
// representation (seq-hash: the 2-bit k-mer, base i at bits 2i, min of both
// strands; rabin-karp128: the rolling hash's min of both strands,
// crates/hashes/src/base/cn_rkhash_base.rs:65-110, computed directly).
struct HostKeys {
    uint32_t k, kind;
    MergeArgs rk{};  // only the rabin-karp128 constants are used
    u128_t key(const uint8_t* s) const {
        if (kind == GGCAT_GPU_HASH_SEQ) {
            u128_t fw = 0, rc = 0;
            for (uint32_t i = 0; i < k; i++) {
                const u128_t c = (s[i] >> 1) & 3;
                fw |= c << (2 * i);
                rc |= (c ^ 2) << (2 * (k - 1 - i));
            }
            return fw < rc ? fw : rc;
        }
        u128_t f = 0, b = 0;
        for (uint32_t i = 0; i < k; i++) f = f * rk.rk_mult + rk.rk_fwd[(s[i] >> 1) & 3];
        for (int i = (int)k - 1; i >= 0; i--) b = b * rk.rk_mult + rk.rk_bkw[(s[i] >> 1) & 3];
        return f < b ? f : b;
    }
};

bool acgt(uint8_t c) {
    const uint8_t u = c & 0xDF;
    return u == 'A' || u == 'C' || u == 'G' || u == 'T';
}

}  // namespace

extern "C" int ggcat_gpu_query_graph(ggcat_gpu_ctx* ctx, const char* input_graph, const char* input_query,
                                     const char* output_file_prefix, uint32_t k, uint32_t threads_count,
                                     int forward_only, uint32_t minimizer_length, int colors,
                                     uint32_t color_output_format, uint32_t hash_kind, char* out_path,
                                     uint64_t out_path_cap) {
    (void)threads_count;
    (void)minimizer_length;
    if (!ctx || !input_graph || !input_query || !output_file_prefix) return GGCAT_GPU_EINVAL;
    if (k < 2 || k > 64) return fail(ctx, GGCAT_GPU_EINVAL, "k must be in [2, 64]");
    if (forward_only) return fail(ctx, GGCAT_GPU_EINVAL, "forward-only queries are not supported (canonical k-mers only)");
    if (color_output_format > 1) return fail(ctx, GGCAT_GPU_EINVAL, "color_output_format must be 0 (numbers) or 1 (names)");
    if (hash_kind != GGCAT_GPU_HASH_SEQ && hash_kind != GGCAT_GPU_HASH_RK128)
        return fail(ctx, GGCAT_GPU_EINVAL, "unknown hash kind");
    if (ctx->server) return fail(ctx, GGCAT_GPU_ESTATE, "close the query server before loading another graph");
    std::string err;
    std::vector<uint8_t> gseq;
    std::vector<gg_host::GraphRecord> recs;
    if (!gg_host::read_graph_fasta(input_graph, gseq, recs, &err)) return fail(ctx, GGCAT_GPU_EINVAL, err);
    gg_host::ColorMap cmap;
    if (colors && !gg_host::colormap_read(gg_host::with_extension(input_graph, "colors.dat"), cmap, &err))
        return fail(ctx, GGCAT_GPU_EINVAL, err);
    const uint32_t n_colors = colors ? (uint32_t)cmap.names.size() : 0;
    if (n_colors > 16384) return fail(ctx, GGCAT_GPU_EINVAL, "more than 16384 colors");
    const uint32_t W = colors ? std::max<uint32_t>(1, (n_colors + 31) / 32) : 0;
    const uint32_t kw = (hash_kind == GGCAT_GPU_HASH_SEQ && k <= 32) ? 1 : 2;
    HostKeys hk{k, hash_kind};
    set_rk_constants(hk.rk, k);
    std::vector<uint64_t> keys;
    std::vector<uint32_t> bitmaps;
    std::vector<uint32_t> sub_ids;
    std::vector<uint32_t> subset_bits;  // per subset, W words
    if (colors) {
        const uint64_t ns = cmap.subset_off.size() - 1;
        subset_bits.assign(ns * W, 0);
        for (uint64_t s = 0; s < ns; s++)
            for (uint64_t i = cmap.subset_off[s]; i < cmap.subset_off[s + 1]; i++) {
                const uint32_t c = cmap.colors[i];
                if (c >= n_colors) return fail(ctx, GGCAT_GPU_EINVAL, "color map subset with an unnamed color");
                subset_bits[s * W + c / 32] |= 1u << (c % 32);
            }
    }
    for (const gg_host::GraphRecord& r : recs) {
        if (r.seq_len < k) continue;
        const uint64_t nk = r.seq_len - k + 1;
        if (colors && !gg_host::header_subsets(r.header, nk, sub_ids))
            return fail(ctx, GGCAT_GPU_EINVAL, "unitig header '" + r.header + "': color runs do not cover its k-mers");
        const uint8_t* s = gseq.data() + r.seq_off;
        uint64_t bad = 0;  // last non-ACGT base + 1
        for (uint64_t i = 0; i < r.seq_len; i++) {
            if (!acgt(s[i])) bad = i + 1;
            if (i + 1 < k || bad > i + 1 - k) continue;
            const uint64_t j = i + 1 - k;
            const u128_t key = hk.key(s + j);
            keys.push_back((uint64_t)key);
            if (kw == 2) keys.push_back((uint64_t)(key >> 64));
            if (colors) {
                if (sub_ids[j] >= cmap.subset_off.size() - 1)
                    return fail(ctx, GGCAT_GPU_EINVAL, "unitig color subset id outside the color map");
                bitmaps.insert(bitmaps.end(), subset_bits.begin() + (uint64_t)sub_ids[j] * W,
                               subset_bits.begin() + (uint64_t)(sub_ids[j] + 1) * W);
            }
        }
    }
    const uint64_t n = keys.size() / kw;
    // the graph k-mers become the context's table (every entry in the tail
    // part; no chunks), then the device index and query path take over
    HIPCHK(hipSetDevice(ctx->device));
    ENSURE(ctx->out_key, n * 8 * kw + 16);
    ENSURE(ctx->out_colors, (uint64_t)n * 4 * std::max<uint32_t>(W, 1) + 16);
    if (n) HIPCHK(hipMemcpy(ctx->out_key.p, keys.data(), n * 8 * kw, hipMemcpyHostToDevice));
    if (n && W) HIPCHK(hipMemcpy(ctx->out_colors.p, bitmaps.data(), (uint64_t)n * 4 * W, hipMemcpyHostToDevice));
    ctx->have_records = ctx->have_unitigs = ctx->have_subsets = ctx->have_index = false;
    ctx->n_chunks_used = 0;
    ctx->chunk_fill.clear();
    ctx->n_tail = n;
    ctx->n_solid = n;
    ctx->key_words = kw;
    ctx->colored = colors != 0;
    ctx->color_words = W;
    ctx->n_color_ids = n_colors;
    ctx->P = ggcat_gpu_params{};
    ctx->P.k = k;
    ctx->P.hash_kind = hash_kind;
    ctx->have_solid = true;
    int r = ggcat_gpu_index_build(ctx);
    if (r != GGCAT_GPU_OK) return r;
    gg_host::ReadResult Q;
    if (!gg_host::read_files({std::string(input_query)}, k, false, Q)) return fail(ctx, GGCAT_GPU_EINVAL, Q.error);
    const uint32_t nq = (uint32_t)Q.records;
    std::vector<uint64_t> qk(nq), mt(nq);
    std::vector<uint32_t> cc((uint64_t)nq * n_colors);
    r = ggcat_gpu_query(ctx, Q.joined.data(), Q.joined.size(), nq, qk.data(), mt.data(), n_colors ? cc.data() : nullptr);
    if (r != GGCAT_GPU_OK) return r;
    std::string out_file = output_file_prefix;
    if (!gg_host::has_extension(out_file)) out_file = gg_host::with_extension(out_file, colors ? "jsonl" : "csv");
    std::string text;
    if (!colors) {
        text = "query_index,matched_kmers,query_kmers,match_percentage\n";
        for (uint32_t q = 0; q < nq; q++)
            text += std::to_string(q) + "," + std::to_string(mt[q]) + "," + std::to_string(qk[q]) + "," +
                    gg_host::ratio2(mt[q], qk[q]) + "\n";
    } else {
        for (uint32_t q = 0; q < nq; q++) {
            const uint32_t* row = cc.data() + (uint64_t)q * n_colors;
            bool any = false;
            std::string line = "{\"query_index\":" + std::to_string(q) + ", \"matches\":{";
            for (uint32_t c = 0; c < n_colors; c++) {
                if (!row[c]) continue;
                if (any) line += ",";
                any = true;
                line += "\"" + (color_output_format ? gg_host::json_escaped_name(cmap.names[c]) : std::to_string(c)) +
                        "\": " + gg_host::ratio2(row[c], qk[q]);
            }
            if (any) text += line + "}}\n";
        }
    }
    if (!gg_host::write_output(out_file, text, &err)) return fail(ctx, GGCAT_GPU_EINVAL, err);
    if (out_path) {
        if (out_path_cap < out_file.size() + 1) return fail(ctx, GGCAT_GPU_EINVAL, "output path buffer too small");
        memcpy(out_path, out_file.c_str(), out_file.size() + 1);
    }
    return GGCAT_GPU_OK;
}
