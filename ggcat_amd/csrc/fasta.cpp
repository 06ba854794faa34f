// fasta.cpp -- host-side input reading into the joined layout.
//
// Restates crates/io/src/sequences_reader.rs:56-241 (process_fasta /
// process_fastq): '>' starts a FASTA record, ';' lines are comments, sequence
// lines are concatenated; FASTQ records are 4 lines.  Records longer than
// max(4 MiB, 2(k-1)) are flushed in chunks that repeat the last k-1 bases
// (:119-173), each chunk a record of its own.  Compressed inputs (gzip, lz4,
// bz2, xz, zstd by the last extension) are decoded by decompress.cpp; plain
// files pass through unchanged.  Normalisation
// (:26-54) is not applied here: the device reads a/c/g/t as A/C/G/T and every
// other byte ends an ACGT run, exactly like the 'N' it would become.
//
// Output: records, each followed by '\n'; with colors, the byte offset where
// each file's records start and the file's color (its index in the list).
#include "decompress.cpp"

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace gg_host {

struct ReadResult {
    std::vector<uint8_t> joined;
    std::vector<uint64_t> color_starts;
    std::vector<uint32_t> colors;
    uint64_t estimated_bases = 0;  // crates/io/src/sequences_stream/fasta.rs:12-30
    uint64_t records = 0;
    std::string error;
};

inline void emit(ReadResult& R, const std::vector<uint8_t>& seq) {
    R.joined.insert(R.joined.end(), seq.begin(), seq.end());
    R.joined.push_back('\n');
    R.records++;
}

// File type from the extension, stripping unknown extensions one at a time
// (sequences_reader.rs:64-88).  0 = FASTA, 1 = FASTQ, -1 = unknown.
inline int file_type(const std::string& path) {
    std::string name = path.substr(path.find_last_of('/') == std::string::npos ? 0 : path.find_last_of('/') + 1);
    while (true) {
        size_t dot = name.find_last_of('.');
        if (dot == std::string::npos || dot == 0) return -1;  // Path::extension() is None
        std::string ext = name.substr(dot + 1);
        if (ext == "fq" || ext == "fastq") return 1;
        if (ext == "fa" || ext == "fasta" || ext == "fna" || ext == "ffn") return 0;
        name = name.substr(0, dot);
    }
}

inline std::string last_ext(const std::string& path) {
    size_t slash = path.find_last_of('/');
    size_t dot = path.find_last_of('.');
    if (dot == std::string::npos || (slash != std::string::npos && dot < slash) || dot == slash + 1) return "";
    return path.substr(dot + 1);
}

// The reference reads files in blocks of DEFAULT_OUTPUT_BUFFER_SIZE bytes
// (crates/config/src/lib.rs:53, 4 MiB); block = 0 selects that size, others exist
// so tests can place block ends in small files.
constexpr size_t READ_BLOCK = 4u << 20;

// Reads the file in blocks (lines_reader.rs:33-56) and calls cb(line, partial, finished) exactly as
// LinesReader::process_lines does (lines_reader.rs:147-219): one call per
// complete line ('\n' and one '\r' before it removed), one call with
// partial = true for a non-empty line piece cut by the block end (a trailing
// '\r' removed), and a final call (empty, false, true).
template <class F>
bool for_each_line_piece(ByteSource& f, F&& cb, size_t block) {
    const size_t BLOCK = block ? block : READ_BLOCK;
    std::vector<uint8_t> buf(BLOCK);
    while (true) {
        const long r = f.read(buf.data(), BLOCK);
        if (r < 0) return false;
        if (r == 0) {
            cb((const uint8_t*)nullptr, (size_t)0, false, true);
            return true;
        }
        const uint8_t* s = buf.data();
        size_t n = (size_t)r;
        while (true) {
            const uint8_t* nl = (const uint8_t*)memchr(s, '\n', n);
            if (nl) {
                size_t len = nl - s;
                size_t out = (len && s[len - 1] == '\r') ? len - 1 : len;
                cb(s, out, false, false);
                s += len + 1;
                n -= len + 1;
            } else {
                size_t out = (n && s[n - 1] == '\r') ? n - 1 : n;
                if (out > 0) cb(s, out, true, false);
                break;
            }
        }
    }
}

inline bool read_file(const std::string& path, uint32_t k, ReadResult& R, size_t block = 0) {
    const int type = file_type(path);
    if (type < 0) {
        R.error = "cannot recognize file type of '" + path + "'";
        return false;
    }
    const std::string ext = last_ext(path);
    std::unique_ptr<ByteSource> src = open_source(path, ext, &R.error);
    if (!src) return false;
    ByteSource& f = *src;
    {
        // FastaFileSequencesStream::get_estimated_bases_count: the file size,
        // halved when the last extension names a compression
        FILE* p = fopen(path.c_str(), "rb");
        if (p) {
            fseek(p, 0, SEEK_END);
            uint64_t sz = (uint64_t)ftell(p);
            fclose(p);
            const bool comp = ext == "gz" || ext == "lz4" || ext == "bz2" || ext == "xz" || ext == "zst" || ext == "zstd";
            R.estimated_bases += comp ? sz / 2 : sz;
        }
    }
    std::vector<uint8_t> seq;
    bool ok;
    if (type == 0) {
        // process_fasta (sequences_reader.rs:106-176), minimizer bucketing
        // passes line_split_copyback = k - 1
        // (crates/assembler_minimizer_bucketing/src/lib.rs:333).  As in the
        // reference, a ';' line sets on_comment, which then stays set across
        // every complete line until a piece cut by a block end arrives; the
        // final call is swallowed too while it is set.
        const size_t copyback = k > 1 ? k - 1 : 0;
        const size_t flush_size = std::max<size_t>(block ? block : READ_BLOCK, 2 * copyback);
        bool on_comment = false, ident = false, new_line = true;
        ok = for_each_line_piece(f, [&](const uint8_t* line, size_t len, bool partial, bool finished) {
            if (on_comment) {
                on_comment = !partial;
            } else if (finished || (new_line && len > 0 && line[0] == '>')) {
                if (!seq.empty()) emit(R, seq);
                seq.clear();
                ident = partial;
            } else if (new_line && len > 0 && line[0] == ';') {
                on_comment = true;
            } else if (ident) {
                if (!partial) ident = false;
            } else {
                seq.insert(seq.end(), line, line + len);
            }
            if (seq.size() >= flush_size) {
                emit(R, seq);
                seq.erase(seq.begin(), seq.end() - copyback);
            }
            new_line = !partial;
        }, block);
    } else {
        // process_fastq (sequences_reader.rs:178-240): identifier, sequence,
        // '+' line, quality; the sequence is emitted when its quality line
        // ends, and a pending sequence at the end of the file
        int state = 0;  // 0 ident, 1 seq, 2 qual
        bool skipped_plus = false;
        ok = for_each_line_piece(f, [&](const uint8_t* line, size_t len, bool partial, bool finished) {
            if (finished) {
                if (!seq.empty()) emit(R, seq);
                return;
            }
            if (state == 2) {
                if (!skipped_plus) {
                    if (!partial) skipped_plus = true;
                    return;
                }
                if (!partial) {
                    emit(R, seq);
                    seq.clear();
                    skipped_plus = false;
                }
            } else if (state == 1) {
                seq.insert(seq.end(), line, line + len);
            }
            if (!partial) state = (state + 1) % 3;
        }, block);
    }
    src.reset();
    if (!ok) {
        R.error = "error while reading '" + path + "'";
        return false;
    }
    return true;
}

// Files are independent: they are read by up to 16 threads, each into its own
// result, and concatenated in the order given (the same bytes, offsets and
// colors as reading them one after another).  The error reported is the
// first failing file's in that order.
inline bool read_files(const std::vector<std::string>& paths, uint32_t k, bool colors, ReadResult& R,
                       size_t block = 0) {
    const size_t n = paths.size();
    std::vector<ReadResult> part(n);
    std::vector<char> ok(n, 0);
    std::atomic<size_t> next{0};
    auto work = [&] {
        for (size_t i; (i = next.fetch_add(1)) < n;) ok[i] = read_file(paths[i], k, part[i], block) ? 1 : 0;
    };
    const size_t nt = std::min<size_t>(n, 16);
    std::vector<std::thread> th;
    for (size_t t = 1; t < nt; t++) th.emplace_back(work);
    work();
    for (auto& t : th) t.join();
    size_t total = R.joined.size();
    for (size_t i = 0; i < n; i++) {
        if (!ok[i]) {
            R.error = part[i].error;
            return false;
        }
        total += part[i].joined.size();
    }
    R.joined.reserve(total);
    for (size_t i = 0; i < n; i++) {
        if (colors) {
            R.color_starts.push_back(R.joined.size());
            R.colors.push_back((uint32_t)i);
        }
        R.joined.insert(R.joined.end(), part[i].joined.begin(), part[i].joined.end());
        R.estimated_bases += part[i].estimated_bases;
        R.records += part[i].records;
        std::vector<uint8_t>().swap(part[i].joined);
    }
    return true;
}

// compute_best_m (crates/utils/src/lib.rs:29-40)
inline uint32_t compute_best_m(uint32_t k) {
    if (k <= 13) return std::max(k / 2, k - 4);
    if (k <= 15) return 9;
    if (k <= 21) return 10;
    if (k <= 30) return 11;
    if (k <= 37) return 12;
    if (k <= 42) return 13;
    if (k <= 64) return 14;
    return (uint32_t)((double)k / 4.0 + 0.5);
}

inline uint64_t npow2(uint64_t x) {
    uint64_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

// compute_stats_from_input_blocks (crates/io/src/lib.rs:67-140): bucket and
// second-bucket logs from the estimated base count.
inline void auto_bucket_logs(uint64_t bases, uint32_t* b_log, uint32_t* sb_log) {
    const uint64_t MIN_BUCKET_SIZE = 512 * 1024, MAX_BUCKET_SIZE = 1024ull * 1024 * 1024;
    uint64_t b = std::min<uint64_t>(1 << 10, bases / MIN_BUCKET_SIZE);
    b = std::max<uint64_t>(b, bases / MAX_BUCKET_SIZE);
    b = npow2(b);
    b = std::min<uint64_t>(b, 1 << 13);
    b = std::max<uint64_t>(b, 1 << 2);
    uint64_t per = bases / b;
    uint64_t sb = std::min<uint64_t>(1 << 6, per / (2 * 1024));
    sb = std::max<uint64_t>(sb, per / (4 * 1024 * 1024));
    sb = npow2(sb);
    sb = std::min<uint64_t>(sb, 1 << 8);
    sb = std::max<uint64_t>(sb, 1 << 1);
    uint32_t bl = 0, sl = 0;
    while ((1ull << bl) < b) bl++;
    while ((1ull << sl) < sb) sl++;
    *b_log = bl;
    *sb_log = sl;
}

}  // namespace gg_host
