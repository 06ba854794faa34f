// decompress.cpp -- byte sources for the input reader: plain and gzip files
// through zlib, and lz4 / bz2 / xz / zstd files through the system's runtime
// libraries (liblz4.so.1, libbz2.so.1, liblzma.so.5, libzstd.so.1), opened
// with dlopen when such a file is first read, so the library has no link-time
// dependency on them.  The choice follows the last extension, as
// crates/io/src/lines_reader.rs:58-130 does (gz, lz4, bz2, xz, zst / zstd,
// anything else plain).  Each decoder continues over concatenated streams /
// frames.  Only the documented public entry points of each library are used;
// their prototypes and the two stream structs are declared here because the
// image has the libraries but not their headers.
//
// Included by fasta.cpp.
#include <dlfcn.h>
#include <zlib.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace gg_host {

// read(): bytes produced (at most n), 0 at the end, -1 on an error.
struct ByteSource {
    virtual ~ByteSource() = default;
    virtual long read(uint8_t* buf, size_t n) = 0;
};

struct GzSource : ByteSource {
    gzFile f = nullptr;
    ~GzSource() override {
        if (f) gzclose(f);
    }
    long read(uint8_t* buf, size_t n) override {
        int r = gzread(f, buf, (unsigned)std::min<size_t>(n, 1u << 30));
        return r < 0 ? -1 : r;
    }
};

// A compressed file decoded through a pull-style stream API: the source keeps
// a 1 MiB input buffer refilled from the file.
struct FileInput {
    FILE* f = nullptr;
    std::vector<uint8_t> in = std::vector<uint8_t>(1u << 20);
    size_t pos = 0, len = 0;
    bool eof = false, err = false;
    ~FileInput() {
        if (f) fclose(f);
    }
    // makes input available; false when the file is exhausted
    bool fill() {
        if (pos < len) return true;
        if (eof) return false;
        len = fread(in.data(), 1, in.size(), f);
        pos = 0;
        if (len == 0) {
            eof = true;
            err = ferror(f) != 0;
            return false;
        }
        return true;
    }
};

void* dl_open(const char* name) {
    static std::mutex mu;
    std::lock_guard<std::mutex> lock(mu);
    return dlopen(name, RTLD_NOW | RTLD_LOCAL);
}

template <class F>
bool dl_sym(void* h, const char* name, F& out) {
    out = reinterpret_cast<F>(dlsym(h, name));
    return out != nullptr;
}

// ---- zstd (libzstd streaming decompression API) ---------------------------
struct ZstdIn {
    const void* src;
    size_t size, pos;
};
struct ZstdOut {
    void* dst;
    size_t size, pos;
};
struct ZstdApi {
    void* (*create)() = nullptr;
    size_t (*init)(void*) = nullptr;
    size_t (*decompress)(void*, ZstdOut*, ZstdIn*) = nullptr;
    size_t (*free_)(void*) = nullptr;
    unsigned (*is_error)(size_t) = nullptr;
    bool ok = false;
};
const ZstdApi& zstd_api() {
    static ZstdApi a = [] {
        ZstdApi x;
        void* h = dl_open("libzstd.so.1");
        x.ok = h && dl_sym(h, "ZSTD_createDStream", x.create) && dl_sym(h, "ZSTD_initDStream", x.init) &&
               dl_sym(h, "ZSTD_decompressStream", x.decompress) && dl_sym(h, "ZSTD_freeDStream", x.free_) &&
               dl_sym(h, "ZSTD_isError", x.is_error);
        return x;
    }();
    return a;
}
struct ZstdSource : ByteSource {
    FileInput in;
    void* ds = nullptr;
    bool frame_done = true;
    ~ZstdSource() override {
        if (ds) zstd_api().free_(ds);
    }
    long read(uint8_t* buf, size_t n) override {
        const ZstdApi& z = zstd_api();
        ZstdOut out{buf, n, 0};
        while (out.pos == 0) {
            const bool more = in.fill();
            if (!more && in.err) return -1;
            if (!more && frame_done) return 0;
            // without input the call still flushes output the decoder holds
            ZstdIn src{more ? in.in.data() + in.pos : nullptr, more ? in.len - in.pos : 0, 0};
            const size_t r = z.decompress(ds, &out, &src);
            if (more) in.pos += src.pos;
            if (z.is_error(r)) return -1;
            frame_done = r == 0;  // a frame is decoded and flushed
            if (!more && out.pos == 0) return frame_done ? 0 : -1;  // else the last frame is truncated
        }
        return (long)out.pos;
    }
};

// ---- xz (liblzma stream decoder) -----------------------------------------
struct LzmaStream {
    const uint8_t* next_in;
    size_t avail_in;
    uint64_t total_in;
    uint8_t* next_out;
    size_t avail_out;
    uint64_t total_out;
    const void* allocator;
    void* internal;
    void *reserved_ptr1, *reserved_ptr2, *reserved_ptr3, *reserved_ptr4;
    uint64_t reserved_int1, reserved_int2;
    size_t reserved_int3, reserved_int4;
    int reserved_enum1, reserved_enum2;
};
struct LzmaApi {
    int (*stream_decoder)(LzmaStream*, uint64_t, uint32_t) = nullptr;
    int (*code)(LzmaStream*, int) = nullptr;
    void (*end)(LzmaStream*) = nullptr;
    bool ok = false;
};
const LzmaApi& lzma_api() {
    static LzmaApi a = [] {
        LzmaApi x;
        void* h = dl_open("liblzma.so.5");
        x.ok = h && dl_sym(h, "lzma_stream_decoder", x.stream_decoder) && dl_sym(h, "lzma_code", x.code) &&
               dl_sym(h, "lzma_end", x.end);
        return x;
    }();
    return a;
}
constexpr uint32_t LZMA_CONCATENATED_FLAG = 0x08;
struct XzSource : ByteSource {
    FileInput in;
    LzmaStream s{};
    bool init = false, done = false;
    ~XzSource() override {
        if (init) lzma_api().end(&s);
    }
    long read(uint8_t* buf, size_t n) override {
        const LzmaApi& z = lzma_api();
        s.next_out = buf;
        s.avail_out = n;
        while (s.avail_out == n && !done) {
            const bool more = in.fill();
            if (!more && in.err) return -1;
            s.next_in = more ? in.in.data() + in.pos : nullptr;
            s.avail_in = more ? in.len - in.pos : 0;
            const int r = z.code(&s, more ? 0 /* LZMA_RUN */ : 3 /* LZMA_FINISH */);
            if (more) in.pos = in.len - s.avail_in;
            if (r == 1 /* LZMA_STREAM_END */) done = true;
            else if (r != 0 /* LZMA_OK */) return -1;
            if (!more && s.avail_out == n && !done) return -1;  // truncated stream
        }
        return (long)(n - s.avail_out);
    }
};

// ---- bz2 (libbz2 decompression) ------------------------------------------
struct BzStream {
    char* next_in;
    unsigned int avail_in, total_in_lo32, total_in_hi32;
    char* next_out;
    unsigned int avail_out, total_out_lo32, total_out_hi32;
    void* state;
    void* (*bzalloc)(void*, int, int);
    void (*bzfree)(void*, void*);
    void* opaque;
};
struct BzApi {
    int (*init)(BzStream*, int, int) = nullptr;
    int (*decompress)(BzStream*) = nullptr;
    int (*end)(BzStream*) = nullptr;
    bool ok = false;
};
const BzApi& bz_api() {
    static BzApi a = [] {
        BzApi x;
        void* h = dl_open("libbz2.so.1");
        if (!h) h = dl_open("libbz2.so.1.0");
        x.ok = h && dl_sym(h, "BZ2_bzDecompressInit", x.init) && dl_sym(h, "BZ2_bzDecompress", x.decompress) &&
               dl_sym(h, "BZ2_bzDecompressEnd", x.end);
        return x;
    }();
    return a;
}
struct Bz2Source : ByteSource {
    FileInput in;
    BzStream s{};
    bool live = false;
    ~Bz2Source() override {
        if (live) bz_api().end(&s);
    }
    long read(uint8_t* buf, size_t n) override {
        const BzApi& z = bz_api();
        const unsigned cap = (unsigned)std::min<size_t>(n, 1u << 30);
        unsigned got = 0;
        while (got == 0) {
            const bool more = in.fill();
            if (!more && in.err) return -1;
            if (!live) {
                if (!more) return 0;  // no further stream
                if (z.init(&s, 0, 0) != 0) return -1;
                live = true;
            }
            s.next_in = more ? reinterpret_cast<char*>(in.in.data() + in.pos) : nullptr;
            s.avail_in = more ? (unsigned)(in.len - in.pos) : 0u;
            s.next_out = reinterpret_cast<char*>(buf);
            s.avail_out = cap;
            const int r = z.decompress(&s);
            if (more) in.pos = in.len - s.avail_in;
            got = cap - s.avail_out;
            if (r == 4 /* BZ_STREAM_END */) {
                z.end(&s);
                live = false;
                s = BzStream{};
            } else if (r != 0 /* BZ_OK */) {
                return -1;
            } else if (!more && got == 0) {
                return -1;  // truncated stream
            }
        }
        return (long)got;
    }
};

// ---- lz4 (liblz4 frame API) ----------------------------------------------
struct Lz4Api {
    size_t (*create)(void**, unsigned) = nullptr;
    size_t (*free_)(void*) = nullptr;
    size_t (*decompress)(void*, void*, size_t*, const void*, size_t*, const void*) = nullptr;
    unsigned (*is_error)(size_t) = nullptr;
    bool ok = false;
};
const Lz4Api& lz4_api() {
    static Lz4Api a = [] {
        Lz4Api x;
        void* h = dl_open("liblz4.so.1");
        x.ok = h && dl_sym(h, "LZ4F_createDecompressionContext", x.create) &&
               dl_sym(h, "LZ4F_freeDecompressionContext", x.free_) && dl_sym(h, "LZ4F_decompress", x.decompress) &&
               dl_sym(h, "LZ4F_isError", x.is_error);
        return x;
    }();
    return a;
}
struct Lz4Source : ByteSource {
    FileInput in;
    void* dctx = nullptr;
    bool frame_done = true;
    ~Lz4Source() override {
        if (dctx) lz4_api().free_(dctx);
    }
    long read(uint8_t* buf, size_t n) override {
        const Lz4Api& z = lz4_api();
        size_t produced = 0;
        while (produced == 0) {
            const bool more = in.fill();
            if (!more && in.err) return -1;
            if (!more && frame_done) return 0;
            size_t dst = n, src = more ? in.len - in.pos : 0;
            const size_t r = z.decompress(dctx, buf, &dst, more ? in.in.data() + in.pos : nullptr, &src, nullptr);
            if (z.is_error(r)) return -1;
            if (more) in.pos += src;
            produced = dst;
            frame_done = r == 0;  // a frame is decoded and flushed
            if (!more && produced == 0) return frame_done ? 0 : -1;  // else the last frame is truncated
        }
        return (long)produced;
    }
};

// Opens `path` with the decoder its last extension names; nullptr with
// *error set on failure.
inline std::unique_ptr<ByteSource> open_source(const std::string& path, const std::string& ext, std::string* error) {
    auto open_file = [&](FileInput& in) {
        in.f = fopen(path.c_str(), "rb");
        if (!in.f) *error = "cannot open '" + path + "'";
        return in.f != nullptr;
    };
    auto missing = [&](const char* lib) {
        *error = std::string("reading '") + path + "' needs " + lib + ", which could not be loaded";
        return nullptr;
    };
    if (ext == "zst" || ext == "zstd") {
        if (!zstd_api().ok) return missing("libzstd.so.1");
        std::unique_ptr<ZstdSource> s(new ZstdSource());
        if (!open_file(s->in)) return nullptr;
        s->ds = zstd_api().create();
        if (!s->ds || zstd_api().is_error(zstd_api().init(s->ds))) {
            *error = "zstd decoder initialisation failed";
            return nullptr;
        }
        return s;
    }
    if (ext == "xz") {
        if (!lzma_api().ok) return missing("liblzma.so.5");
        std::unique_ptr<XzSource> s(new XzSource());
        if (!open_file(s->in)) return nullptr;
        if (lzma_api().stream_decoder(&s->s, UINT64_MAX, LZMA_CONCATENATED_FLAG) != 0) {
            *error = "xz decoder initialisation failed";
            return nullptr;
        }
        s->init = true;
        return s;
    }
    if (ext == "bz2") {
        if (!bz_api().ok) return missing("libbz2.so.1");
        std::unique_ptr<Bz2Source> s(new Bz2Source());
        if (!open_file(s->in)) return nullptr;
        return s;
    }
    if (ext == "lz4") {
        if (!lz4_api().ok) return missing("liblz4.so.1");
        std::unique_ptr<Lz4Source> s(new Lz4Source());
        if (!open_file(s->in)) return nullptr;
        if (lz4_api().is_error(lz4_api().create(&s->dctx, 100 /* LZ4F_VERSION */))) {
            *error = "lz4 decoder initialisation failed";
            return nullptr;
        }
        return s;
    }
    // gzip and plain files: zlib reads plain files unchanged
    std::unique_ptr<GzSource> s(new GzSource());
    s->f = gzopen(path.c_str(), "rb");
    if (!s->f) {
        *error = "cannot open '" + path + "'";
        return nullptr;
    }
    gzbuffer(s->f, 1 << 20);
    return s;
}

}  // namespace gg_host
