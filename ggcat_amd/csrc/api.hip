// api.hip -- C ABI (include/ggcat_gpu.h) and host orchestration of the
// bucketing and merge kernels.  Unity build: the kernel files are included.
#include "../../include/ggcat_gpu.h"
#include "bucketing.hip"
#include "merge.hip"
#include "query.hip"
#include "unitigs.hip"
#include "parse.hip"
#include "resplit.hip"
#include "fasta.cpp"
#include "colormap.cpp"

#include <algorithm>
#include <cstdio>
#include <array>
#include <deque>
#include <map>
#include <memory>
#include <chrono>
#include <mutex>
#include <thread>
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>
#include <cstring>
#include <string>
#include <tuple>
#include <vector>

using namespace gg;

__global__ void fill32_kernel(uint32_t* p, uint32_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}


namespace {

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
};

constexpr uint64_t PIN_CHUNK = 64ull << 20;  // bytes per pinned staging buffer (two of them)
constexpr uint32_t HOST_COPY_THREADS = 16;  // host threads reading or copying input into a staging buffer

}  // namespace

struct ggcat_gpu_ctx {
    int device = 0;
    int n_cu = 256;
    hipStream_t stream = nullptr;
    std::string err;
    hipEvent_t ev[16];
    // bucketing buffers
    DevBuf input, seg_bits, unit_count, unit_cursor, unit_off, unit_kmers, scan_part, slots, pool, pool_cursor,
        slot_color, color_starts, colors, error_flag, stage_slots, stage_unit, stage_rank, stage_color, stage_cursor, defer,
        defer_count, kmer_total, spill_slots, spill_color,  // stage_*: the unit regions, the spill list's units and ranks
        unit_dcount, unit_dregion, unit_skmers;  // per unit after deduplication
    // merge buffers
    DevBuf items, n_items, work, parts, redo, redo_count, out_key, out_val, out_sub, chunk_cursor, chunk_fill_dev,
        fix_key, fix_val, fix_sub, fix_cursor, out_colors, fix_colors;
    // state
    ggcat_gpu_params P{};
    bool have_records = false, have_solid = false;
    uint32_t n_sub = 0, n_units = 0, xb = 0, slot_words = 0, key_words = 1;
    uint64_t n_records = 0, pool_words = 0, n_solid = 0, n_input = 0, kmer_occ = 0, n_big = 0, n_fix = 0;
    uint64_t n_stored = 0;  // arena slots: records with identical copies stored once
    uint32_t n_sub_all = 0;  // sub-buckets including the resplit buckets of outliers
    uint32_t dd_max = 0;  // the compaction deduplicated the units of up to this many records
    bool resplit_done = false;  // the arena holds resplit buckets (ggcat_gpu_resplit_outliers)
    uint64_t n_spill = 0;  // records of the last bucketing past their unit's region
    uint32_t n_chunks_used = 0;  // output chunks of the last merge
    uint32_t n_passes = 0;  // kernel passes of the last merge
    std::vector<uint32_t> chunk_fill;  // host copy of their fill counts
    uint64_t n_tail = 0;  // entries appended after the chunks (combined fixups)
    bool colored = false;
    uint32_t color_words = 0;  // bitmap words per solid entry of a colored build
    double ms[9] = {0};
    uint64_t input_path = 0, flush_cuts = 0;  // last build_from_files: 1 parsed in HBM, 2 host reader
    ggcat_gpu_tuning tune{};  // zero fields select the defaults
    hipGraphExec_t graph_a = nullptr;  // captured bucketing staging pass (graphs enabled)
    struct GraphKeyA {
        gg::BucketArgs args;  // every kernel argument of the captured launches
        uint64_t seg_words;
        uint32_t n_units;
    } graph_a_key{};
    uint32_t n_color_ids = 0;  // largest color id + 1 of the colored build
    // query index over the solid table
    DevBuf idx_key, idx_pos, idx_boff, idx_bcount, idx_bcursor;
    bool have_index = false;
    uint32_t idx_log_buckets = 0;
    DevBuf q_seq, q_off, q_matched, q_counts, q_work;
    struct Server* server = nullptr;
    // unitig chains
    DevBuf u_nb, u_ext, u_link, u_hbytes, u_hcount, u_hoff, u_hidx, u_seq, u_sub, u_flags, u_visited, u_rest,
        u_rest_n, u_rep, u_tab;
    bool have_unitigs = false;
    bool u_joined = false;
    uint64_t u_n = 0, u_bytes = 0, u_npos = 0;
    std::vector<uint8_t> cyc_seq, cyc_flags;
    std::vector<uint32_t> cyc_sub;
    DevBuf u_path;  // colored builds: table position of every unitig k-mer, in written order
    DevBuf rk_pos;  // rabin-karp128 per-position start terms of the merge's rolling hash
    uint32_t rk_pos_k = 0;  // k they were computed for (0: none)
    // input files read straight into HBM: raw bytes, parse state per chunk,
    // and a pinned two-buffer staging ring for the uploads
    DevBuf raw, p_stats, p_carry, p_keep, p_keep_off, p_flags, p_bits;
    void* pin_ring = nullptr;
    hipStream_t copy_stream = nullptr;  // host-to-device uploads of host input, beside the compute stream
    hipEvent_t up_ev[2] = {nullptr, nullptr};  // last upload out of each staging buffer
    hipEvent_t dd_ev[2] = {nullptr, nullptr};  // the wide deduplication's start and end on the copy stream
    std::vector<uint32_t> cyc_pos;  // the same for the unitigs of cycles (walked on the host)
    // color subsets of the solid table: ids for the color map and the colored
    // FASTA headers (assigned on first use after a merge)
    bool have_subsets = false;
    std::vector<uint32_t> pos_subset;  // per table position, ~0u for an empty position
    std::vector<std::vector<uint32_t>> subsets;
};

namespace {

// Error text of calls made without a context (ggcat_gpu_read_files accepts
// ctx == NULL so that inputs can be read on a machine without a GPU).
thread_local std::string g_ctxless_err;

int fail(ggcat_gpu_ctx* c, int code, const std::string& msg) {
    if (c)
        c->err = msg;
    else
        g_ctxless_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(ctx, GGCAT_GPU_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

int ensure(ggcat_gpu_ctx* ctx, DevBuf& b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b.cap >= bytes) return GGCAT_GPU_OK;
    if (b.p) {
        hipError_t e = hipFree(b.p);
        if (e != hipSuccess) return fail(ctx, GGCAT_GPU_EHIP, std::string("hipFree: ") + hipGetErrorString(e));
        b.p = nullptr;
        b.cap = 0;
    }
    size_t want = bytes + bytes / 8;
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(ctx, GGCAT_GPU_ENOMEM, "hipMalloc(" + std::to_string(want) + "): " + hipGetErrorString(e));
    }
    b.cap = want;
    return GGCAT_GPU_OK;
}

#define ENSURE(buf, bytes)                        \
    do {                                          \
        int r_ = ensure(ctx, (buf), (bytes));     \
        if (r_ != GGCAT_GPU_OK) return r_;        \
    } while (0)

template <class T>
T* P_(DevBuf& b) {
    return reinterpret_cast<T*>(b.p);
}

// ntHash rolling tables, index out*4+in with out == 4 meaning "no base leaves"
// (crates/hashes/src/nthash_base.rs:49-61 on 2-bit codes A=0,C=1,T=2,G=3).
void nthash_tables(uint32_t m, uint64_t* gtab) {
    auto rotl = [](uint64_t x, unsigned r) { r &= 63; return r ? (x << r) | (x >> (64 - r)) : x; };
    auto rotr = [](uint64_t x, unsigned r) { r &= 63; return r ? (x >> r) | (x << (64 - r)) : x; };
    auto h = [](uint32_t c) { return (uint64_t)((c << 1) + 1) * NT_MULT; };
    auto rcc = [](uint32_t c) { return (uint64_t)(((c << 1) ^ 4) + 1) * NT_MULT; };
    for (uint32_t out = 0; out < 5; out++) {
        for (uint32_t in = 0; in < 4; in++) {
            uint32_t e = out * 4 + in;
            uint64_t f = h(in), r = rotl(rcc(in), m - 1);
            if (out < 4) {
                f ^= rotl(h(out), m);
                r ^= rotr(rcc(out), 1);
            }
            gtab[2 * e] = f;
            gtab[2 * e + 1] = r;
        }
    }
}

u128_t mk128(uint64_t hi, uint64_t lo) { return ((u128_t)hi << 64) | lo; }

const uint32_t TILE_T = 1024;
// Colored merges keep one bit per color in each table entry.
const uint32_t MAX_COLORS = 1024;

// Device-wide exclusive scan counts[0..n) -> off[0..n].
int device_scan(ggcat_gpu_ctx* ctx, const uint32_t* counts, uint64_t* off, uint32_t n) {
    const uint32_t nb = std::max<uint32_t>(1, (n + SCAN_CHUNK - 1) / SCAN_CHUNK);
    ENSURE(ctx->scan_part, ((size_t)nb + 1) * 8);
    uint64_t* part = P_<uint64_t>(ctx->scan_part);
    hipLaunchKernelGGL(scan_reduce_kernel, dim3(nb), dim3(1024), 0, ctx->stream, counts, part, n);
    hipLaunchKernelGGL(scan_partials_kernel, dim3(1), dim3(1024), 0, ctx->stream, part, nb);
    hipLaunchKernelGGL(scan_apply_kernel, dim3(nb), dim3(1024), 0, ctx->stream, counts, part, off, n, nb);
    HIPCHK(hipGetLastError());
    return GGCAT_GPU_OK;
}

// rabin-karp128 constants (crates/hashes/src/cn_rkhash.rs:65-78) and
// rmmult = MULT^(k-1) (crates/hashes/src/lib.rs:167-191).
template <class Args>
void set_rk_constants(Args& M, uint32_t k) {
    M.rk_mult = mk128(0x3eb9402f3e733993ULL, 0xadd64d3ca00e1b6bULL);
    M.rk_mult_inv = mk128(0x09cb6ff6f1b1a6d7ULL, 0x33e0952e899c3943ULL);
    const u128_t A_ = mk128(0x4751137d01d863c5ULL, 0xb8c36de2b7d399dfULL);
    const u128_t C_ = mk128(0x37ea3a13226503fbULL, 0x783f5cb69f4552bdULL);
    const u128_t G_ = mk128(0x50796b285343f09aULL, 0x0c53113ae736572bULL);
    const u128_t T_ = mk128(0x1e62d96a5e1f5adeULL, 0x2d4e68d8f88110b7ULL);
    M.rk_fwd[0] = A_; M.rk_fwd[1] = C_; M.rk_fwd[2] = T_; M.rk_fwd[3] = G_;
    M.rk_bkw[0] = T_; M.rk_bkw[1] = G_; M.rk_bkw[2] = A_; M.rk_bkw[3] = C_;
    u128_t r = 1, b = M.rk_mult;
    for (uint32_t e = k - 1; e; e >>= 1, b *= b)
        if (e & 1) r *= b;
    M.rk_rmmult = r;
    for (uint32_t c = 0; c < 4; c++) {
        M.rk_fwd_rm[c] = M.rk_fwd[c] * r;
        M.rk_bkw_rm[c] = M.rk_bkw[c] * r;
    }
}

// Query records of a joined buffer: start offsets [n + 1] (entry q + 1 is one
// past query q's separator).  False when the separator count differs from n
// or the buffer does not end with a separator.
bool query_offsets(const uint8_t* q, uint64_t n_bytes, uint32_t n, std::vector<uint64_t>& off) {
    off.clear();
    off.reserve((size_t)n + 1);
    off.push_back(0);
    const uint8_t* p = q;
    const uint8_t* end = q + n_bytes;
    while (p < end) {
        const void* f = memchr(p, '\n', (size_t)(end - p));
        if (!f) break;
        p = static_cast<const uint8_t*>(f) + 1;
        off.push_back((uint64_t)(p - q));
    }
    return off.size() == (size_t)n + 1 && (n_bytes == 0 || q[n_bytes - 1] == '\n');
}

void query_kmer_counts(const std::vector<uint64_t>& off, uint32_t k, uint64_t* out) {
    for (size_t q = 0; q + 1 < off.size(); q++) {
        const uint64_t len = off[q + 1] - off[q] - 1;
        out[q] = len >= k ? len - (k - 1) : 0;
    }
}

QueryArgs query_args(ggcat_gpu_ctx* ctx) {
    QueryArgs Q{};
    Q.k = ctx->P.k;
    Q.kind = ctx->P.hash_kind;
    Q.key_words = ctx->key_words;
    Q.idx_key = P_<uint64_t>(ctx->idx_key);
    Q.idx_pos = P_<uint32_t>(ctx->idx_pos);
    Q.boff = P_<uint64_t>(ctx->idx_boff);
    Q.log_buckets = ctx->idx_log_buckets;
    Q.color_words = ctx->colored ? ctx->color_words : 0;
    Q.out_colors = ctx->colored ? P_<uint32_t>(ctx->out_colors) : nullptr;
    Q.n_colors = ctx->colored ? ctx->n_color_ids : 0;
    set_rk_constants(Q, ctx->P.k);
    return Q;
}

// First-fit allocator over one pre-allocated range (offsets 256-aligned).
struct RangeAlloc {
    std::map<uint64_t, uint64_t> free_;  // offset -> size
    void init(uint64_t size) {
        free_.clear();
        free_[0] = size;
    }
    bool alloc(uint64_t want, uint64_t* off) {
        want = (want + 255) & ~255ull;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < want) continue;
            *off = it->first;
            const uint64_t rest = it->second - want, at = it->first + want;
            free_.erase(it);
            if (rest) free_[at] = rest;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t size) {
        size = (size + 255) & ~255ull;
        auto it = free_.emplace(off, size).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct ServeReq {
    std::vector<uint8_t> data;  // own copy of the query bytes while the request waits
    const uint8_t* src = nullptr;  // the query bytes (the caller's buffer during submit)
    uint64_t n_bytes = 0;
    std::vector<uint64_t> off;
    uint32_t n_queries = 0;
    bool admitted = false, done = false;
    int stream = -1;
    uint64_t dev_off = 0, pin_off = 0, bytes = 0;
    // layout inside both ranges
    uint64_t o_seq = 0, o_off = 0, o_work = 0, o_matched = 0, o_counts = 0;
    hipEvent_t ev = nullptr;
};

}  // namespace

struct Server {
    std::vector<hipStream_t> streams;
    std::vector<uint32_t> load;  // admitted requests per stream
    uint8_t* dev = nullptr;
    uint8_t* pin = nullptr;
    uint64_t budget = 0;
    RangeAlloc dalloc, palloc;
    std::deque<uint64_t> waiting;
    std::map<uint64_t, ServeReq> reqs;
    std::vector<hipEvent_t> free_events;
    uint64_t next_ticket = 1;
    uint32_t n_admitted = 0;
};

namespace {

// Layout of one request inside its device / pinned range (same offsets in
// both): query bytes, offsets, work counter, then the results.
void serve_layout(const ggcat_gpu_ctx* ctx, ServeReq& r) {
    const uint32_t nq = r.n_queries, nc = ctx->colored ? ctx->n_color_ids : 0;
    uint64_t o = 0;
    auto place = [&](uint64_t bytes) {
        const uint64_t at = o;
        o = (o + bytes + 255) & ~255ull;
        return at;
    };
    r.o_seq = place(r.n_bytes + 16);
    r.o_off = place(((uint64_t)nq + 1) * 8);
    r.o_work = place(8);
    r.o_matched = place((uint64_t)nq * 8);
    r.o_counts = place((uint64_t)nq * nc * 4 + 4);
    r.bytes = o;
}

// Admits waiting requests in arrival order while their memory fits.  Every
// waiting request fits the budget on its own (serve_submit rejects the
// others), so the queue always drains once earlier requests are fetched.
int serve_admit(ggcat_gpu_ctx* ctx) {
    Server& S = *ctx->server;
    while (!S.waiting.empty()) {
        ServeReq& r = S.reqs[S.waiting.front()];
        const uint32_t nq = r.n_queries, nc = ctx->colored ? ctx->n_color_ids : 0;
        uint64_t doff, poff;
        if (!S.dalloc.alloc(r.bytes, &doff)) break;
        if (!S.palloc.alloc(r.bytes, &poff)) {
            S.dalloc.release(doff, r.bytes);
            break;
        }
        r.dev_off = doff;
        r.pin_off = poff;
        uint8_t* hp = S.pin + poff;
        memcpy(hp + r.o_seq, r.src, r.n_bytes);
        memcpy(hp + r.o_off, r.off.data(), r.off.size() * 8);
        memset(hp + r.o_work, 0, 8);
        int si = 0;
        for (int i = 1; i < (int)S.streams.size(); i++)
            if (S.load[i] < S.load[si]) si = i;
        r.stream = si;
        hipStream_t st = S.streams[si];
        uint8_t* dp = S.dev + doff;
        HIPCHK(hipMemcpyAsync(dp, hp, r.o_matched, hipMemcpyHostToDevice, st));  // seq, offsets, work counter
        QueryArgs Q = query_args(ctx);
        Q.seq = dp + r.o_seq;
        Q.qoff = reinterpret_cast<const uint64_t*>(dp + r.o_off);
        Q.n_queries = nq;
        Q.matched = reinterpret_cast<uint64_t*>(dp + r.o_matched);
        Q.color_counts = reinterpret_cast<uint32_t*>(dp + r.o_counts);
        Q.work = reinterpret_cast<unsigned int*>(dp + r.o_work);
        if (nq) {
            const uint32_t grid = std::min<uint32_t>(nq, ctx->n_cu * 4);
            hipLaunchKernelGGL(query_kernel, dim3(grid), dim3(256), (size_t)nc * 4 + 16, st, Q);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipMemcpyAsync(hp + r.o_matched, dp + r.o_matched, r.bytes - r.o_matched, hipMemcpyDeviceToHost, st));
        if (S.free_events.empty()) {
            hipEvent_t e;
            HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            S.free_events.push_back(e);
        }
        r.ev = S.free_events.back();
        S.free_events.pop_back();
        HIPCHK(hipEventRecord(r.ev, st));
        r.admitted = true;
        S.load[si]++;
        S.n_admitted++;
        S.waiting.pop_front();
    }
    return GGCAT_GPU_OK;
}

}  // namespace

extern "C" {

namespace {
static int noalloc_scratch_setup(int device);
}  // namespace

int ggcat_gpu_create(int device, ggcat_gpu_ctx** out) {
    if (!out) return GGCAT_GPU_EINVAL;
    *out = nullptr;
    ggcat_gpu_ctx* ctx = new ggcat_gpu_ctx();
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking);
    for (int i = 0; i < 16 && e == hipSuccess; i++) e = hipEventCreate(&ctx->ev[i]);
    for (int i = 0; i < 2 && e == hipSuccess; i++) e = hipEventCreateWithFlags(&ctx->up_ev[i], hipEventDisableTiming);
    for (int i = 0; i < 2 && e == hipSuccess; i++) e = hipEventCreateWithFlags(&ctx->dd_ev[i], hipEventDisableTiming);
    hipDeviceProp_t prop;
    if (e == hipSuccess && hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->n_cu = prop.multiProcessorCount;
    // the no-allocation entry point's scratch array: its address on this device
    if (e == hipSuccess && noalloc_scratch_setup(device) != GGCAT_GPU_OK) e = hipErrorInvalidSymbol;
    if (e != hipSuccess) {
        delete ctx;
        return GGCAT_GPU_EHIP;
    }
    *out = ctx;
    return GGCAT_GPU_OK;
}

void ggcat_gpu_destroy(ggcat_gpu_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->copy_stream) (void)hipStreamSynchronize(ctx->copy_stream);
    DevBuf* bufs[] = {&ctx->input, &ctx->seg_bits, &ctx->unit_count, &ctx->unit_cursor, &ctx->unit_off,
                      &ctx->unit_kmers, &ctx->scan_part, &ctx->slots, &ctx->pool, &ctx->pool_cursor,
                      &ctx->unit_dcount, &ctx->unit_dregion, &ctx->unit_skmers,
                      &ctx->slot_color, &ctx->color_starts, &ctx->colors, &ctx->error_flag, &ctx->items,
                      &ctx->n_items, &ctx->work, &ctx->parts, &ctx->redo, &ctx->redo_count, &ctx->out_key,
                      &ctx->out_val, &ctx->out_sub, &ctx->chunk_cursor, &ctx->chunk_fill_dev, &ctx->fix_key,
                      &ctx->fix_val, &ctx->fix_sub, &ctx->fix_cursor, &ctx->out_colors, &ctx->fix_colors,
                      &ctx->idx_key, &ctx->idx_pos, &ctx->idx_boff, &ctx->idx_bcount, &ctx->idx_bcursor,
                      &ctx->q_seq, &ctx->q_off, &ctx->q_matched, &ctx->q_counts, &ctx->q_work,
                      &ctx->u_nb, &ctx->u_ext, &ctx->u_link, &ctx->u_hbytes, &ctx->u_hcount, &ctx->u_hoff,
                      &ctx->u_hidx, &ctx->u_seq, &ctx->u_sub, &ctx->u_flags, &ctx->u_visited, &ctx->u_rest,
                      &ctx->u_rest_n, &ctx->u_rep, &ctx->u_tab, &ctx->u_path, &ctx->stage_slots, &ctx->stage_unit,
                      &ctx->stage_rank, &ctx->stage_color, &ctx->stage_cursor, &ctx->defer, &ctx->defer_count,
                      &ctx->kmer_total, &ctx->spill_slots, &ctx->spill_color, &ctx->raw, &ctx->p_stats,
                      &ctx->p_carry, &ctx->p_keep, &ctx->p_keep_off, &ctx->p_flags, &ctx->p_bits, &ctx->rk_pos};
    if (ctx->pin_ring) (void)hipHostFree(ctx->pin_ring);
    if (ctx->server) (void)ggcat_gpu_serve_close(ctx);
    if (ctx->graph_a) (void)hipGraphExecDestroy(ctx->graph_a);
    for (DevBuf* b : bufs)
        if (b->p) (void)hipFree(b->p);
    for (int i = 0; i < 16; i++) (void)hipEventDestroy(ctx->ev[i]);
    for (int i = 0; i < 2; i++) {
        if (ctx->up_ev[i]) (void)hipEventDestroy(ctx->up_ev[i]);
        if (ctx->dd_ev[i]) (void)hipEventDestroy(ctx->dd_ev[i]);
    }
    if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* ggcat_gpu_last_error(const ggcat_gpu_ctx* ctx) {
    if (ctx) return ctx->err.c_str();
    return g_ctxless_err.empty() ? "null context" : g_ctxless_err.c_str();
}

int ggcat_gpu_bucketing(ggcat_gpu_ctx* ctx, const uint8_t* data, uint64_t n_bytes, int input_on_device,
                        const uint64_t* color_starts, const uint32_t* colors, uint32_t n_colors,
                        const ggcat_gpu_params* params) {
    if (!ctx) return GGCAT_GPU_EINVAL;
    if (!params || (!data && n_bytes)) return fail(ctx, GGCAT_GPU_EINVAL, "null params or data");
    const ggcat_gpu_params p = *params;
    if (p.k < 2 || p.k > 64) return fail(ctx, GGCAT_GPU_EINVAL, "k must be in [2, 64]");
    if (p.m < 1 || p.m + 2 > p.k) return fail(ctx, GGCAT_GPU_EINVAL, "m must satisfy 1 <= m <= k-2");
    if (p.buckets_log < 1 || p.buckets_log > 13) return fail(ctx, GGCAT_GPU_EINVAL, "buckets_log must be in [1, 13]");
    if (p.second_buckets_log > 8) return fail(ctx, GGCAT_GPU_EINVAL, "second_buckets_log must be <= 8");
    if (p.hash_kind > 1) return fail(ctx, GGCAT_GPU_EINVAL, "unknown hash kind");
    if (n_colors && (!color_starts || !colors)) return fail(ctx, GGCAT_GPU_EINVAL, "null color arrays");
    if (input_on_device && (reinterpret_cast<uintptr_t>(data) & 15))
        return fail(ctx, GGCAT_GPU_EINVAL, "device input must be 16-byte aligned");
    HIPCHK(hipSetDevice(ctx->device));
    if (ctx->server) return fail(ctx, GGCAT_GPU_ESTATE, "close the query server before a new build");
    ctx->have_records = ctx->have_solid = ctx->have_index = ctx->have_unitigs = false;
    ctx->P = p;
    const uint32_t w = p.k - p.m;
    const uint32_t n_sub = ((1u << p.buckets_log) + 1) << p.second_buckets_log;
    // physical units: about `unit_bytes` input bytes per unit (default 2000,
    // 1000 for colored builds),
    // at most 2^25 units (unit sizes then stay near unit_bytes up to ~30 GB
    // of input, so a schedule tuned on a sample holds at full size)
    uint32_t xb = 0;
    // (colored builds have smaller LDS tables per workgroup: smaller units)
    const uint64_t unit_bytes = ctx->tune.unit_bytes ? ctx->tune.unit_bytes : (n_colors ? 1000 : 2000);
    while (xb < 10 && (n_bytes >> xb) / n_sub > unit_bytes && ((uint64_t)n_sub << (xb + 1)) <= (1u << 25)) xb++;
    const uint32_t n_units = n_sub << xb;
    // slot: meta word + inline bases for the longest unique-minimizer record (2k - m)
    uint32_t sw = 1 + (2 * p.k - p.m + 15) / 16;
    uint32_t slot_words = 4;
    while (slot_words < sw) slot_words <<= 1;
    ctx->n_sub = n_sub;
    ctx->n_sub_all = n_sub;
    ctx->resplit_done = false;
    ctx->xb = xb;
    ctx->n_units = n_units;
    ctx->slot_words = slot_words;
    ctx->n_input = n_bytes;
    ctx->colored = n_colors > 0;
    ctx->color_words = 0;
    ctx->n_color_ids = 0;
    if (n_colors) {
        uint32_t max_color = 0;
        for (uint32_t i = 0; i < n_colors; i++) max_color = std::max(max_color, colors[i]);
        if (max_color >= MAX_COLORS)
            return fail(ctx, GGCAT_GPU_EINVAL, "color ids must be < " + std::to_string(MAX_COLORS));
        ctx->color_words = max_color / 32 + 1;
        ctx->n_color_ids = max_color + 1;
    }

    const uint8_t* d_in = data;
    // Host input of at least two upload chunks goes through the pinned
    // staging ring on the copy stream, chunk by chunk, while the staging pass
    // runs on the compute stream over the tiles already resident (first
    // attempt of pass 1); smaller input is copied whole first.
    const uint64_t up_chunk = ctx->tune.upload_chunk ? std::min<uint64_t>(ctx->tune.upload_chunk, PIN_CHUNK) : PIN_CHUNK;
    bool chunked = false;
    if (!input_on_device) {
        ENSURE(ctx->input, n_bytes + 16);
        d_in = P_<uint8_t>(ctx->input);
        chunked = n_bytes >= 2 * up_chunk;
        if (chunked && !ctx->pin_ring &&
            hipHostMalloc(&ctx->pin_ring, 2 * PIN_CHUNK, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            ctx->pin_ring = nullptr;
            chunked = false;
        }
        if (!chunked) HIPCHK(hipMemcpy(ctx->input.p, data, n_bytes, hipMemcpyHostToDevice));
    }
    // w == 16 forms window minima in registers; its tile of 960 windows (plus
    // the w + 1 halo) is exactly 62 blocks of 16, one per lane.  It keeps the
    // low 32 bits of each window minimum, which must name the unique bit,
    // bucket, second bucket and unit.
    const bool regs16 = (w == 16 && p.k <= 47 && 1 + p.buckets_log + p.second_buckets_log + xb <= 32);
    const uint32_t T = regs16 ? 960 : TILE_T;
    const uint64_t n_tiles = (n_bytes + T - 1) / T;
    const uint64_t seg_words = n_tiles * T / 32 + 2;
    ENSURE(ctx->seg_bits, seg_words * 4);
    ENSURE(ctx->unit_count, (size_t)n_units * 4);
    ENSURE(ctx->unit_cursor, (size_t)n_units * 8);
    ENSURE(ctx->unit_kmers, (size_t)n_units * 4);
    ENSURE(ctx->unit_off, ((size_t)n_units + 1) * 8);
    ENSURE(ctx->unit_dcount, (size_t)n_units * 4);
    ENSURE(ctx->unit_dregion, (size_t)n_units * 4);
    ENSURE(ctx->unit_skmers, (size_t)n_units * 4);
    ENSURE(ctx->pool_cursor, 8);
    ENSURE(ctx->error_flag, 8);
    if (n_colors) {
        ENSURE(ctx->color_starts, (size_t)n_colors * 8);
        ENSURE(ctx->colors, (size_t)n_colors * 4);
        HIPCHK(hipMemcpy(ctx->color_starts.p, color_starts, (size_t)n_colors * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->colors.p, colors, (size_t)n_colors * 4, hipMemcpyHostToDevice));
    }

    BucketArgs A;
    memset(&A, 0, sizeof A);  // padding included: the struct is part of the graph key
    A.seq = d_in;
    A.n = n_bytes;
    A.k = p.k;
    A.m = p.m;
    A.w = w;
    A.b = p.buckets_log;
    A.sb = p.second_buckets_log;
    A.xb = xb;
    A.canonical = p.forward_only ? 0 : 1;
    A.T = T;
    A.n_tiles = n_tiles;
    A.seg_bits = P_<uint32_t>(ctx->seg_bits);
    A.sub_count = P_<uint32_t>(ctx->unit_count);
    A.unit_ck = P_<unsigned long long>(ctx->unit_cursor);
    A.sub_off = P_<uint64_t>(ctx->unit_off);
    A.sub_kmers = P_<uint32_t>(ctx->unit_kmers);
    A.unit_dcount = P_<uint32_t>(ctx->unit_dcount);
    A.unit_dregion = P_<uint32_t>(ctx->unit_dregion);
    A.unit_skmers = P_<uint32_t>(ctx->unit_skmers);
    A.slot_words = slot_words;
    A.pool_cursor = P_<unsigned long long>(ctx->pool_cursor);
    A.color_starts = n_colors ? P_<uint64_t>(ctx->color_starts) : nullptr;
    A.colors = n_colors ? P_<uint32_t>(ctx->colors) : nullptr;
    A.n_colors = n_colors;
    A.error_flag = P_<uint32_t>(ctx->error_flag);
    nthash_tables(p.m, A.gtab);
    const TileGeom G = tile_geom(T, w, p.k, p.m, regs16);
    ENSURE(ctx->stage_cursor, 8);
    ENSURE(ctx->defer_count, 8);
    ENSURE(ctx->kmer_total, 16);  // k-mer occurrences, records
    ENSURE(ctx->defer, std::max<uint64_t>(n_tiles, 1) * 24);
    A.spill_cursor = P_<unsigned long long>(ctx->stage_cursor);
    A.defer = P_<uint64_t>(ctx->defer);
    A.defer_count = P_<uint32_t>(ctx->defer_count);
    A.kmer_total = P_<unsigned long long>(ctx->kmer_total);
    ENSURE(ctx->scan_part, (((size_t)n_units + SCAN_CHUNK - 1) / SCAN_CHUNK + 1) * 8);

    // Pass 1: clear the unit counts, k-mer counts, group-start bitmap and
    // cursors; build every record, count it in its unit and write it to its
    // unit's region; resolve the deferred far-start records.  With graphs
    // enabled this sequence is captured once per argument set and replayed
    // (fills are kernels, never memsets).  A unit's region holds 2.5x the
    // records expected per unit (2 per w+1 windows on random sequence); the
    // rest spill to a list sized for 1/64 of the input bytes, grown to the
    // counted size when that was short.
    const uint64_t est_records = n_bytes * 2 / (w + 1) + 1;
    const uint32_t unit_cap =
        ctx->tune.unit_cap ? ctx->tune.unit_cap
                           : (uint32_t)std::max<uint64_t>(16, std::min<uint64_t>(1u << 16, ((est_records * 5 / 2) / n_units + 15) / 8 * 8));
    A.unit_cap = unit_cap;
    uint64_t spill_cap = std::max<uint64_t>(n_bytes / 64 + 65536, ctx->stage_unit.cap / 4);
    uint64_t pool_cap = std::max<uint64_t>(ctx->pool.cap / 4, n_bytes / 64 + 65536);
    const uint32_t small_grid = (uint32_t)std::min<uint64_t>((std::max<uint64_t>(n_tiles, 1) + 255) / 256, 65535);
    auto pass_1 = [&]() {
        hipLaunchKernelGGL(fill32_kernel, dim3(256), dim3(256), 0, ctx->stream, reinterpret_cast<uint32_t*>(A.unit_ck),
                           0u, 2 * (uint64_t)n_units);
        hipLaunchKernelGGL(fill32_kernel, dim3(256), dim3(256), 0, ctx->stream, A.seg_bits, 0u, (uint64_t)seg_words);
        hipLaunchKernelGGL(bucket_reset_kernel, dim3(1), dim3(64), 0, ctx->stream, A);
        if (n_tiles) {
            if (regs16)
                hipLaunchKernelGGL(bucket_stage_kernel<16>, dim3((uint32_t)n_tiles), dim3(64), G.bytes, ctx->stream, A);
            else
                hipLaunchKernelGGL(bucket_stage_kernel<0>, dim3((uint32_t)n_tiles), dim3(64), G.bytes, ctx->stream, A);
            hipLaunchKernelGGL(bucket_defer_kernel, dim3(small_grid), dim3(256), 0, ctx->stream, A);
        }
        hipLaunchKernelGGL(unit_split_kernel, dim3(std::min<uint32_t>((n_units + 255) / 256, 4096)), dim3(256), 0,
                           ctx->stream, A.unit_ck, A.sub_count, A.sub_kmers, n_units, A.kmer_total);
    };
    // The same pass over host input uploaded chunk by chunk: chunk j is copied
    // into staging buffer j % 2 by host threads (once the upload that last
    // read that buffer is done) and uploaded on the copy stream; the compute
    // stream waits for it and stages every tile whose bytes, with the margin
    // a tile reads past its windows, are now resident.  Tiles are launched
    // with their absolute index (A.tile0), so the records are the ones of a
    // single launch.
    auto pass_1_chunked = [&]() -> int {
        hipLaunchKernelGGL(fill32_kernel, dim3(256), dim3(256), 0, ctx->stream, reinterpret_cast<uint32_t*>(A.unit_ck),
                           0u, 2 * (uint64_t)n_units);
        hipLaunchKernelGGL(fill32_kernel, dim3(256), dim3(256), 0, ctx->stream, A.seg_bits, 0u, (uint64_t)seg_words);
        hipLaunchKernelGGL(bucket_reset_kernel, dim3(1), dim3(64), 0, ctx->stream, A);
        HIPCHK(hipGetLastError());
        // no upload of an earlier call may still read the staging buffers
        HIPCHK(hipStreamSynchronize(ctx->copy_stream));
        uint8_t* ring = static_cast<uint8_t*>(ctx->pin_ring);
        uint8_t* d_dst = P_<uint8_t>(ctx->input);
        const uint64_t margin = 2 * (uint64_t)T + 4 * (uint64_t)p.k + 4096;
        const uint32_t n_threads = std::max(1u, std::min(HOST_COPY_THREADS, std::thread::hardware_concurrency()));
        bool pending[2] = {false, false};
        uint64_t tiles_done = 0;
        uint32_t j = 0;
        for (uint64_t o = 0; o < n_bytes; o += up_chunk, j++) {
            const uint32_t b = j & 1;
            if (pending[b]) HIPCHK(hipEventSynchronize(ctx->up_ev[b]));
            const uint64_t len = std::min<uint64_t>(up_chunk, n_bytes - o);
            uint8_t* stage_buf = ring + b * PIN_CHUNK;
            const uint64_t slice = (len + n_threads - 1) / n_threads;
            std::vector<std::thread> th;
            for (uint32_t q = 1; q < n_threads; q++)
                th.emplace_back([&, q] {
                    const uint64_t lo = std::min(len, q * slice), hi = std::min(len, (q + 1) * slice);
                    if (hi > lo) memcpy(stage_buf + lo, data + o + lo, hi - lo);
                });
            memcpy(stage_buf, data + o, std::min(len, slice));
            for (auto& x : th) x.join();
            HIPCHK(hipMemcpyAsync(d_dst + o, stage_buf, len, hipMemcpyHostToDevice, ctx->copy_stream));
            HIPCHK(hipEventRecord(ctx->up_ev[b], ctx->copy_stream));
            pending[b] = true;
            HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->up_ev[b], 0));
            const uint64_t resident = o + len;
            uint64_t safe = resident == n_bytes ? n_tiles : (resident > margin ? (resident - margin) / T : 0);
            safe = std::min<uint64_t>(safe, n_tiles);
            if (safe > tiles_done) {
                A.tile0 = tiles_done;
                const dim3 grid((uint32_t)(safe - tiles_done));
                if (regs16)
                    hipLaunchKernelGGL(bucket_stage_kernel<16>, grid, dim3(64), G.bytes, ctx->stream, A);
                else
                    hipLaunchKernelGGL(bucket_stage_kernel<0>, grid, dim3(64), G.bytes, ctx->stream, A);
                HIPCHK(hipGetLastError());
                tiles_done = safe;
            }
        }
        A.tile0 = 0;
        if (n_tiles) hipLaunchKernelGGL(bucket_defer_kernel, dim3(small_grid), dim3(256), 0, ctx->stream, A);
        hipLaunchKernelGGL(unit_split_kernel, dim3(std::min<uint32_t>((n_units + 255) / 256, 4096)), dim3(256), 0,
                           ctx->stream, A.unit_ck, A.sub_count, A.sub_kmers, n_units, A.kmer_total);
        HIPCHK(hipGetLastError());
        return GGCAT_GPU_OK;
    };
    uint64_t staged = 0, pool_used = 0;
    uint32_t errf = 0;
    for (int attempt = 0;; attempt++) {
        const uint64_t region_slots = (uint64_t)n_units * unit_cap;
        ENSURE(ctx->stage_slots, region_slots * slot_words * 4);
        ENSURE(ctx->spill_slots, spill_cap * slot_words * 4);
        ENSURE(ctx->stage_unit, spill_cap * 4);
        ENSURE(ctx->stage_rank, spill_cap * 4);
        if (n_colors) {
            ENSURE(ctx->stage_color, region_slots * 4);
            ENSURE(ctx->spill_color, spill_cap * 4);
        }
        ENSURE(ctx->pool, pool_cap * 4);
        A.region = P_<uint32_t>(ctx->stage_slots);
        A.region_color = n_colors ? P_<uint32_t>(ctx->stage_color) : nullptr;
        A.spill_slots = P_<uint32_t>(ctx->spill_slots);
        A.spill_unit = P_<uint32_t>(ctx->stage_unit);
        A.spill_rank = P_<uint32_t>(ctx->stage_rank);
        A.spill_color = n_colors ? P_<uint32_t>(ctx->spill_color) : nullptr;
        A.spill_cap = spill_cap;
        A.pool = P_<uint32_t>(ctx->pool);
        A.pool_cap = pool_cap;
        HIPCHK(hipEventRecord(ctx->ev[0], ctx->stream));
        if (attempt == 0 && chunked) {
            const int r = pass_1_chunked();
            if (r != GGCAT_GPU_OK) return r;
        } else if (ctx->tune.use_graphs) {
            ggcat_gpu_ctx::GraphKeyA key;
            memset(&key, 0, sizeof key);
            key.args = A;
            key.seg_words = seg_words;
            key.n_units = n_units;
            if (!ctx->graph_a || memcmp(&key, &ctx->graph_a_key, sizeof key) != 0) {
                if (ctx->graph_a) (void)hipGraphExecDestroy(ctx->graph_a);
                ctx->graph_a = nullptr;
                hipGraph_t g;
                HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
                pass_1();
                hipError_t e = hipStreamEndCapture(ctx->stream, &g);
                if (e != hipSuccess) return fail(ctx, GGCAT_GPU_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
                e = hipGraphInstantiate(&ctx->graph_a, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                if (e != hipSuccess) {
                    ctx->graph_a = nullptr;
                    return fail(ctx, GGCAT_GPU_EHIP, std::string("graph instantiate: ") + hipGetErrorString(e));
                }
                ctx->graph_a_key = key;
            }
            HIPCHK(hipGraphLaunch(ctx->graph_a, ctx->stream));
        } else {
            pass_1();
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(ctx->ev[1], ctx->stream));
        HIPCHK(hipMemcpyAsync(&staged, ctx->stage_cursor.p, 8, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(&pool_used, ctx->pool_cursor.p, 8, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(&errf, ctx->error_flag.p, 4, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        if (errf & 8) return fail(ctx, GGCAT_GPU_ENOMEM, "a physical unit exceeds 2^32 k-mer occurrences");
        if (errf) return fail(ctx, GGCAT_GPU_EHIP, "bucketing: a group start was not found");
        if (staged <= spill_cap && pool_used <= pool_cap) break;
        if (attempt == 1) return fail(ctx, GGCAT_GPU_ENOMEM, "bucketing spill list overflow");
        spill_cap = std::max(spill_cap, staged + 1024);
        pool_cap = std::max(pool_cap, pool_used + 1024);
    }
    ctx->pool_words = pool_used;
    ctx->n_spill = staged;

    // Pass 2: identical records of a unit counted once, unit offsets of the
    // stored records, then every stored record to its unit's slots.
    if (n_units) {
        // units of more than DEDUP_SMALL region records: 256-thread workgroups
        // (their region must fit the registers that stage it: 8 pieces per
        // thread) on the copy stream, beside the wavefront-per-unit launch
        // (the two take disjoint units); both persistent grids are as large
        // as the CUs hold at once
        const bool wide = unit_cap > DEDUP_SMALL && DEDUP_WIDE * slot_words / 4 <= 8 * 256;
        ctx->dd_max = wide ? DEDUP_WIDE : DEDUP_SMALL;
        const size_t lds_small = dedup_lds_bytes(DEDUP_SMALL, slot_words, n_colors != 0);
        const size_t lds_wide = dedup_lds_bytes(DEDUP_WIDE, slot_words, n_colors != 0);
        int occ = 0;
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(unit_dedup_kernel), 64,
                                                            lds_small));
        const uint32_t grid = (uint32_t)std::min<uint64_t>(n_units, (uint64_t)ctx->n_cu * std::max(occ, 1));
        if (wide) {
            HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &occ, reinterpret_cast<const void*>(unit_dedup_wide_kernel), 256, lds_wide));
            const uint32_t wgrid = (uint32_t)std::min<uint64_t>(n_units, (uint64_t)ctx->n_cu * std::max(occ, 1));
            HIPCHK(hipEventRecord(ctx->dd_ev[0], ctx->stream));
            HIPCHK(hipStreamWaitEvent(ctx->copy_stream, ctx->dd_ev[0], 0));
            hipLaunchKernelGGL(unit_dedup_wide_kernel, dim3(wgrid), dim3(256), lds_wide, ctx->copy_stream, A, n_units);
            HIPCHK(hipGetLastError());
            HIPCHK(hipEventRecord(ctx->dd_ev[1], ctx->copy_stream));
        }
        hipLaunchKernelGGL(unit_dedup_kernel, dim3(grid), dim3(64), lds_small, ctx->stream, A, n_units, wide ? 1u : 0u);
        HIPCHK(hipGetLastError());
        if (wide) HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->dd_ev[1], 0));
    }
    HIPCHK(hipEventRecord(ctx->ev[15], ctx->stream));
    const int rs = device_scan(ctx, A.unit_dcount, P_<uint64_t>(ctx->unit_off), n_units);
    if (rs != GGCAT_GPU_OK) return rs;
    HIPCHK(hipEventRecord(ctx->ev[2], ctx->stream));
    uint64_t total = 0, kmer_total[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(&total, P_<uint64_t>(ctx->unit_off) + n_units, 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(kmer_total, A.kmer_total, 16, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->n_records = kmer_total[1];
    ctx->n_stored = total;
    ENSURE(ctx->slots, std::max<uint64_t>(total, 1) * slot_words * 4);
    if (n_colors) ENSURE(ctx->slot_color, std::max<uint64_t>(total, 1) * 4);
    A.slots = P_<uint32_t>(ctx->slots);
    A.n_slots = total;
    A.slot_color = n_colors ? P_<uint32_t>(ctx->slot_color) : nullptr;
    HIPCHK(hipEventRecord(ctx->ev[3], ctx->stream));
    if (total) {
        const uint32_t grid = (uint32_t)std::min<uint64_t>(((uint64_t)n_units + 3) / 4, (uint64_t)ctx->n_cu * 32);
        hipLaunchKernelGGL(unit_compact_kernel, dim3(grid), dim3(256), 0, ctx->stream, A, n_units);
    }
    if (staged) {
        const uint32_t grid = (uint32_t)std::min<uint64_t>((staged + 255) / 256, (uint64_t)ctx->n_cu * 64);
        hipLaunchKernelGGL(spill_place_kernel, dim3(grid), dim3(256), 0, ctx->stream, A, staged);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ctx->ev[4], ctx->stream));
    HIPCHK(hipMemcpyAsync(&errf, ctx->error_flag.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (errf) return fail(ctx, GGCAT_GPU_EHIP, "bucketing: a record rank outside its unit");
    float t;
    HIPCHK(hipEventElapsedTime(&t, ctx->ev[0], ctx->ev[1]));
    ctx->ms[0] = t;
    HIPCHK(hipEventElapsedTime(&t, ctx->ev[15], ctx->ev[2]));
    ctx->ms[1] = t;
    float t_dedup;
    HIPCHK(hipEventElapsedTime(&t_dedup, ctx->ev[1], ctx->ev[15]));
    HIPCHK(hipEventElapsedTime(&t, ctx->ev[3], ctx->ev[4]));
    ctx->ms[2] = t + t_dedup;
    ctx->ms[5] = ctx->ms[0] + ctx->ms[1] + ctx->ms[2];
    ctx->kmer_occ = kmer_total[0];  // sizes the merge output
    ctx->have_records = true;
    return GGCAT_GPU_OK;
}

// Outlier sub-buckets of the last bucketing as crates/kmers_transform/src/lib.rs:171-371
// would find them: buckets ordered by size (largest first, after the
// min(normal / 8, 2 * threads) smallest ones, then alternating a large one
// with small ones that match its size, :195-243), a running average of
// records per non-empty sub-bucket over the buckets taken so far (at least
// MIN_AVERAGE_CAP = 5000), and a sub-bucket above 8x that average
// (MAX_SUBBUCKET_AVERAGE_MULTIPLIER, crates/config/src/lib.rs:62-66) is
// resplit into next_pow2(records / average) clamped to [4, 1024] buckets.
// The reference orders by compressed bucket-file size and counts compacted
// records; this uses the serialized record bytes and the records as staged.
namespace {
struct OutlierInfo {
    uint32_t sub, resplit;
    uint64_t records;
};

static int find_outliers(ggcat_gpu_ctx* ctx, uint32_t threads, std::vector<OutlierInfo>& found) {
    found.clear();
    const uint32_t n_sub = ctx->n_sub, sb = ctx->P.second_buckets_log;
    std::vector<uint64_t> rec(n_sub), bas(n_sub);
    DevBuf tmp;
    if (ensure(ctx, tmp, (size_t)n_sub * 16) != GGCAT_GPU_OK) return GGCAT_GPU_ENOMEM;
    hipLaunchKernelGGL(sub_totals_kernel, dim3((n_sub + 255) / 256), dim3(256), 0, ctx->stream,
                       P_<uint32_t>(ctx->unit_count), P_<uint32_t>(ctx->unit_kmers), n_sub, ctx->xb, ctx->P.k,
                       P_<uint64_t>(tmp), P_<uint64_t>(tmp) + n_sub);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(rec.data(), tmp.p, (size_t)n_sub * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(bas.data(), P_<uint64_t>(tmp) + n_sub, (size_t)n_sub * 8, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(tmp.p);
    if (e != hipSuccess) return fail(ctx, GGCAT_GPU_EHIP, std::string("outlier counts: ") + hipGetErrorString(e));
    const uint32_t n_buckets = n_sub >> sb, normal = n_buckets - 1;
    // bucket sizes: serialized record bytes (second-bucket byte, two length
    // bytes, packed bases)
    std::vector<std::pair<uint64_t, uint32_t>> sizes(n_buckets);
    for (uint32_t b = 0; b < n_buckets; b++) {
        uint64_t sz = 0;
        for (uint32_t s = b << sb; s < (b + 1) << sb; s++) sz += rec[s] * 3 + (bas[s] + 3 * rec[s]) / 4;
        sizes[b] = {sz, b};
    }
    // sort_by_key (stable, ascending) then reverse
    std::stable_sort(sizes.begin(), sizes.end(),
                     [](const std::pair<uint64_t, uint32_t>& a, const std::pair<uint64_t, uint32_t>& b) {
                         return a.first < b.first;
                     });
    std::reverse(sizes.begin(), sizes.end());
    std::vector<uint32_t> order;
    size_t start = 0, end = sizes.size();
    int64_t matched = 0;
    uint64_t estimators = std::min<uint64_t>(normal / 8, 2ull * std::max<uint32_t>(threads, 1));
    while (start != end && estimators > 0) {
        end--;
        estimators--;
        order.push_back(sizes[end].second);
    }
    while (start != end) {
        if (matched <= 0) {
            matched = (int64_t)sizes[start].first;
            order.push_back(sizes[start++].second);
        } else {
            matched -= (int64_t)sizes[end - 1].first;
            order.push_back(sizes[--end].second);
        }
    }
    uint64_t tot_seq = 0, tot_cnt = 0;
    for (uint32_t b : order) {
        for (uint32_t s = b << sb; s < (b + 1) << sb; s++)
            if (rec[s]) {
                tot_seq += rec[s];
                tot_cnt++;
            }
        const uint64_t avg = std::max<uint64_t>(tot_seq / std::max<uint64_t>(tot_cnt, 1), 5000);
        for (uint32_t s = b << sb; s < (b + 1) << sb; s++) {
            if (!rec[s] || rec[s] <= avg * 8) continue;
            uint64_t c = 1;
            while (c < rec[s] / avg) c <<= 1;
            c = std::min<uint64_t>(std::max<uint64_t>(c, 4), 1024);
            found.push_back(OutlierInfo{s, (uint32_t)c, rec[s]});
        }
    }
    return GGCAT_GPU_OK;
}

static void free_buf(DevBuf& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}
}  // namespace

int ggcat_gpu_outliers(ggcat_gpu_ctx* ctx, uint32_t threads, uint32_t* n_outliers, uint32_t* sub, uint32_t* resplit,
                       uint64_t* records, uint32_t cap) {
    if (!ctx || !n_outliers) return GGCAT_GPU_EINVAL;
    if (!ctx->have_records) return fail(ctx, GGCAT_GPU_ESTATE, "no bucketing result");
    if (ctx->resplit_done) return fail(ctx, GGCAT_GPU_ESTATE, "the bucketing's outliers were resplit already");
    HIPCHK(hipSetDevice(ctx->device));
    std::vector<OutlierInfo> found;
    const int rc = find_outliers(ctx, threads, found);
    if (rc != GGCAT_GPU_OK) return rc;
    for (uint32_t i = 0; i < found.size() && i < cap; i++) {
        if (sub) sub[i] = found[i].sub;
        if (resplit) resplit[i] = found[i].resplit;
        if (records) records[i] = found[i].records;
    }
    *n_outliers = (uint32_t)found.size();
    return GGCAT_GPU_OK;
}

namespace {
// The listed sub-buckets resplit as crates/kmers_transform/src/resplitter.rs:47-183
// does (resplit.hip): the arena is rebuilt with their units emptied and one
// unit per resplit bucket appended after the bucketing's units, as extra
// sub-buckets n_sub + base_i .. n_sub + base_i + count_i (duplicates bucket
// last).  Counts are powers of two.
static int resplit_subs(ggcat_gpu_ctx* ctx, const std::vector<OutlierInfo>& found) {
    int rc = GGCAT_GPU_OK;
    const uint32_t n_out = (uint32_t)found.size();
    const uint32_t n_sub = ctx->n_sub, xb = ctx->xb, n_units = ctx->n_units, SW = ctx->slot_words, k = ctx->P.k;
    std::vector<uint32_t> logs(n_out), base(n_out);
    uint64_t n_extra = 0;
    for (uint32_t i = 0; i < n_out; i++) {
        uint32_t lg = 0;
        while ((1u << lg) < found[i].resplit) lg++;
        logs[i] = lg;
        base[i] = (uint32_t)n_extra;
        n_extra += (1u << lg) + 1;
    }
    if (n_out == 0) return GGCAT_GPU_OK;
    const uint64_t n_units_all = (uint64_t)(n_sub + n_extra) << xb;
    if (n_units_all > (1ull << 31)) return fail(ctx, GGCAT_GPU_ENOMEM, "too many resplit units");
    // each outlier's stored records: one contiguous range of the unit-major arena
    std::vector<uint64_t> prefix(n_out + 1, 0), first(n_out);
    std::vector<uint8_t> is_out(n_sub, 0);
    for (uint32_t i = 0; i < n_out; i++) {
        uint64_t ab[2];
        HIPCHK(hipMemcpy(&ab[0], P_<uint64_t>(ctx->unit_off) + ((uint64_t)found[i].sub << xb), 8, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(&ab[1], P_<uint64_t>(ctx->unit_off) + ((uint64_t)(found[i].sub + 1) << xb), 8,
                         hipMemcpyDeviceToHost));
        first[i] = ab[0];
        prefix[i + 1] = prefix[i] + (ab[1] - ab[0]);
        is_out[found[i].sub] = 1;
    }
    const uint32_t m2 = k > 11 ? k - 10 : std::min<uint32_t>(ctx->P.m, 2);
    if (k - m2 > RS_RING) return fail(ctx, GGCAT_GPU_EINVAL, "resplit window wider than supported");
    DevBuf meta_b, isout_b, ext_b, need_b, cur_b;
    auto cleanup = [&]() {
        free_buf(meta_b);
        free_buf(isout_b);
        free_buf(ext_b);
        free_buf(need_b);
        free_buf(cur_b);
    };
    // host tables: prefix [n_out + 1], first [n_out], logs [n_out], base [n_out]
    const size_t meta_bytes = (size_t)(2 * n_out + 1) * 8 + (size_t)2 * n_out * 4;
    std::vector<uint8_t> mt(meta_bytes);
    memcpy(mt.data(), prefix.data(), (n_out + 1) * 8);
    memcpy(mt.data() + (n_out + 1) * 8, first.data(), n_out * 8);
    memcpy(mt.data() + (2 * n_out + 1) * 8, logs.data(), n_out * 4);
    memcpy(mt.data() + (2 * n_out + 1) * 8 + n_out * 4, base.data(), n_out * 4);
    if ((rc = ensure(ctx, meta_b, meta_bytes)) != GGCAT_GPU_OK || (rc = ensure(ctx, isout_b, n_sub)) != GGCAT_GPU_OK ||
        (rc = ensure(ctx, ext_b, (size_t)n_extra * 12)) != GGCAT_GPU_OK ||
        (rc = ensure(ctx, need_b, 48)) != GGCAT_GPU_OK) {
        cleanup();
        return rc;
    }
#define RS_CHK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) {                                                                        \
            cleanup();                                                                                 \
            return fail(ctx, GGCAT_GPU_EHIP, std::string("resplit: ") + #expr + ": " + hipGetErrorString(e_)); \
        }                                                                                              \
    } while (0)
    RS_CHK(hipMemcpyAsync(meta_b.p, mt.data(), meta_bytes, hipMemcpyHostToDevice, ctx->stream));
    RS_CHK(hipMemcpyAsync(isout_b.p, is_out.data(), n_sub, hipMemcpyHostToDevice, ctx->stream));
    RS_CHK(hipMemsetAsync(ext_b.p, 0, (size_t)n_extra * 12, ctx->stream));
    RS_CHK(hipMemsetAsync(need_b.p, 0, 48, ctx->stream));
    RS_CHK(hipMemsetAsync(ctx->error_flag.p, 0, 4, ctx->stream));
    ResplitArgs R{};
    R.slots = P_<uint32_t>(ctx->slots);
    R.slot_color = ctx->colored ? P_<uint32_t>(ctx->slot_color) : nullptr;
    R.pool = P_<uint32_t>(ctx->pool);
    R.pool_words = ctx->pool_words;
    R.SW = SW;
    R.k = k;
    R.m2 = m2;
    R.w2 = k - m2;
    R.n_out = n_out;
    R.rec_prefix = reinterpret_cast<const uint64_t*>(meta_b.p);
    R.rec_first = R.rec_prefix + n_out + 1;
    R.log_buckets = reinterpret_cast<const uint32_t*>(R.rec_first + n_out);
    R.ext_base = R.log_buckets + n_out;
    R.ext_records = P_<uint32_t>(ext_b);
    R.ext_kmers = R.ext_records + n_extra;
    R.ext_cursor = R.ext_kmers + n_extra;
    R.pool_need = P_<unsigned long long>(need_b);
    R.sums = R.pool_need + 2;
    R.error_flag = P_<uint32_t>(ctx->error_flag);
    const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((prefix[n_out] + 255) / 256, 4096));
    hipLaunchKernelGGL(resplit_count_kernel, dim3(grid), dim3(256), 0, ctx->stream, R);
    RS_CHK(hipGetLastError());
    uint64_t pool_need = 0, sums[4] = {0, 0, 0, 0};
    uint32_t errf = 0;
    RS_CHK(hipMemcpyAsync(&pool_need, need_b.p, 8, hipMemcpyDeviceToHost, ctx->stream));
    RS_CHK(hipMemcpyAsync(sums, R.sums, 32, hipMemcpyDeviceToHost, ctx->stream));
    RS_CHK(hipMemcpyAsync(&errf, ctx->error_flag.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    RS_CHK(hipStreamSynchronize(ctx->stream));
    if (errf) {
        cleanup();
        return fail(ctx, GGCAT_GPU_EHIP, "resplit: a record does not fit its slot or the pool");
    }
    // the rebuilt unit arrays and arena
    DevBuf dcount2, count2, kmers2, skmers2, off2, slots2, color2, pool2;
    auto cleanup2 = [&]() {
        free_buf(dcount2);
        free_buf(count2);
        free_buf(kmers2);
        free_buf(skmers2);
        free_buf(off2);
        free_buf(slots2);
        free_buf(color2);
        free_buf(pool2);
    };
    if ((rc = ensure(ctx, dcount2, n_units_all * 4)) != GGCAT_GPU_OK ||
        (rc = ensure(ctx, count2, n_units_all * 4)) != GGCAT_GPU_OK ||
        (rc = ensure(ctx, kmers2, n_units_all * 4)) != GGCAT_GPU_OK ||
        (rc = ensure(ctx, skmers2, n_units_all * 4)) != GGCAT_GPU_OK ||
        (rc = ensure(ctx, off2, (n_units_all + 1) * 8)) != GGCAT_GPU_OK) {
        cleanup();
        cleanup2();
        return rc;
    }
    hipLaunchKernelGGL(resplit_units_kernel, dim3((uint32_t)std::min<uint64_t>((n_units_all + 255) / 256, 65535)),
                       dim3(256), 0, ctx->stream, P_<uint32_t>(ctx->unit_dcount), P_<uint32_t>(ctx->unit_count),
                       P_<uint32_t>(ctx->unit_kmers), P_<uint32_t>(ctx->unit_skmers), P_<uint8_t>(isout_b), n_units,
                       (uint32_t)n_units_all, xb, n_sub, R.ext_records, R.ext_kmers, P_<uint32_t>(dcount2),
                       P_<uint32_t>(count2), P_<uint32_t>(kmers2), P_<uint32_t>(skmers2));
    RS_CHK(hipGetLastError());
    if ((rc = device_scan(ctx, P_<uint32_t>(dcount2), P_<uint64_t>(off2), (uint32_t)n_units_all)) != GGCAT_GPU_OK) {
        cleanup();
        cleanup2();
        return rc;
    }
    uint64_t total2 = 0;
    RS_CHK(hipMemcpyAsync(&total2, P_<uint64_t>(off2) + n_units_all, 8, hipMemcpyDeviceToHost, ctx->stream));
    RS_CHK(hipStreamSynchronize(ctx->stream));
    const uint64_t pool_cap2 = ctx->pool_words + pool_need;
    if ((rc = ensure(ctx, slots2, std::max<uint64_t>(total2, 1) * SW * 4)) != GGCAT_GPU_OK ||
        (ctx->colored && (rc = ensure(ctx, color2, std::max<uint64_t>(total2, 1) * 4)) != GGCAT_GPU_OK) ||
        (pool_need && pool_cap2 * 4 > ctx->pool.cap && (rc = ensure(ctx, pool2, pool_cap2 * 4)) != GGCAT_GPU_OK)) {
        cleanup();
        cleanup2();
        return rc;
    }
    if (pool2.p && ctx->pool_words)
        RS_CHK(hipMemcpyAsync(pool2.p, ctx->pool.p, ctx->pool_words * 4, hipMemcpyDeviceToDevice, ctx->stream));
    {
        const uint32_t g = (uint32_t)std::min<uint64_t>(((uint64_t)n_units + 3) / 4, (uint64_t)ctx->n_cu * 32);
        hipLaunchKernelGGL(resplit_copy_kernel, dim3(std::max<uint32_t>(g, 1)), dim3(256), 0, ctx->stream,
                           P_<uint32_t>(ctx->slots), R.slot_color, P_<uint64_t>(ctx->unit_off), P_<uint64_t>(off2),
                           P_<uint8_t>(isout_b), n_units, xb, SW, P_<uint32_t>(slots2),
                           ctx->colored ? P_<uint32_t>(color2) : nullptr);
        RS_CHK(hipGetLastError());
    }
    const uint64_t pool_used = ctx->pool_words;
    RS_CHK(hipMemcpyAsync(R.pool_need + 1, &pool_used, 8, hipMemcpyHostToDevice, ctx->stream));
    R.unit_off = P_<uint64_t>(off2);
    R.first_extra_unit = n_sub;
    R.xb = xb;
    R.out_slots = P_<uint32_t>(slots2);
    R.out_color = ctx->colored ? P_<uint32_t>(color2) : nullptr;
    R.out_pool = pool2.p ? P_<uint32_t>(pool2) : P_<uint32_t>(ctx->pool);
    R.pool_cursor = R.pool_need + 1;
    R.pool_cap = pool2.p ? pool2.cap / 4 : ctx->pool.cap / 4;
    hipLaunchKernelGGL(resplit_write_kernel, dim3(grid), dim3(256), 0, ctx->stream, R);
    RS_CHK(hipGetLastError());
    RS_CHK(hipMemcpyAsync(&errf, ctx->error_flag.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    RS_CHK(hipStreamSynchronize(ctx->stream));
#undef RS_CHK
    if (errf) {
        cleanup();
        cleanup2();
        return fail(ctx, GGCAT_GPU_EHIP, "resplit: a piece outside its unit or the pool");
    }
    cleanup();
    std::swap(ctx->unit_dcount, dcount2);
    std::swap(ctx->unit_count, count2);
    std::swap(ctx->unit_kmers, kmers2);
    std::swap(ctx->unit_skmers, skmers2);
    std::swap(ctx->unit_off, off2);
    std::swap(ctx->slots, slots2);
    if (ctx->colored) std::swap(ctx->slot_color, color2);
    if (pool2.p) std::swap(ctx->pool, pool2);
    cleanup2();  // the replaced buffers
    // unit_dregion / unit_cursor keep the bucketing's size; the next
    // bucketing sizes every unit array again
    ctx->pool_words = pool_used + pool_need;
    ctx->n_units = (uint32_t)n_units_all;
    ctx->n_sub_all = (uint32_t)(n_sub + n_extra);
    ctx->n_stored = total2;
    // record copies and k-mer occurrences: the outliers' records out, their pieces in
    ctx->n_records = ctx->n_records - sums[0] + sums[1];
    ctx->kmer_occ = ctx->kmer_occ - sums[2] + sums[3];
    ctx->resplit_done = true;
    ctx->have_solid = ctx->have_index = ctx->have_unitigs = false;
    return GGCAT_GPU_OK;
}

// First resplit sub-bucket of each listed sub-bucket (resplit_subs numbering).
static void resplit_firsts(const ggcat_gpu_ctx* ctx, const std::vector<OutlierInfo>& found, uint32_t* first_sub,
                    uint32_t cap) {
    uint32_t next = ctx->n_sub;
    for (uint32_t i = 0; i < found.size(); i++) {
        uint32_t c = 1;
        while (c < found[i].resplit) c <<= 1;
        if (first_sub && i < cap) first_sub[i] = next;
        next += c + 1;
    }
}
}  // namespace

int ggcat_gpu_resplit_outliers(ggcat_gpu_ctx* ctx, uint32_t threads, uint32_t* n_outliers, uint32_t* sub,
                               uint32_t* resplit, uint32_t* first_sub, uint32_t cap) {
    if (!ctx || !n_outliers) return GGCAT_GPU_EINVAL;
    if (!ctx->have_records) return fail(ctx, GGCAT_GPU_ESTATE, "no bucketing result");
    if (ctx->resplit_done) return fail(ctx, GGCAT_GPU_ESTATE, "the bucketing's outliers were resplit already");
    HIPCHK(hipSetDevice(ctx->device));
    std::vector<OutlierInfo> found;
    int rc = find_outliers(ctx, threads, found);
    if (rc != GGCAT_GPU_OK) return rc;
    *n_outliers = (uint32_t)found.size();
    for (uint32_t i = 0; i < found.size() && i < cap; i++) {
        if (sub) sub[i] = found[i].sub;
        if (resplit) resplit[i] = found[i].resplit;
    }
    resplit_firsts(ctx, found, first_sub, cap);
    return resplit_subs(ctx, found);
}

int ggcat_gpu_resplit(ggcat_gpu_ctx* ctx, uint32_t n, const uint32_t* sub, const uint32_t* count,
                      uint32_t* first_sub) {
    if (!ctx || (n && (!sub || !count))) return GGCAT_GPU_EINVAL;
    if (!ctx->have_records) return fail(ctx, GGCAT_GPU_ESTATE, "no bucketing result");
    if (ctx->resplit_done) return fail(ctx, GGCAT_GPU_ESTATE, "the bucketing's outliers were resplit already");
    HIPCHK(hipSetDevice(ctx->device));
    std::vector<OutlierInfo> list;
    std::vector<uint8_t> seen(ctx->n_sub, 0);
    for (uint32_t i = 0; i < n; i++) {
        if (sub[i] >= ctx->n_sub || seen[sub[i]]) return fail(ctx, GGCAT_GPU_EINVAL, "resplit: sub-bucket out of range or listed twice");
        if (count[i] < 1 || count[i] > 1024 || (count[i] & (count[i] - 1)))
            return fail(ctx, GGCAT_GPU_EINVAL, "resplit: bucket count must be a power of two in [1, 1024]");
        seen[sub[i]] = 1;
        list.push_back(OutlierInfo{sub[i], count[i], 0});
    }
    resplit_firsts(ctx, list, first_sub, n);
    return resplit_subs(ctx, list);
}

int ggcat_gpu_records_info(ggcat_gpu_ctx* ctx, uint64_t* n_records, uint32_t* n_sub, uint64_t* n_bases) {
    if (!ctx) return GGCAT_GPU_EINVAL;
    if (!ctx->have_records) return fail(ctx, GGCAT_GPU_ESTATE, "no bucketing result");
    if (n_records) *n_records = ctx->n_records;
    if (n_sub) *n_sub = ctx->n_sub_all;
    if (n_bases) *n_bases = ctx->kmer_occ + ctx->n_records * (ctx->P.k - 1);
    return GGCAT_GPU_OK;
}

int ggcat_gpu_records_copy(ggcat_gpu_ctx* ctx, uint32_t* sub, uint32_t* len, uint8_t* flags, uint16_t* minpos,
                           uint32_t* color, uint64_t* off, uint8_t* bases) {
    if (!ctx) return GGCAT_GPU_EINVAL;
    if (!ctx->have_records) return fail(ctx, GGCAT_GPU_ESTATE, "no bucketing result");
    HIPCHK(hipSetDevice(ctx->device));
    const uint32_t SW = ctx->slot_words;
    std::vector<uint32_t> slots(ctx->n_stored * SW), pool(ctx->pool_words), cols;
    std::vector<uint64_t> uoff(ctx->n_units + 1);
    if (!slots.empty()) HIPCHK(hipMemcpy(slots.data(), ctx->slots.p, slots.size() * 4, hipMemcpyDeviceToHost));
    if (!pool.empty()) HIPCHK(hipMemcpy(pool.data(), ctx->pool.p, pool.size() * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(uoff.data(), ctx->unit_off.p, uoff.size() * 8, hipMemcpyDeviceToHost));
    if (ctx->colored && ctx->n_stored) {
        cols.resize(ctx->n_stored);
        HIPCHK(hipMemcpy(cols.data(), ctx->slot_color.p, cols.size() * 4, hipMemcpyDeviceToHost));
    }
    static const char LET[4] = {'A', 'C', 'T', 'G'};
    // every stored record as many times as its multiplicity
    uint64_t bpos = 0, out_r = 0;
    uint32_t u = 0;
    for (uint64_t s = 0; s < ctx->n_stored; s++) {
        const uint64_t r = s;
        while (u + 1 < uoff.size() && uoff[u + 1] <= r) u++;
        const uint32_t* sl = slots.data() + r * SW;
        uint32_t meta = sl[0];
        uint32_t L;
        const uint32_t* bw;
        if (meta & META_LONG) {
            L = sl[1];
            bw = pool.data() + ((uint64_t)sl[2] | ((uint64_t)sl[3] << 32));
        } else {
            L = meta & 0xFF;
            bw = sl + 1;
        }
        for (uint32_t c = meta_mult(meta); c--; out_r++) {
            if (out_r >= ctx->n_records) return fail(ctx, GGCAT_GPU_EHIP, "stored records exceed the record count");
            if (sub) sub[out_r] = u >> ctx->xb;
            if (len) len[out_r] = L;
            if (flags) flags[out_r] = (meta >> 8) & 3;
            if (minpos) minpos[out_r] = (meta >> 10) & 0xFF;
            if (color) color[out_r] = ctx->colored ? cols[r] : 0;
            if (off) off[out_r] = bpos;
            if (bases)
                for (uint32_t i = 0; i < L; i++) bases[bpos + i] = LET[(bw[i >> 4] >> (2 * (i & 15))) & 3];
            bpos += L;
        }
    }
    if (out_r != ctx->n_records) return fail(ctx, GGCAT_GPU_EHIP, "stored records do not add up to the record count");
    return GGCAT_GPU_OK;
}

namespace {

// crates/io/src/varint.rs:8-22 (encode_varint): 7 bits per byte, low first,
// bit 7 set while more bytes follow.
static void put_varint(std::vector<uint8_t>& o, uint64_t v) {
    do {
        o.push_back((uint8_t)((v & 0x7F) | (v > 0x7F ? 0x80 : 0)));
        v >>= 7;
    } while (v);
}

// crates/io/src/varint.rs:26-57 (encode_varint_flags) with FlagsCount = 2:
// the flags in bits 6-7 of the first byte, 5 value bits and a continuation
// bit below them, then plain varint bytes.
static void put_varint_flags2(std::vector<uint8_t>& o, uint64_t v, uint8_t flags) {
    o.push_back((uint8_t)((flags << 6) | (v & 0x1F) | (v > 0x1F ? 0x20 : 0)));
    v >>= 5;
    while (v) {
        o.push_back((uint8_t)((v & 0x7F) | (v > 0x7F ? 0x80 : 0)));
        v >>= 7;
    }
}

}  // namespace

int ggcat_gpu_records_serialize(ggcat_gpu_ctx* ctx, uint8_t* out, uint64_t cap, uint64_t* bucket_off,
                                uint64_t* n_bytes) {
    if (!ctx) return GGCAT_GPU_EINVAL;
    if (!ctx->have_records) return fail(ctx, GGCAT_GPU_ESTATE, "no bucketing result");
    if (ctx->resplit_done)
        return fail(ctx, GGCAT_GPU_ESTATE, "resplit buckets have no bucket file of the first bucketing");
    HIPCHK(hipSetDevice(ctx->device));
    const uint32_t SW = ctx->slot_words, k = ctx->P.k, sb = ctx->P.second_buckets_log;
    const uint32_t n_buckets = ctx->n_sub >> sb;
    uint32_t min_size_log = 0;
    while ((1u << min_size_log) < k) min_size_log++;  // k.next_power_of_two().ilog2()
    std::vector<uint32_t> slots(ctx->n_stored * SW), pool(ctx->pool_words), cols;
    std::vector<uint64_t> uoff(ctx->n_units + 1);
    if (!slots.empty()) HIPCHK(hipMemcpy(slots.data(), ctx->slots.p, slots.size() * 4, hipMemcpyDeviceToHost));
    if (!pool.empty()) HIPCHK(hipMemcpy(pool.data(), ctx->pool.p, pool.size() * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(uoff.data(), ctx->unit_off.p, uoff.size() * 8, hipMemcpyDeviceToHost));
    if (ctx->colored && ctx->n_stored) {
        cols.resize(ctx->n_stored);
        HIPCHK(hipMemcpy(cols.data(), ctx->slot_color.p, cols.size() * 4, hipMemcpyDeviceToHost));
    }
    // Records are stored by unit, so the records of a bucket (its 2^sb
    // sub-buckets) are consecutive and each bucket is one stream.
    std::vector<uint8_t> o;
    std::vector<uint64_t> boff(n_buckets + 1, 0);
    uint32_t u = 0, cur_bucket = 0;
    uint32_t last_color = 0xFFFFFFFFu;  // MinBkSingleColor::default()
    for (uint64_t r = 0; r < ctx->n_stored; r++) {
        while (u + 1 < uoff.size() && uoff[u + 1] <= r) u++;
        const uint32_t sub = u >> ctx->xb, bucket = sub >> sb;
        if (bucket >= n_buckets) return fail(ctx, GGCAT_GPU_EHIP, "record outside the bucket range");
        while (cur_bucket < bucket) {
            boff[++cur_bucket] = o.size();
            last_color = 0xFFFFFFFFu;  // each bucket stream starts afresh
        }
        const uint32_t* sl = slots.data() + r * SW;
        const uint32_t meta = sl[0];
        uint32_t L;
        const uint32_t* bw;
        if (meta & META_LONG) {
            L = sl[1];
            bw = pool.data() + ((uint64_t)sl[2] | ((uint64_t)sl[3] << 32));
        } else {
            L = meta & 0xFF;
            bw = sl + 1;
        }
        if (L < k) return fail(ctx, GGCAT_GPU_EHIP, "record shorter than k");
        for (uint32_t c = meta_mult(meta); c--;) {  // the record once per copy (no multiplicity field)
        // creads_utils.rs:374-432 write_to: second bucket byte, no
        // multiplicity, length with minimizer position and flags
        // (compressed_read.rs:435-467 encode_length), packed bases, color.
        o.push_back((uint8_t)(sub & ((1u << sb) - 1)));
        put_varint_flags2(o, ((uint64_t)(L - k) << min_size_log) | ((meta >> 10) & 0xFF), (uint8_t)((meta >> 8) & 3));
        const uint8_t* bb = reinterpret_cast<const uint8_t*>(bw);
        o.insert(o.end(), bb, bb + (L + 3) / 4);
        const uint32_t tail = L & 3;  // bases past the end of the last byte are zero
        if (tail) o.back() &= (uint8_t)((1u << (2 * tail)) - 1);
        if (ctx->colored) {
            // crates/colors/src/parsers/separate.rs:98-115: 0 repeats the
            // previous record's color, otherwise color + 1
            put_varint(o, cols[r] == last_color ? 0 : (uint64_t)cols[r] + 1);
            last_color = cols[r];
        }
        }
    }
    while (cur_bucket < n_buckets) boff[++cur_bucket] = o.size();
    if (n_bytes) *n_bytes = o.size();
    if (bucket_off) memcpy(bucket_off, boff.data(), boff.size() * 8);
    if (out) {
        if (cap < o.size()) return fail(ctx, GGCAT_GPU_EINVAL, "serialization buffer too small");
        memcpy(out, o.data(), o.size());
    }
    return GGCAT_GPU_OK;
}

int ggcat_gpu_kmers_merge(ggcat_gpu_ctx* ctx) {
    if (!ctx) return GGCAT_GPU_EINVAL;
    if (!ctx->have_records) return fail(ctx, GGCAT_GPU_ESTATE, "no bucketing result");
    if (ctx->server) return fail(ctx, GGCAT_GPU_ESTATE, "close the query server before a new merge");
    ctx->have_index = ctx->have_unitigs = ctx->have_subsets = false;
    HIPCHK(hipSetDevice(ctx->device));
    const ggcat_gpu_params& p = ctx->P;
    const uint32_t key_words = (p.hash_kind == GGCAT_GPU_HASH_SEQ && p.k <= 32) ? 1 : 2;
    ctx->key_words = key_words;
    MergeArgs M{};
    M.slots = P_<uint32_t>(ctx->slots);
    M.slot_words = ctx->slot_words;
    M.pool = P_<uint32_t>(ctx->pool);
    M.k = p.k;
    M.min_mult = p.min_multiplicity;
    M.kind = p.hash_kind;
    M.key_words = key_words;
    const uint32_t W = ctx->colored ? ctx->color_words : 0;
    const size_t entry_bytes = 12 * key_words + 4 * W + 2;  // + claimed-slot list entry
    M.lds_cap = key_words == 1 ? 2048 : 1024;
    if (ctx->tune.lds_table_log) M.lds_cap = 1u << std::min<uint32_t>(std::max<uint32_t>(ctx->tune.lds_table_log, 8), 13);
    while (M.lds_cap > 256 && M.lds_cap * entry_bytes > 48 * 1024) M.lds_cap >>= 1;
    M.table_pct = ctx->tune.table_pct ? ctx->tune.table_pct : 200;
    M.color_words = W;
    M.slot_color = W ? P_<uint32_t>(ctx->slot_color) : nullptr;
    M.lds_try = ctx->tune.lds_max_occ ? ctx->tune.lds_max_occ : M.lds_cap * 16;
    const uint32_t item_occ = ctx->tune.item_occ ? ctx->tune.item_occ : M.lds_cap / 2;
    // identical records counted once (seq-hash keys up to k = 32, inline slots)
    M.dedup = (ctx->slot_words <= 8 && ctx->tune.no_dedup != 1) ? 1 : 0;
    M.tag_kmers = (p.hash_kind == GGCAT_GPU_HASH_SEQ && p.k <= 32 && ctx->tune.no_dedup != 2) ? 1 : 0;
    // work items whose units the compaction deduplicated skip the record hashing
    M.rec_direct = ctx->tune.no_dedup != 3 ? 1 : 0;
    // workgroup size of the merge (128 or 256 threads); a chunk of records
    // (rec_cap / 2) has at most one record per thread
    const uint32_t merge_threads = ctx->tune.merge_threads == 128 ? 128u : 256u;
    M.rec_cap = std::min<uint32_t>(ctx->tune.rec_cap ? ctx->tune.rec_cap : REC_CAP_DEFAULT, 2 * merge_threads);
    M.rec_lds_off = (uint32_t)(((size_t)M.lds_cap * entry_bytes + 64 + 15) / 16 * 16);
    // the uncolored rabin-karp128 variant spreads the distinct records' k-mers
    // over the workgroup in segments (tags in the record table's LDS)
    const bool rk_seg = M.dedup && key_words == 2 && W == 0 && !M.tag_kmers && ctx->slot_words == 8 &&
                        p.hash_kind == GGCAT_GPU_HASH_RK128;
    const size_t lds_bytes =
        M.dedup ? M.rec_lds_off + rec_table_bytes(ctx->slot_words, (M.tag_kmers || rk_seg) ? p.k : 0xFFFFu, M.rec_cap, rk_seg)
                : (size_t)M.lds_cap * entry_bytes + 64;
    M.pool_words = ctx->pool_words;
    set_rk_constants(M, p.k);
    if (p.hash_kind == GGCAT_GPU_HASH_RK128) {
        // start terms of the rolling hash per base position and base
        if (ctx->rk_pos_k != p.k) {
            std::vector<u128_t> tab((size_t)8 * (p.k - 1));
            for (uint32_t i = 0; i + 1 < p.k; i++) {
                u128_t pf = 1, pb = M.rk_mult;  // M^(k-2-i), M^(i+1)
                for (uint32_t e = 0; e + 2 + i < p.k; e++) pf *= M.rk_mult;
                for (uint32_t e = 0; e < i; e++) pb *= M.rk_mult;
                for (uint32_t c = 0; c < 4; c++) {
                    tab[2 * (4 * i + c)] = M.rk_fwd[c] * pf;
                    tab[2 * (4 * i + c) + 1] = M.rk_bkw[c] * pb;
                }
            }
            ctx->rk_pos_k = 0;
            ENSURE(ctx->rk_pos, tab.size() * sizeof(u128_t));
            HIPCHK(hipMemcpy(ctx->rk_pos.p, tab.data(), tab.size() * sizeof(u128_t), hipMemcpyHostToDevice));
            ctx->rk_pos_k = p.k;
        }
        M.rk_pos = P_<u128_t>(ctx->rk_pos);
    }

    // work items: runs of units of one sub-bucket
    ENSURE(ctx->items, (size_t)ctx->n_units * sizeof(WorkItem));
    ENSURE(ctx->n_items, 8);
    ENSURE(ctx->work, 8);
    ENSURE(ctx->redo_count, 8);
    ENSURE(ctx->chunk_cursor, 8);
    ENSURE(ctx->fix_cursor, 8);
    ENSURE(ctx->error_flag, 8);
    M.error_flag = P_<uint32_t>(ctx->error_flag);
    HIPCHK(hipEventRecord(ctx->ev[5], ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->n_items.p, 0, 8, ctx->stream));
    hipLaunchKernelGGL(make_items_kernel, dim3((ctx->n_sub_all + 255) / 256), dim3(256), 0, ctx->stream,
                       P_<uint32_t>(ctx->unit_skmers), P_<uint64_t>(ctx->unit_off), ctx->n_sub_all, ctx->xb, item_occ,
                       ctx->dd_max, P_<WorkItem>(ctx->items), P_<unsigned int>(ctx->n_items));
    HIPCHK(hipGetLastError());
    uint32_t n_items = 0;
    HIPCHK(hipMemcpyAsync(&n_items, ctx->n_items.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    M.items = P_<WorkItem>(ctx->items);

    // resident workgroups per CU: what the 160 KiB LDS holds with the
    // kernel's static LDS, at most 8 (32 waves of 64)
    // the seq-hash k <= 32 dedup path runs a variant with its switches compiled in,
    const bool mode1 = key_words == 1 && W == 0 && M.dedup && M.tag_kmers && ctx->slot_words == 4;
    // and the uncolored rabin-karp128 path another
    const bool mode2 = key_words == 2 && W == 0 && !M.tag_kmers && ctx->slot_words == 8 &&
                       p.hash_kind == GGCAT_GPU_HASH_RK128;
    void (*merge_kernel)(MergeArgs) =
        mode1 ? merge_lds_seq64_kernel : mode2 ? merge_lds_rk128_kernel : merge_lds_kernel;
    hipFuncAttributes fa{};
    HIPCHK(hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(merge_kernel)));
    const size_t wg_lds = (lds_bytes + fa.sharedSizeBytes + 511) / 512 * 512;  // allocation granule
    const uint32_t fit =
        (uint32_t)std::max<size_t>(1, std::min<size_t>(2048 / merge_threads, (160 * 1024) / wg_lds));
    const uint32_t lds_wgs = ctx->n_cu * (ctx->tune.merge_wgs_per_cu ? ctx->tune.merge_wgs_per_cu : fit);
    // every wavefront keeps one partly filled chunk open
    uint64_t n_chunks = std::max<uint64_t>((ctx->kmer_occ / 6) / OUT_CHUNK + 8 * lds_wgs + 4096 + 64,
                                           ctx->out_val.cap / 4 / OUT_CHUNK);
    uint64_t out_cap = n_chunks * OUT_CHUNK;
    uint64_t fix_cap = std::max<uint64_t>(ctx->n_records / 64 + 4096, ctx->fix_val.cap / 4);
    for (int attempt = 0; attempt < 2; attempt++) {
        out_cap = n_chunks * OUT_CHUNK;
        ENSURE(ctx->out_key, out_cap * 8 * key_words);
        ENSURE(ctx->out_val, out_cap * 4);
        ENSURE(ctx->out_sub, out_cap * 4);
        ENSURE(ctx->chunk_fill_dev, n_chunks * 4);
        ENSURE(ctx->fix_key, fix_cap * 8 * key_words);
        ENSURE(ctx->fix_val, fix_cap * 4);
        ENSURE(ctx->fix_sub, fix_cap * 4);
        if (W) {
            ENSURE(ctx->out_colors, out_cap * 4 * W);
            ENSURE(ctx->fix_colors, fix_cap * 4 * W);
        }
        M.out_colors = W ? P_<uint32_t>(ctx->out_colors) : nullptr;
        M.fix_colors = W ? P_<uint32_t>(ctx->fix_colors) : nullptr;
        M.work = P_<unsigned int>(ctx->work);
        M.redo_count = P_<unsigned int>(ctx->redo_count);
        M.out_key = P_<uint64_t>(ctx->out_key);
        M.out_val = P_<uint32_t>(ctx->out_val);
        M.out_sub = P_<uint32_t>(ctx->out_sub);
        M.chunk_cursor = P_<unsigned int>(ctx->chunk_cursor);
        M.chunk_fill = P_<uint32_t>(ctx->chunk_fill_dev);
        M.n_chunks = (uint32_t)n_chunks;
        M.fix_key = P_<uint64_t>(ctx->fix_key);
        M.fix_val = P_<uint32_t>(ctx->fix_val);
        M.fix_sub = P_<uint32_t>(ctx->fix_sub);
        M.fix_cursor = P_<unsigned long long>(ctx->fix_cursor);
        M.fix_cap = fix_cap;
        HIPCHK(hipMemsetAsync(ctx->chunk_cursor.p, 0, 8, ctx->stream));
        HIPCHK(hipMemsetAsync(ctx->fix_cursor.p, 0, 8, ctx->stream));
        HIPCHK(hipMemsetAsync(ctx->error_flag.p, 0, 8, ctx->stream));
        HIPCHK(hipEventRecord(ctx->ev[6], ctx->stream));
        HIPCHK(hipEventRecord(ctx->ev[7], ctx->stream));
        // Pass 0: one work entry per item.  Entries whose table overflows (or
        // items above lds_try) come back in `redo` and are processed again as
        // 2^L hash partitions, refined x4 while a partition still overflows.
        std::vector<PartWork> next;
        uint32_t n_work = n_items, n_split_items = 0, n_passes_run = 0;
        double ms_parts = 0;
        for (uint32_t pass = 0; n_work; pass++) {
            ENSURE(ctx->redo, (size_t)n_work * sizeof(PartWork));
            M.redo = P_<PartWork>(ctx->redo);
            M.n_work = n_work;
            // batches of up to 8 entries per claim, while every workgroup
            // still gets about 32 claims
            M.claim = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(8, n_work / ((uint64_t)lds_wgs * 32)));
            M.parts = pass ? P_<PartWork>(ctx->parts) : nullptr;
            HIPCHK(hipMemsetAsync(ctx->work.p, 0, 8, ctx->stream));
            HIPCHK(hipMemsetAsync(ctx->redo_count.p, 0, 8, ctx->stream));
            if (pass) HIPCHK(hipEventRecord(ctx->ev[8], ctx->stream));
            hipLaunchKernelGGL(merge_kernel, dim3(lds_wgs), dim3(merge_threads), lds_bytes, ctx->stream, M);
            HIPCHK(hipGetLastError());
            n_passes_run++;
            HIPCHK(hipEventRecord(ctx->ev[pass ? 9 : 7], ctx->stream));
            uint32_t n_redo = 0;
            HIPCHK(hipMemcpyAsync(&n_redo, ctx->redo_count.p, 4, hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));
            if (pass) {
                float tp;
                HIPCHK(hipEventElapsedTime(&tp, ctx->ev[8], ctx->ev[9]));
                ms_parts += tp;
            }
            if (n_redo == 0) break;
            std::vector<PartWork> redo(n_redo);
            HIPCHK(hipMemcpy(redo.data(), ctx->redo.p, (size_t)n_redo * sizeof(PartWork), hipMemcpyDeviceToHost));
            next.clear();
            for (const PartWork& w : redo) {
                if (w.item >= n_items) return fail(ctx, GGCAT_GPU_EHIP, "corrupt merge work list");
                uint32_t L, base, stride;
                if (w.log_parts == 0) {
                    // about half a table of occurrences per part
                    const uint64_t per = ctx->tune.part_occ ? ctx->tune.part_occ : M.lds_cap / 2;
                    const uint64_t want = (w.occ + per - 1) / per;
                    L = 1;
                    while ((1ull << L) < want) L++;
                    base = 0;
                    stride = 1;
                    n_split_items++;
                } else {
                    L = w.log_parts + 2;
                    base = w.part;
                    stride = 1u << w.log_parts;
                }
                if (L > MAX_LOG_PARTS)
                    return fail(ctx, GGCAT_GPU_EHIP, "k-mer table overflow: a work item needs more than 2^16 parts");
                const uint32_t n_new = w.log_parts == 0 ? (1u << L) : 4u;
                for (uint32_t j = 0; j < n_new; j++) next.push_back(PartWork{w.item, base + j * stride, L, w.occ});
            }
            n_work = (uint32_t)next.size();
            ENSURE(ctx->parts, (size_t)n_work * sizeof(PartWork));
            HIPCHK(hipMemcpy(ctx->parts.p, next.data(), (size_t)n_work * sizeof(PartWork), hipMemcpyHostToDevice));
        }
        ctx->n_big = n_split_items;
        ctx->n_passes = n_passes_run;
        uint64_t n_fix = 0;
        uint32_t used_chunks = 0, errf = 0;
        HIPCHK(hipMemcpyAsync(&used_chunks, ctx->chunk_cursor.p, 4, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(&n_fix, ctx->fix_cursor.p, 8, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipMemcpyAsync(&errf, ctx->error_flag.p, 4, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        if (errf & 1) return fail(ctx, GGCAT_GPU_EHIP, "k-mer table overflow");
        if (errf & 2) return fail(ctx, GGCAT_GPU_EHIP, "malformed record slot");
        float t;
        HIPCHK(hipEventElapsedTime(&t, ctx->ev[6], ctx->ev[7]));
        ctx->ms[3] = t;
        ctx->ms[4] = ms_parts;
        if (used_chunks > n_chunks || n_fix > fix_cap) {
            if (attempt == 1) return fail(ctx, GGCAT_GPU_ENOMEM, "solid output overflow");
            n_chunks = std::max<uint64_t>(n_chunks, (uint64_t)used_chunks + 8 * lds_wgs);
            fix_cap = std::max(fix_cap, n_fix + 1024);
            continue;
        }
        ctx->chunk_fill.resize(used_chunks);
        if (used_chunks)
            HIPCHK(hipMemcpy(ctx->chunk_fill.data(), ctx->chunk_fill_dev.p, (size_t)used_chunks * 4, hipMemcpyDeviceToHost));
        uint64_t n_out = 0;
        for (uint32_t f : ctx->chunk_fill) n_out += f;
        // combined fixups are appended after the last used chunk
        const uint64_t tail0 = (uint64_t)used_chunks * OUT_CHUNK;
        uint64_t n_add = 0;
        // combine the two halves of cross-unit boundary k-mers
        if (n_fix) {
            const uint32_t kw = key_words;
            std::vector<uint64_t> fk(n_fix * kw);
            std::vector<uint32_t> fv(n_fix), fs(n_fix), fc(n_fix * W);
            if (W) HIPCHK(hipMemcpy(fc.data(), ctx->fix_colors.p, fc.size() * 4, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(fk.data(), ctx->fix_key.p, fk.size() * 8, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(fv.data(), ctx->fix_val.p, n_fix * 4, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(fs.data(), ctx->fix_sub.p, n_fix * 4, hipMemcpyDeviceToHost));
            std::vector<uint32_t> ord(n_fix);
            for (uint32_t i = 0; i < n_fix; i++) ord[i] = i;
            auto key_of = [&](uint32_t i) {
                return std::make_tuple(fs[i], kw == 2 ? fk[2 * i + 1] : 0ull, fk[kw * i]);
            };
            std::sort(ord.begin(), ord.end(), [&](uint32_t a2, uint32_t b2) { return key_of(a2) < key_of(b2); });
            std::vector<uint64_t> ok_;
            std::vector<uint32_t> ov, os, oc;
            for (uint64_t i = 0; i < n_fix;) {
                uint64_t j = i + 1;
                const uint32_t a2 = ord[i];
                uint32_t v = fv[a2];
                const size_t c0 = oc.size();
                for (uint32_t q = 0; q < W; q++) oc.push_back(fc[(size_t)a2 * W + q]);
                while (j < n_fix && key_of(ord[j]) == key_of(a2)) {
                    const uint32_t b2 = ord[j++];
                    v |= fv[b2] & 0xC0000000u;
                    for (uint32_t q = 0; q < W; q++) oc[c0 + q] |= fc[(size_t)b2 * W + q];
                }
                for (uint32_t q = 0; q < kw; q++) ok_.push_back(fk[kw * a2 + q]);
                ov.push_back(v);
                os.push_back(fs[a2]);
                i = j;
            }
            n_add = ov.size();
            if (tail0 + n_add > out_cap) {
                if (attempt == 1) return fail(ctx, GGCAT_GPU_ENOMEM, "solid output overflow");
                n_chunks = (tail0 + n_add) / OUT_CHUNK + 64;
                continue;
            }
            HIPCHK(hipMemcpy(P_<uint64_t>(ctx->out_key) + tail0 * kw, ok_.data(), n_add * kw * 8, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(P_<uint32_t>(ctx->out_val) + tail0, ov.data(), n_add * 4, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(P_<uint32_t>(ctx->out_sub) + tail0, os.data(), n_add * 4, hipMemcpyHostToDevice));
            if (W)
                HIPCHK(hipMemcpy(P_<uint32_t>(ctx->out_colors) + tail0 * W, oc.data(), n_add * 4 * W,
                                 hipMemcpyHostToDevice));
        }
        ctx->n_chunks_used = used_chunks;
        ctx->n_tail = n_add;
        n_out += n_add;
        ctx->n_fix = n_fix;
        ctx->n_solid = n_out;
        HIPCHK(hipEventRecord(ctx->ev[10], ctx->stream));
        HIPCHK(hipEventSynchronize(ctx->ev[10]));
        HIPCHK(hipEventElapsedTime(&t, ctx->ev[5], ctx->ev[10]));
        ctx->ms[6] = t;
        ctx->have_solid = true;
        return GGCAT_GPU_OK;
    }
    return fail(ctx, GGCAT_GPU_ENOMEM, "solid output overflow");
}

int ggcat_gpu_solid_info(ggcat_gpu_ctx* ctx, uint64_t* n_solid, uint32_t* key_words) {
    if (!ctx) return GGCAT_GPU_EINVAL;
    if (!ctx->have_solid) return fail(ctx, GGCAT_GPU_ESTATE, "no merge result");
    if (n_solid) *n_solid = ctx->n_solid;
    if (key_words) *key_words = ctx->key_words;
    return GGCAT_GPU_OK;
}

int ggcat_gpu_solid_copy(ggcat_gpu_ctx* ctx, uint64_t* keys, uint32_t* multiplicity, uint8_t* flags, uint32_t* sub) {
    if (!ctx) return GGCAT_GPU_EINVAL;
    if (!ctx->have_solid) return fail(ctx, GGCAT_GPU_ESTATE, "no merge result");
    HIPCHK(hipSetDevice(ctx->device));
    const uint32_t kw = ctx->key_words;
    // the output is the valid prefix of every used chunk, then the tail of
    // combined cross-unit entries
    std::vector<std::pair<uint64_t, uint64_t>> runs;  // (first entry, count)
    for (uint32_t c = 0; c < ctx->n_chunks_used; c++)
        if (ctx->chunk_fill[c]) runs.emplace_back((uint64_t)c * OUT_CHUNK, ctx->chunk_fill[c]);
    if (ctx->n_tail) runs.emplace_back((uint64_t)ctx->n_chunks_used * OUT_CHUNK, ctx->n_tail);
    std::vector<uint32_t> v;
    uint64_t o = 0;
    for (auto& r : runs) {
        if (keys)
            HIPCHK(hipMemcpy(keys + o * kw, P_<uint64_t>(ctx->out_key) + r.first * kw, r.second * 8 * kw,
                             hipMemcpyDeviceToHost));
        if (sub) HIPCHK(hipMemcpy(sub + o, P_<uint32_t>(ctx->out_sub) + r.first, r.second * 4, hipMemcpyDeviceToHost));
        if (multiplicity || flags) {
            v.resize(r.second);
            HIPCHK(hipMemcpy(v.data(), P_<uint32_t>(ctx->out_val) + r.first, r.second * 4, hipMemcpyDeviceToHost));
            for (uint64_t i = 0; i < r.second; i++) {
                if (multiplicity) multiplicity[o + i] = v[i] & 0x3FFFFFFFu;
                if (flags) flags[o + i] = (uint8_t)(v[i] >> 30);
            }
        }
        o += r.second;
    }
    return GGCAT_GPU_OK;
}

int ggcat_gpu_solid_colors(ggcat_gpu_ctx* ctx, uint32_t* color_words, uint32_t* bitmaps) {
    if (!ctx) return GGCAT_GPU_EINVAL;
    if (!ctx->have_solid) return fail(ctx, GGCAT_GPU_ESTATE, "no merge result");
    if (!ctx->colored) return fail(ctx, GGCAT_GPU_ESTATE, "the last build was not colored");
    const uint32_t W = ctx->color_words;
    if (color_words) *color_words = W;
    if (!bitmaps) return GGCAT_GPU_OK;
    HIPCHK(hipSetDevice(ctx->device));
    uint64_t o = 0;
    auto copy = [&](uint64_t first, uint64_t count) -> int {
        HIPCHK(hipMemcpy(bitmaps + o * W, P_<uint32_t>(ctx->out_colors) + first * W, count * 4 * W,
                         hipMemcpyDeviceToHost));
        o += count;
        return GGCAT_GPU_OK;
    };
    for (uint32_t c = 0; c < ctx->n_chunks_used; c++)
        if (ctx->chunk_fill[c] && copy((uint64_t)c * OUT_CHUNK, ctx->chunk_fill[c]) != GGCAT_GPU_OK) return GGCAT_GPU_EHIP;
    if (ctx->n_tail && copy((uint64_t)ctx->n_chunks_used * OUT_CHUNK, ctx->n_tail) != GGCAT_GPU_OK)
        return GGCAT_GPU_EHIP;
    return GGCAT_GPU_OK;
}

int ggcat_gpu_set_tuning(ggcat_gpu_ctx* ctx, const ggcat_gpu_tuning* t) {
    if (!ctx || !t) return GGCAT_GPU_EINVAL;
    if (t->merge_wgs_per_cu > 64) return fail(ctx, GGCAT_GPU_EINVAL, "merge_wgs_per_cu must be <= 64");
    if (t->merge_threads && t->merge_threads != 128 && t->merge_threads != 256)
        return fail(ctx, GGCAT_GPU_EINVAL, "merge_threads must be 0, 128 or 256");
    if (t->rec_cap && (t->rec_cap < 16 || t->rec_cap > 512 || (t->rec_cap & (t->rec_cap - 1))))
        return fail(ctx, GGCAT_GPU_EINVAL, "rec_cap must be 0 or a power of two in [16, 512]");
    if (t->read_block && t->read_block < 256) return fail(ctx, GGCAT_GPU_EINVAL, "read_block must be 0 or >= 256");
    if (t->upload_chunk && t->upload_chunk < 4096) return fail(ctx, GGCAT_GPU_EINVAL, "upload_chunk must be 0 or >= 4096");
    ctx->tune = *t;
    return GGCAT_GPU_OK;
}

int ggcat_gpu_timings(ggcat_gpu_ctx* ctx, double* ms, uint32_t n) {
    if (!ctx || !ms) return GGCAT_GPU_EINVAL;
    for (uint32_t i = 0; i < n && i < 9; i++) ms[i] = ctx->ms[i];
    return GGCAT_GPU_OK;
}

int ggcat_gpu_counters(ggcat_gpu_ctx* ctx, uint64_t* out, uint32_t n) {
    if (!ctx || !out) return GGCAT_GPU_EINVAL;
    uint64_t v[11] = {ctx->n_input, ctx->n_records, ctx->kmer_occ + ctx->n_records * (ctx->P.k - 1), ctx->kmer_occ,
                      ctx->n_big,   ctx->n_solid,   ctx->n_fix,   ctx->n_passes, ctx->input_path, ctx->flush_cuts,
                      ctx->n_stored};
    for (uint32_t i = 0; i < n && i < 11; i++) out[i] = v[i];
    return GGCAT_GPU_OK;
}

int ggcat_gpu_device_alloc(ggcat_gpu_ctx* ctx, uint64_t bytes, void** out) {
    if (!ctx || !out) return GGCAT_GPU_EINVAL;
    *out = nullptr;
    HIPCHK(hipSetDevice(ctx->device));
    hipError_t e = hipMalloc(out, bytes ? bytes : 16);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *out = nullptr;
        return fail(ctx, GGCAT_GPU_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    return GGCAT_GPU_OK;
}

int ggcat_gpu_device_free(ggcat_gpu_ctx* ctx, void* p) {
    if (!ctx) return GGCAT_GPU_EINVAL;
    HIPCHK(hipSetDevice(ctx->device));
    if (p) HIPCHK(hipFree(p));
    return GGCAT_GPU_OK;
}

// Valid solid positions that no path walk visited (nodes on cycles).
__global__ void unvisited_kernel(UnitigArgs A, uint64_t n_pos, uint32_t* rest, unsigned int* n_rest, uint32_t cap) {
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < n_pos; p += (uint64_t)gridDim.x * blockDim.x)
        if (unitig_valid(A, p) && !A.visited[p]) {
            const uint32_t i = atomicAdd(n_rest, 1u);
            if (i < cap) rest[i] = (uint32_t)p;
        }
}

// Plain device-to-device copy (16-byte vectors, grid-stride).
__global__ void copy16_kernel(uint4* __restrict__ dst, const uint4* __restrict__ src, uint64_t n16) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int ggcat_gpu_copy_device(ggcat_gpu_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return GGCAT_GPU_EINVAL;
    if ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src) | bytes) & 15)
        return fail(ctx, GGCAT_GPU_EINVAL, "device copy needs 16-byte aligned pointers and size");
    HIPCHK(hipSetDevice(ctx->device));
    if (bytes) {
        hipLaunchKernelGGL(copy16_kernel, dim3(ctx->n_cu * 8), dim3(256), 0, ctx->stream, reinterpret_cast<uint4*>(dst),
                           reinterpret_cast<const uint4*>(src), bytes / 16);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return GGCAT_GPU_OK;
}

int ggcat_gpu_copy_to_device(ggcat_gpu_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return GGCAT_GPU_EINVAL;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return GGCAT_GPU_OK;
}

int ggcat_gpu_copy_to_host(ggcat_gpu_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return GGCAT_GPU_EINVAL;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return GGCAT_GPU_OK;
}

int ggcat_gpu_read_files(ggcat_gpu_ctx* ctx, const char* const* paths, uint32_t n_paths, uint32_t k, int colors,
                         uint8_t** joined, uint64_t* n_bytes, uint64_t** color_starts, uint32_t** color_ids,
                         uint32_t* n_colors, uint64_t* estimated_bases) {
    if ((!paths && n_paths) || !joined || !n_bytes) return fail(ctx, GGCAT_GPU_EINVAL, "null argument");
    std::vector<std::string> ps;
    for (uint32_t i = 0; i < n_paths; i++) {
        if (!paths[i]) return fail(ctx, GGCAT_GPU_EINVAL, "null path");
        ps.emplace_back(paths[i]);
    }
    gg_host::ReadResult R;
    if (!gg_host::read_files(ps, k, colors != 0, R, ctx ? ctx->tune.read_block : 0))
        return fail(ctx, GGCAT_GPU_EINVAL, R.error);
    *joined = (uint8_t*)malloc(R.joined.size() + 16);
    if (!*joined) return fail(ctx, GGCAT_GPU_ENOMEM, "host allocation failed");
    memcpy(*joined, R.joined.data(), R.joined.size());
    *n_bytes = R.joined.size();
    if (estimated_bases) *estimated_bases = R.estimated_bases;
    if (n_colors) *n_colors = (uint32_t)R.colors.size();
    if (color_starts) {
        *color_starts = (uint64_t*)malloc(R.color_starts.size() * 8 + 8);
        memcpy(*color_starts, R.color_starts.data(), R.color_starts.size() * 8);
    }
    if (color_ids) {
        *color_ids = (uint32_t*)malloc(R.colors.size() * 4 + 4);
        memcpy(*color_ids, R.colors.data(), R.colors.size() * 4);
    }
    return GGCAT_GPU_OK;
}

void ggcat_gpu_free_host(void* p) { free(p); }

int ggcat_gpu_plan_build(const ggcat_gpu_build_options* o, uint64_t estimated_bases, ggcat_gpu_params* out) {
    if (!o || !out || o->k < 2 || (o->m == 0 && o->k < 4)) return GGCAT_GPU_EINVAL;
    ggcat_gpu_params p{};
    p.k = o->k;
    p.m = o->m ? o->m : gg_host::compute_best_m(o->k);
    uint32_t b_auto = 0, sb_auto = 0;
    gg_host::auto_bucket_logs(estimated_bases, &b_auto, &sb_auto);
    p.buckets_log = o->buckets_log ? o->buckets_log : b_auto;
    p.second_buckets_log = sb_auto;
    p.forward_only = o->forward_only;
    p.hash_kind = o->hash_kind;
    p.min_multiplicity = o->min_multiplicity;
    *out = p;
    return GGCAT_GPU_OK;
}

namespace {


// Plain FASTA / FASTQ files read straight into HBM (parallel reads into a
// pinned two-buffer ring, each upload overlapping the next read) and parsed
// there (parse.hip) into the joined layout at ctx->input; every file's parse
// is queued behind its upload, so it overlaps the reads of the files after
// it.  Returns GGCAT_GPU_EAGAIN when the files need the host reader:
// compressed, not a regular file, or a layout parse.hip leaves to it.
static int device_read_files(ggcat_gpu_ctx* ctx, const std::vector<std::string>& paths, uint32_t k, bool colors,
                      uint64_t* n_joined, std::vector<uint64_t>& color_starts, std::vector<uint32_t>& color_ids,
                      uint64_t* est_bases) {
    const size_t nf = paths.size();
    std::vector<uint64_t> sizes(nf);
    std::vector<int> types(nf);
    uint64_t total = 0, max_chunks = 1;
    *est_bases = 0;
    for (size_t i = 0; i < nf; i++) {
        types[i] = gg_host::file_type(paths[i]);
        if (types[i] < 0) return fail(ctx, GGCAT_GPU_EINVAL, "cannot recognize file type of '" + paths[i] + "'");
        const std::string ext = gg_host::last_ext(paths[i]);
        if (ext == "gz" || ext == "lz4" || ext == "bz2" || ext == "xz" || ext == "zst" || ext == "zstd")
            return GGCAT_GPU_EAGAIN;
        struct stat st;
        if (stat(paths[i].c_str(), &st) != 0 || !S_ISREG(st.st_mode)) return GGCAT_GPU_EAGAIN;
        sizes[i] = (uint64_t)st.st_size;
        total += sizes[i];
        *est_bases += sizes[i];  // get_estimated_bases_count of an uncompressed file
        max_chunks = std::max<uint64_t>(max_chunks, (sizes[i] + PARSE_CHUNK - 1) / PARSE_CHUNK);
    }
    if (max_chunks >= 0xFFFFFFFFull) return GGCAT_GPU_EAGAIN;
    const uint64_t block = ctx->tune.read_block ? ctx->tune.read_block : (uint64_t)gg_host::READ_BLOCK;
    const uint64_t copyback = k > 1 ? k - 1 : 0;
    const uint64_t flush = std::max<uint64_t>(block, 2 * copyback);
    const uint64_t out_cap = total + nf + 16;  // every file's bytes and its closing '\n'
    const uint64_t bit_words = out_cap / 64 + 2;
    HIPCHK(hipSetDevice(ctx->device));
    ENSURE(ctx->raw, total + 16);
    ENSURE(ctx->input, out_cap);
    // counters: running output offset, then per file its flags and output start
    ENSURE(ctx->p_flags, 8 * (1 + 2 * nf));
    ENSURE(ctx->p_bits, bit_words * 8);
    ENSURE(ctx->p_stats, max_chunks * sizeof(ChunkStats));
    ENSURE(ctx->p_carry, max_chunks * sizeof(ChunkCarry));
    ENSURE(ctx->p_keep, max_chunks * 4);
    ENSURE(ctx->p_keep_off, (max_chunks + 1) * 8);
    if (!ctx->pin_ring) {
        if (hipHostMalloc(&ctx->pin_ring, 2 * PIN_CHUNK, hipHostMallocDefault) != hipSuccess) {
            ctx->pin_ring = nullptr;
            return GGCAT_GPU_EAGAIN;
        }
    }
    uint64_t* d_base = P_<uint64_t>(ctx->p_flags);
    unsigned int* d_flags = reinterpret_cast<unsigned int*>(d_base + 1);  // nf words (2 per u64 slot)
    uint64_t* d_file_start = d_base + 1 + nf;
    const uint32_t zgrid = (uint32_t)std::min<uint64_t>((bit_words + 255) / 256, (uint64_t)ctx->n_cu * 16);
    hipLaunchKernelGGL(parse_zero_kernel, dim3(1), dim3(256), 0, ctx->stream, P_<unsigned long long>(ctx->p_flags),
                       (uint64_t)(1 + 2 * nf));
    hipLaunchKernelGGL(parse_zero_kernel, dim3(zgrid), dim3(256), 0, ctx->stream, P_<unsigned long long>(ctx->p_bits),
                       bit_words);
    HIPCHK(hipGetLastError());
    uint8_t* ring = static_cast<uint8_t*>(ctx->pin_ring);
    uint8_t* d_raw = P_<uint8_t>(ctx->raw);
    bool pending[2] = {false, false};
    uint64_t file_off = 0;
    uint32_t j = 0;
    const uint32_t n_threads = std::max(1u, std::min(HOST_COPY_THREADS, std::thread::hardware_concurrency()));
    for (size_t i = 0; i < nf; i++) {
        // upload: chunk j of the concatenated files goes through ring buffer j % 2
        const int fd = open(paths[i].c_str(), O_RDONLY);
        if (fd < 0) return fail(ctx, GGCAT_GPU_EINVAL, "cannot open '" + paths[i] + "'");
        bool ok = true;
        for (uint64_t o = 0; o < sizes[i] && ok; o += PIN_CHUNK, j++) {
            const uint32_t b = j & 1;
            if (pending[b]) HIPCHK(hipEventSynchronize(ctx->ev[13 + b]));
            const uint64_t len = std::min<uint64_t>(PIN_CHUNK, sizes[i] - o);
            const uint64_t slice = (len + n_threads - 1) / n_threads;
            std::vector<char> good(n_threads, 1);
            std::vector<std::thread> th;
            for (uint32_t q = 0; q < n_threads; q++) {
                th.emplace_back([&, q] {
                    const uint64_t hi = std::min(len, (q + 1) * slice);
                    for (uint64_t a = q * slice; a < hi;) {
                        const ssize_t r = pread(fd, ring + b * PIN_CHUNK + a, hi - a, (off_t)(o + a));
                        if (r <= 0) {
                            good[q] = 0;
                            return;
                        }
                        a += (uint64_t)r;
                    }
                });
            }
            for (auto& x : th) x.join();
            for (char g : good) ok = ok && g;
            if (!ok) break;
            HIPCHK(hipMemcpyAsync(d_raw + file_off + o, ring + b * PIN_CHUNK, len, hipMemcpyHostToDevice, ctx->stream));
            HIPCHK(hipEventRecord(ctx->ev[13 + b], ctx->stream));
            pending[b] = true;
        }
        close(fd);
        if (!ok) return fail(ctx, GGCAT_GPU_EINVAL, "error while reading '" + paths[i] + "'");
        // parse, queued behind the upload
        ParseArgs A{};
        A.raw = d_raw + file_off;
        A.n = sizes[i];
        A.fastq = types[i] == 1 ? 1u : 0u;
        A.n_chunks = (sizes[i] + PARSE_CHUNK - 1) / PARSE_CHUNK;
        A.block = block;
        A.flush = flush;
        A.stats = P_<ChunkStats>(ctx->p_stats);
        A.carry = P_<ChunkCarry>(ctx->p_carry);
        A.keep = P_<uint32_t>(ctx->p_keep);
        A.keep_off = P_<uint64_t>(ctx->p_keep_off);
        A.out = P_<uint8_t>(ctx->input);
        A.base = d_base;
        A.bits = P_<unsigned long long>(ctx->p_bits);
        A.flags = d_flags + i;
        if (A.n_chunks) {
            const uint32_t grid = (uint32_t)std::min<uint64_t>(A.n_chunks, (uint64_t)ctx->n_cu * 8);
            hipLaunchKernelGGL(parse_stats_kernel, dim3(grid), dim3(256), 0, ctx->stream, A);
            hipLaunchKernelGGL(parse_carry_kernel, dim3(1), dim3(1024), 0, ctx->stream, A);
            hipLaunchKernelGGL(parse_emit_kernel<false>, dim3(grid), dim3(256), 0, ctx->stream, A);
            HIPCHK(hipGetLastError());
            const int r = device_scan(ctx, A.keep, P_<uint64_t>(ctx->p_keep_off), (uint32_t)A.n_chunks);
            if (r != GGCAT_GPU_OK) return r;
            hipLaunchKernelGGL(parse_emit_kernel<true>, dim3(grid), dim3(256), 0, ctx->stream, A);
        }
        hipLaunchKernelGGL(parse_finish_kernel, dim3(1), dim3(64), 0, ctx->stream, A, d_file_start + i);
        HIPCHK(hipGetLastError());
        file_off += sizes[i];
    }
    std::vector<uint64_t> counters(1 + 2 * nf);
    HIPCHK(hipMemcpyAsync(counters.data(), d_base, counters.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    uint64_t n_out = counters[0];
    const unsigned int* flags = reinterpret_cast<const unsigned int*>(counters.data() + 1);
    std::vector<uint64_t> starts(counters.begin() + 1 + nf, counters.end());
    unsigned int any = 0;
    for (size_t i = 0; i < nf; i++) any |= flags[i];
    if (any & 1u) return GGCAT_GPU_EAGAIN;
    if (any & 2u) {
        // FASTA records of F bytes or more: find them from per-window newline
        // summaries, walk their flush points, insert the cuts
        const uint64_t window = std::min<uint64_t>(1ull << 20, flush);
        const uint64_t n_win = (n_out + window - 1) / window;
        ENSURE(ctx->p_stats, n_win * 16);
        unsigned long long* d_first = P_<unsigned long long>(ctx->p_stats);
        hipLaunchKernelGGL(newline_summary_kernel, dim3((uint32_t)n_win), dim3(256), 0, ctx->stream,
                           P_<uint8_t>(ctx->input), n_out, window, d_first, d_first + n_win);
        HIPCHK(hipGetLastError());
        std::vector<uint64_t> fl(2 * n_win);
        HIPCHK(hipMemcpyAsync(fl.data(), d_first, fl.size() * 8, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        std::vector<uint64_t> rs, rl;
        uint64_t prev_end = 0;  // start of the record after the last '\n' seen
        for (uint64_t w = 0; w < n_win; w++) {
            if (!fl[w]) continue;
            const uint64_t nl = fl[w] - 1;
            if (nl - prev_end >= flush) {
                rs.push_back(prev_end);
                rl.push_back(nl - prev_end);
            }
            prev_end = fl[n_win + w];  // last '\n' + 1
        }
        if (!rs.empty()) {
            // each cut advances a record by at least F - (k - 1) >= F / 2
            const uint64_t cap = 2 * n_out / flush + rs.size() + 16;
            if (cap >= 0xFFFFFFFFull) return GGCAT_GPU_EAGAIN;
            ENSURE(ctx->p_carry, (2 * rs.size() + cap) * 8);
            ENSURE(ctx->p_keep, 8);
            uint64_t* d_rs = P_<uint64_t>(ctx->p_carry);
            uint64_t* d_cuts = d_rs + 2 * rs.size();
            unsigned int* d_n = P_<unsigned int>(ctx->p_keep);
            HIPCHK(hipMemcpyAsync(d_rs, rs.data(), rs.size() * 8, hipMemcpyHostToDevice, ctx->stream));
            HIPCHK(hipMemcpyAsync(d_rs + rs.size(), rl.data(), rl.size() * 8, hipMemcpyHostToDevice, ctx->stream));
            hipLaunchKernelGGL(parse_zero_kernel, dim3(1), dim3(64), 0, ctx->stream,
                               reinterpret_cast<unsigned long long*>(d_n), (uint64_t)1);
            hipLaunchKernelGGL(flush_split_kernel, dim3((uint32_t)std::min<size_t>(rs.size(), 65535)), dim3(64), 0,
                               ctx->stream, d_rs, d_rs + rs.size(), (uint32_t)rs.size(),
                               P_<const unsigned long long>(ctx->p_bits), flush, copyback, d_cuts, d_n,
                               (uint32_t)cap);
            HIPCHK(hipGetLastError());
            unsigned int n_cuts = 0;
            HIPCHK(hipMemcpyAsync(&n_cuts, d_n, 4, hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));
            if (n_cuts > cap) return fail(ctx, GGCAT_GPU_EHIP, "flush cut list overflow");
            std::vector<uint64_t> cuts(n_cuts);
            HIPCHK(hipMemcpy(cuts.data(), d_cuts, (size_t)n_cuts * 8, hipMemcpyDeviceToHost));
            std::sort(cuts.begin(), cuts.end());
            HIPCHK(hipMemcpyAsync(d_cuts, cuts.data(), (size_t)n_cuts * 8, hipMemcpyHostToDevice, ctx->stream));
            const uint64_t grow = (uint64_t)n_cuts * (1 + copyback);
            ENSURE(ctx->raw, n_out + grow + 16);  // the raw files are parsed: reuse their buffer
            const uint32_t egrid = (uint32_t)std::min<uint64_t>((n_out + 4095) / 4096, (uint64_t)ctx->n_cu * 16);
            hipLaunchKernelGGL(flush_expand_kernel, dim3(std::max(1u, egrid)), dim3(256), 0, ctx->stream,
                               P_<uint8_t>(ctx->input), n_out, d_cuts, n_cuts, copyback, P_<uint8_t>(ctx->raw));
            if (n_cuts)
                hipLaunchKernelGGL(flush_insert_kernel, dim3(std::min(n_cuts, 65535u)), dim3(256), 0, ctx->stream,
                                   P_<uint8_t>(ctx->input), d_cuts, n_cuts, copyback, P_<uint8_t>(ctx->raw));
            HIPCHK(hipGetLastError());
            HIPCHK(hipStreamSynchronize(ctx->stream));
            std::swap(ctx->raw, ctx->input);
            for (auto& s : starts) {
                const uint64_t before = (uint64_t)(std::upper_bound(cuts.begin(), cuts.end(), s) - cuts.begin());
                s += before * (1 + copyback);
            }
            n_out += grow;
            ctx->flush_cuts = n_cuts;
        }
    }
    color_starts.clear();
    color_ids.clear();
    if (colors) {
        color_starts = starts;
        for (size_t i = 0; i < nf; i++) color_ids.push_back((uint32_t)i);
    }
    *n_joined = n_out;
    return GGCAT_GPU_OK;
}

}  // namespace

int ggcat_gpu_build_from_files(ggcat_gpu_ctx* ctx, const char* const* paths, uint32_t n_paths,
                               const ggcat_gpu_build_options* options, ggcat_gpu_params* used) {
    if (!ctx || !options) return GGCAT_GPU_EINVAL;
    const ggcat_gpu_build_options o = *options;
    std::vector<std::string> ps;
    for (uint32_t i = 0; i < n_paths; i++) {
        if (!paths || !paths[i]) return fail(ctx, GGCAT_GPU_EINVAL, "null path");
        ps.emplace_back(paths[i]);
    }
    if (ctx->server) return fail(ctx, GGCAT_GPU_ESTATE, "close the query server before a new build");
    const auto t_read = std::chrono::steady_clock::now();
    auto read_ms = [&] {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_read).count();
    };
    ctx->input_path = 0;
    ctx->flush_cuts = 0;
    {
        // plain files: read into HBM and parsed there
        uint64_t n_joined = 0, est = 0;
        std::vector<uint64_t> cs;
        std::vector<uint32_t> ci;
        int r = ctx->tune.host_reader ? GGCAT_GPU_EAGAIN
                                      : device_read_files(ctx, ps, o.k, o.colors != 0, &n_joined, cs, ci, &est);
        if (r == GGCAT_GPU_OK) {
            ctx->ms[8] = read_ms();
            ctx->input_path = 1;
            ggcat_gpu_params p{};
            r = ggcat_gpu_plan_build(&o, est, &p);
            if (r != GGCAT_GPU_OK) return fail(ctx, r, "invalid build options");
            if (used) *used = p;
            r = ggcat_gpu_bucketing(ctx, P_<uint8_t>(ctx->input), n_joined, 1, o.colors ? cs.data() : nullptr,
                                    o.colors ? ci.data() : nullptr, o.colors ? (uint32_t)ci.size() : 0, &p);
            if (r != GGCAT_GPU_OK) return r;
            return ggcat_gpu_kmers_merge(ctx);
        }
        if (r != GGCAT_GPU_EAGAIN) return r;
    }
    gg_host::ReadResult R;
    if (!gg_host::read_files(ps, o.k, o.colors != 0, R, ctx->tune.read_block)) return fail(ctx, GGCAT_GPU_EINVAL, R.error);
    ctx->ms[8] = read_ms();
    ctx->input_path = 2;
    ggcat_gpu_params p{};
    int r = ggcat_gpu_plan_build(&o, R.estimated_bases, &p);
    if (r != GGCAT_GPU_OK) return fail(ctx, r, "invalid build options");
    if (used) *used = p;
    r = ggcat_gpu_bucketing(ctx, R.joined.data(), R.joined.size(), 0, o.colors ? R.color_starts.data() : nullptr,
                                o.colors ? R.colors.data() : nullptr, o.colors ? (uint32_t)R.colors.size() : 0, &p);
    if (r != GGCAT_GPU_OK) return r;
    return ggcat_gpu_kmers_merge(ctx);
}

int ggcat_gpu_index_build(ggcat_gpu_ctx* ctx) {
    if (!ctx) return GGCAT_GPU_EINVAL;
    if (!ctx->have_solid) return fail(ctx, GGCAT_GPU_ESTATE, "no merge result");
    if (ctx->server) return fail(ctx, GGCAT_GPU_ESTATE, "close the query server before rebuilding the index");
    HIPCHK(hipSetDevice(ctx->device));
    const uint32_t kw = ctx->key_words;
    const uint64_t n = ctx->n_solid;
    uint32_t L = 0;
    while (L < 30 && (1ull << (L + 2)) < n) L++;  // about 4 entries per bucket
    const uint32_t nb = 1u << L;
    ENSURE(ctx->idx_bcount, (size_t)nb * 4);
    ENSURE(ctx->idx_bcursor, (size_t)nb * 4);
    ENSURE(ctx->idx_boff, ((size_t)nb + 1) * 8);
    ENSURE(ctx->idx_key, n * 8 * kw + 8);
    ENSURE(ctx->idx_pos, n * 4 + 4);
    if (ctx->n_chunks_used) {
        ENSURE(ctx->chunk_fill_dev, (size_t)ctx->n_chunks_used * 4);
        HIPCHK(hipMemcpy(ctx->chunk_fill_dev.p, ctx->chunk_fill.data(), (size_t)ctx->n_chunks_used * 4,
                         hipMemcpyHostToDevice));
    }
    IndexArgs I{};
    I.out_key = P_<uint64_t>(ctx->out_key);
    I.chunk_fill = P_<uint32_t>(ctx->chunk_fill_dev);
    I.n_chunks = ctx->n_chunks_used;
    I.n_tail = ctx->n_tail;
    I.chunk = OUT_CHUNK;
    I.key_words = kw;
    I.log_buckets = L;
    I.bcount = P_<uint32_t>(ctx->idx_bcount);
    I.boff = P_<uint64_t>(ctx->idx_boff);
    I.bcursor = P_<uint32_t>(ctx->idx_bcursor);
    I.idx_key = P_<uint64_t>(ctx->idx_key);
    I.idx_pos = P_<uint32_t>(ctx->idx_pos);
    const uint64_t n_pos = (uint64_t)ctx->n_chunks_used * OUT_CHUNK + ctx->n_tail;
    if (n_pos >= 0x7FFFFFFFull) return fail(ctx, GGCAT_GPU_EINVAL, "solid table too large for the query index");
    HIPCHK(hipMemsetAsync(ctx->idx_bcount.p, 0, (size_t)nb * 4, ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->idx_bcursor.p, 0, (size_t)nb * 4, ctx->stream));
    const uint32_t grid = (uint32_t)std::min<uint64_t>((n_pos + 255) / 256 + 1, (uint64_t)ctx->n_cu * 16);
    if (n_pos) hipLaunchKernelGGL(index_count_kernel, dim3(grid), dim3(256), 0, ctx->stream, I, n_pos);
    HIPCHK(hipGetLastError());
    {
        int r = device_scan(ctx, I.bcount, P_<uint64_t>(ctx->idx_boff), nb);
        if (r != GGCAT_GPU_OK) return r;
    }
    if (n_pos) hipLaunchKernelGGL(index_fill_kernel, dim3(grid), dim3(256), 0, ctx->stream, I, n_pos);
    HIPCHK(hipGetLastError());
    uint64_t total = 0;
    HIPCHK(hipMemcpyAsync(&total, P_<uint64_t>(ctx->idx_boff) + nb, 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (total != n) return fail(ctx, GGCAT_GPU_EHIP, "query index holds a different number of entries than the solid table");
    ctx->idx_log_buckets = L;
    ctx->have_index = true;
    return GGCAT_GPU_OK;
}

int ggcat_gpu_index_colors(ggcat_gpu_ctx* ctx, uint32_t* n_colors) {
    if (!ctx || !n_colors) return GGCAT_GPU_EINVAL;
    if (!ctx->have_index) return fail(ctx, GGCAT_GPU_ESTATE, "no query index");
    *n_colors = ctx->colored ? ctx->n_color_ids : 0;
    return GGCAT_GPU_OK;
}

int ggcat_gpu_query(ggcat_gpu_ctx* ctx, const uint8_t* queries, uint64_t n_bytes, uint32_t n_queries,
                    uint64_t* query_kmers, uint64_t* matched, uint32_t* color_counts) {
    if (!ctx || (!queries && n_bytes)) return GGCAT_GPU_EINVAL;
    if (!ctx->have_index) return fail(ctx, GGCAT_GPU_ESTATE, "no query index (call ggcat_gpu_index_build)");
    std::vector<uint64_t> off;
    if (!query_offsets(queries, n_bytes, n_queries, off))
        return fail(ctx, GGCAT_GPU_EINVAL, "queries must be n_queries records, each ending with '\\n'");
    if (query_kmers) query_kmer_counts(off, ctx->P.k, query_kmers);
    HIPCHK(hipSetDevice(ctx->device));
    const uint32_t nc = ctx->colored ? ctx->n_color_ids : 0;
    ENSURE(ctx->q_seq, n_bytes + 16);
    ENSURE(ctx->q_off, ((size_t)n_queries + 1) * 8);
    ENSURE(ctx->q_matched, (size_t)n_queries * 8 + 8);
    ENSURE(ctx->q_counts, (size_t)n_queries * nc * 4 + 4);
    ENSURE(ctx->q_work, 8);
    if (n_bytes) HIPCHK(hipMemcpy(ctx->q_seq.p, queries, n_bytes, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->q_off.p, off.data(), off.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemsetAsync(ctx->q_work.p, 0, 8, ctx->stream));
    QueryArgs Q = query_args(ctx);
    Q.seq = P_<uint8_t>(ctx->q_seq);
    Q.qoff = P_<uint64_t>(ctx->q_off);
    Q.n_queries = n_queries;
    Q.matched = P_<uint64_t>(ctx->q_matched);
    Q.color_counts = P_<uint32_t>(ctx->q_counts);
    Q.work = P_<unsigned int>(ctx->q_work);
    if (n_queries) {
        const uint32_t grid = std::min<uint32_t>(n_queries, ctx->n_cu * 4);
        hipLaunchKernelGGL(query_kernel, dim3(grid), dim3(256), (size_t)nc * 4 + 16, ctx->stream, Q);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (matched && n_queries) HIPCHK(hipMemcpy(matched, ctx->q_matched.p, (size_t)n_queries * 8, hipMemcpyDeviceToHost));
    if (color_counts && nc && n_queries)
        HIPCHK(hipMemcpy(color_counts, ctx->q_counts.p, (size_t)n_queries * nc * 4, hipMemcpyDeviceToHost));
    return GGCAT_GPU_OK;
}

int ggcat_gpu_serve_open(ggcat_gpu_ctx* ctx, uint32_t streams, uint64_t memory_budget) {
    if (!ctx || streams == 0 || streams > 16 || memory_budget < 4096) return GGCAT_GPU_EINVAL;
    if (!ctx->have_index) return fail(ctx, GGCAT_GPU_ESTATE, "no query index (call ggcat_gpu_index_build)");
    if (ctx->server) return fail(ctx, GGCAT_GPU_ESTATE, "query server already open");
    HIPCHK(hipSetDevice(ctx->device));
    std::unique_ptr<Server> S(new Server());
    S->budget = memory_budget;
    hipError_t e = hipMalloc(&S->dev, memory_budget);
    if (e != hipSuccess) return fail(ctx, GGCAT_GPU_ENOMEM, "serving device memory budget not available");
    e = hipHostMalloc(&S->pin, memory_budget, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipFree(S->dev);
        return fail(ctx, GGCAT_GPU_ENOMEM, "serving pinned staging not available");
    }
    for (uint32_t i = 0; i < streams; i++) {
        hipStream_t st;
        e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e != hipSuccess) break;
        S->streams.push_back(st);
    }
    S->load.assign(S->streams.size(), 0);
    S->dalloc.init(memory_budget);
    S->palloc.init(memory_budget);
    ctx->server = S.release();
    if (e != hipSuccess) {
        (void)ggcat_gpu_serve_close(ctx);
        return fail(ctx, GGCAT_GPU_EHIP, "stream creation failed");
    }
    return GGCAT_GPU_OK;
}

int ggcat_gpu_serve_submit(ggcat_gpu_ctx* ctx, const uint8_t* queries, uint64_t n_bytes, uint32_t n_queries,
                           uint64_t* ticket) {
    if (!ctx || !ticket || (!queries && n_bytes)) return GGCAT_GPU_EINVAL;
    if (!ctx->server) return fail(ctx, GGCAT_GPU_ESTATE, "query server not open");
    Server& S = *ctx->server;
    const uint64_t t = S.next_ticket++;
    ServeReq& r = S.reqs[t];
    if (!query_offsets(queries, n_bytes, n_queries, r.off)) {
        S.reqs.erase(t);
        return fail(ctx, GGCAT_GPU_EINVAL, "queries must be n_queries records, each ending with '\\n'");
    }
    r.src = queries;
    r.n_bytes = n_bytes;
    r.n_queries = n_queries;
    serve_layout(ctx, r);
    if (r.bytes > S.budget) {
        // could never be admitted: rejected here, so it cannot block the queue
        const uint64_t need = r.bytes;
        S.reqs.erase(t);
        return fail(ctx, GGCAT_GPU_ENOMEM, "request larger than the serving memory budget (" +
                                               std::to_string(need) + " > " + std::to_string(S.budget) + " bytes)");
    }
    S.waiting.push_back(t);
    *ticket = t;
    HIPCHK(hipSetDevice(ctx->device));
    // admitted at once: the bytes go straight from the caller's buffer to the
    // pinned staging; a request left waiting keeps its own copy
    const int rc = serve_admit(ctx);
    if (!r.admitted) {
        r.data.assign(queries, queries + n_bytes);
        r.src = r.data.data();
    }
    return rc;
}

int ggcat_gpu_serve_poll(ggcat_gpu_ctx* ctx, int block, uint64_t* ticket) {
    if (!ctx || !ticket) return GGCAT_GPU_EINVAL;
    if (!ctx->server) return fail(ctx, GGCAT_GPU_ESTATE, "query server not open");
    Server& S = *ctx->server;
    HIPCHK(hipSetDevice(ctx->device));
    int r0 = serve_admit(ctx);
    if (r0 != GGCAT_GPU_OK) return r0;
    uint64_t oldest = 0;
    for (auto& kv : S.reqs) {
        ServeReq& r = kv.second;
        if (!r.admitted || r.done) continue;
        if (!oldest) oldest = kv.first;
        hipError_t q = hipEventQuery(r.ev);
        if (q == hipSuccess) {
            r.done = true;
            *ticket = kv.first;
            return GGCAT_GPU_OK;
        }
        if (q != hipErrorNotReady) return fail(ctx, GGCAT_GPU_EHIP, std::string("query request: ") + hipGetErrorString(q));
    }
    if (block && oldest) {
        ServeReq& r = S.reqs[oldest];
        HIPCHK(hipEventSynchronize(r.ev));
        r.done = true;
        *ticket = oldest;
        return GGCAT_GPU_OK;
    }
    return GGCAT_GPU_EAGAIN;
}

int ggcat_gpu_serve_fetch(ggcat_gpu_ctx* ctx, uint64_t ticket, uint64_t* query_kmers, uint64_t* matched,
                          uint32_t* color_counts) {
    if (!ctx) return GGCAT_GPU_EINVAL;
    if (!ctx->server) return fail(ctx, GGCAT_GPU_ESTATE, "query server not open");
    Server& S = *ctx->server;
    auto it = S.reqs.find(ticket);
    if (it == S.reqs.end() || !it->second.done) return fail(ctx, GGCAT_GPU_ESTATE, "request not finished");
    ServeReq& r = it->second;
    const uint32_t nq = r.n_queries, nc = ctx->colored ? ctx->n_color_ids : 0;
    const uint8_t* hp = S.pin + r.pin_off;
    if (query_kmers) query_kmer_counts(r.off, ctx->P.k, query_kmers);
    if (matched) memcpy(matched, hp + r.o_matched, (size_t)nq * 8);
    if (color_counts && nc) memcpy(color_counts, hp + r.o_counts, (size_t)nq * nc * 4);
    S.dalloc.release(r.dev_off, r.bytes);
    S.palloc.release(r.pin_off, r.bytes);
    S.free_events.push_back(r.ev);
    S.load[r.stream]--;
    S.n_admitted--;
    S.reqs.erase(it);
    HIPCHK(hipSetDevice(ctx->device));
    return serve_admit(ctx);
}

int ggcat_gpu_serve_stats(ggcat_gpu_ctx* ctx, uint32_t* outstanding, uint32_t* admitted) {
    if (!ctx) return GGCAT_GPU_EINVAL;
    if (!ctx->server) return fail(ctx, GGCAT_GPU_ESTATE, "query server not open");
    if (outstanding) *outstanding = (uint32_t)ctx->server->reqs.size();
    if (admitted) *admitted = ctx->server->n_admitted;
    return GGCAT_GPU_OK;
}

int ggcat_gpu_serve_close(ggcat_gpu_ctx* ctx) {
    if (!ctx) return GGCAT_GPU_EINVAL;
    Server* S = ctx->server;
    if (!S) return GGCAT_GPU_OK;
    (void)hipSetDevice(ctx->device);
    for (hipStream_t st : S->streams) (void)hipStreamSynchronize(st);
    for (auto& kv : S->reqs)
        if (kv.second.ev) S->free_events.push_back(kv.second.ev);
    for (hipEvent_t e : S->free_events) (void)hipEventDestroy(e);
    for (hipStream_t st : S->streams) (void)hipStreamDestroy(st);
    if (S->dev) (void)hipFree(S->dev);
    if (S->pin) (void)hipHostFree(S->pin);
    delete S;
    ctx->server = nullptr;
    return GGCAT_GPU_OK;
}

namespace {

// Unitigs of the last merge's solid table: per sub-bucket (joined == false)
// or over the whole table, joined across sub-buckets (joined == true).
static int build_unitigs(ggcat_gpu_ctx* ctx, bool joined, uint64_t* n_unitigs, uint64_t* n_bytes) {
    if (!ctx->have_solid) return fail(ctx, GGCAT_GPU_ESTATE, "no merge result");
    if (ctx->P.hash_kind != GGCAT_GPU_HASH_SEQ)
        return fail(ctx, GGCAT_GPU_EINVAL, "unitig extension needs seq-hash keys (the k-mer is the key)");
    HIPCHK(hipSetDevice(ctx->device));
    const uint64_t n_pos = (uint64_t)ctx->n_chunks_used * OUT_CHUNK + ctx->n_tail;
    const uint64_t ne = 2 * n_pos;
    if (ne >= (uint64_t)NO_LINK) return fail(ctx, GGCAT_GPU_EINVAL, "solid table too large for unitig links");
    ENSURE(ctx->u_nb, ne * 4 + 8);
    ENSURE(ctx->u_link, ne * 4 + 8);
    ENSURE(ctx->u_ext, n_pos + 8);
    ENSURE(ctx->u_hbytes, ne * 4 + 8);
    ENSURE(ctx->u_hcount, ne * 4 + 8);
    ENSURE(ctx->u_hoff, (ne + 1) * 8);
    ENSURE(ctx->u_hidx, (ne + 1) * 8);
    ENSURE(ctx->u_visited, n_pos + 8);
    ENSURE(ctx->u_rest_n, 8);
    ENSURE(ctx->u_rep, n_pos + 8);
    uint64_t tab_slots = 1024;
    while (tab_slots < 2 * n_pos) tab_slots <<= 1;
    ENSURE(ctx->u_tab, tab_slots * 8);
    ctx->have_unitigs = false;
    UnitigArgs U{};
    U.out_key = P_<uint64_t>(ctx->out_key);
    U.out_val = P_<uint32_t>(ctx->out_val);
    U.out_sub = P_<uint32_t>(ctx->out_sub);
    U.chunk_fill = P_<uint32_t>(ctx->chunk_fill_dev);
    U.n_chunks = ctx->n_chunks_used;
    U.n_tail = ctx->n_tail;
    U.chunk = OUT_CHUNK;
    U.key_words = ctx->key_words;
    U.k = ctx->P.k;
    U.tab = P_<uint64_t>(ctx->u_tab);
    U.tab_mask = tab_slots - 1;
    U.nb = P_<uint32_t>(ctx->u_nb);
    U.ext = P_<uint8_t>(ctx->u_ext);
    U.link = P_<uint32_t>(ctx->u_link);
    U.head_bytes = P_<uint32_t>(ctx->u_hbytes);
    U.head_count = P_<uint32_t>(ctx->u_hcount);
    U.head_off = P_<uint64_t>(ctx->u_hoff);
    U.head_idx = P_<uint64_t>(ctx->u_hidx);
    U.visited = P_<uint8_t>(ctx->u_visited);
    U.global = joined ? 1u : 0u;
    U.rep = P_<uint8_t>(ctx->u_rep);
    U.n_nodes = P_<unsigned int>(ctx->u_rest_n) + 1;
    U.mark_visited = 0;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((ne + 255) / 256 + 1, (uint64_t)ctx->n_cu * 16);
    HIPCHK(hipEventRecord(ctx->ev[11], ctx->stream));
    HIPCHK(hipMemsetAsync(ctx->u_rest_n.p, 0, 8, ctx->stream));  // [0] cycle nodes, [1] graph nodes
    if (n_pos) {
        HIPCHK(hipMemsetAsync(ctx->u_tab.p, 0, tab_slots * 8, ctx->stream));
        hipLaunchKernelGGL(utab_insert_kernel, dim3(grid), dim3(256), 0, ctx->stream, U, n_pos);
        if (joined) hipLaunchKernelGGL(rep_kernel, dim3(grid), dim3(256), 0, ctx->stream, U, n_pos);
        hipLaunchKernelGGL(ext_kernel, dim3(grid), dim3(256), 0, ctx->stream, U, n_pos);
        hipLaunchKernelGGL(link_kernel, dim3(grid), dim3(256), 0, ctx->stream, U, n_pos);
        hipLaunchKernelGGL(walk_kernel<false>, dim3(grid), dim3(256), 0, ctx->stream, U, n_pos);
        HIPCHK(hipGetLastError());
    }
    int r = device_scan(ctx, U.head_bytes, P_<uint64_t>(ctx->u_hoff), (uint32_t)ne);
    if (r != GGCAT_GPU_OK) return r;
    r = device_scan(ctx, U.head_count, P_<uint64_t>(ctx->u_hidx), (uint32_t)ne);
    if (r != GGCAT_GPU_OK) return r;
    uint64_t tot_bytes = 0, tot_n = 0;
    uint32_t counts[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(&tot_bytes, P_<uint64_t>(ctx->u_hoff) + ne, 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(&tot_n, P_<uint64_t>(ctx->u_hidx) + ne, 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(counts, ctx->u_rest_n.p, 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    // Paths cover every node unless there are cycles: only then are the path
    // nodes marked, so that the remaining ones can be collected.
    uint64_t n_nodes = counts[1];
    if (!joined) {
        n_nodes = ctx->n_tail;
        for (uint32_t c = 0; c < ctx->n_chunks_used; c++) n_nodes += ctx->chunk_fill[c];
    }
    const bool cycles = tot_bytes - (uint64_t)ctx->P.k * tot_n != n_nodes;
    U.mark_visited = cycles ? 1u : 0u;
    ENSURE(ctx->u_seq, tot_bytes + 8);
    ENSURE(ctx->u_sub, tot_n * 4 + 8);
    ENSURE(ctx->u_flags, tot_n + 8);
    U.seq = P_<uint8_t>(ctx->u_seq);
    U.uni_sub = P_<uint32_t>(ctx->u_sub);
    U.uni_flags = P_<uint8_t>(ctx->u_flags);
    if (ctx->colored) {
        // the k-mers' table positions give each unitig's color runs
        ENSURE(ctx->u_path, (tot_bytes - (uint64_t)ctx->P.k * tot_n) * 4 + 8);
        U.path_pos = P_<uint32_t>(ctx->u_path);
    }
    if (cycles) HIPCHK(hipMemsetAsync(ctx->u_visited.p, 0, n_pos + 8, ctx->stream));
    const uint32_t cap = 1u << 20;
    ENSURE(ctx->u_rest, (size_t)cap * 4);
    if (n_pos) {
        hipLaunchKernelGGL(walk_kernel<true>, dim3(grid), dim3(256), 0, ctx->stream, U, n_pos);
        if (cycles)
            hipLaunchKernelGGL(unvisited_kernel, dim3(grid), dim3(256), 0, ctx->stream, U, n_pos,
                               P_<uint32_t>(ctx->u_rest), P_<unsigned int>(ctx->u_rest_n), cap);
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipEventRecord(ctx->ev[12], ctx->stream));
    uint32_t n_rest = 0;
    HIPCHK(hipMemcpyAsync(&n_rest, ctx->u_rest_n.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (n_rest > cap) return fail(ctx, GGCAT_GPU_ENOMEM, "too many k-mers on cycles");
    // Cycles: every remaining k-mer has both sides linked.  Walk each from its
    // smallest key, forward (entering at side 0), on the host.
    ctx->cyc_seq.clear();
    ctx->cyc_sub.clear();
    ctx->cyc_flags.clear();
    ctx->cyc_pos.clear();
    if (n_rest) {
        std::vector<uint32_t> rest(n_rest);
        HIPCHK(hipMemcpy(rest.data(), ctx->u_rest.p, (size_t)n_rest * 4, hipMemcpyDeviceToHost));
        const uint32_t kw = ctx->key_words, k = ctx->P.k;
        std::map<uint32_t, std::pair<u128_t, std::array<uint32_t, 2>>> node;  // pos -> key, links
        for (uint32_t p : rest) {
            uint64_t kk[2] = {0, 0};
            uint32_t lk[2];
            HIPCHK(hipMemcpy(kk, P_<uint64_t>(ctx->out_key) + (uint64_t)p * kw, 8 * kw, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(lk, P_<uint32_t>(ctx->u_link) + 2 * (uint64_t)p, 8, hipMemcpyDeviceToHost));
            node[p] = {((u128_t)kk[1] << 64) | kk[0], {lk[0], lk[1]}};
        }
        std::vector<uint32_t> order(rest);
        std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return node[a].first < node[b].first; });
        std::map<uint32_t, bool> seen;
        const char B[4] = {'A', 'C', 'T', 'G'};
        auto rc = [&](u128_t x) {
            u128_t y = 0;
            for (uint32_t i = 0; i < k; i++) y |= (u128_t)((((uint32_t)(x >> (2 * i))) & 3u) ^ 2u) << (2 * (k - 1 - i));
            return y;
        };
        for (uint32_t p0 : order) {
            if (seen[p0]) continue;
            seen[p0] = true;
            const u128_t x = node[p0].first;
            for (uint32_t i = 0; i < k; i++) ctx->cyc_seq.push_back((uint8_t)B[(uint32_t)(x >> (2 * i)) & 3u]);
            ctx->cyc_pos.push_back(p0);
            uint64_t cur = 2 * (uint64_t)p0;  // entered at side 0
            uint32_t sub = 0;
            HIPCHK(hipMemcpy(&sub, P_<uint32_t>(ctx->out_sub) + p0, 4, hipMemcpyDeviceToHost));
            for (uint64_t guard = 0; guard <= n_rest; guard++) {
                auto it = node.find((uint32_t)(cur >> 1));
                if (it == node.end()) return fail(ctx, GGCAT_GPU_EHIP, "cycle walk left the cycle set");
                const uint32_t nx = it->second.second[(cur ^ 1) & 1];
                if (nx == NO_LINK) return fail(ctx, GGCAT_GPU_EHIP, "cycle walk met an open side");
                const uint32_t q = nx >> 1;
                if (seen[q]) break;
                seen[q] = true;
                auto jt = node.find(q);
                if (jt == node.end()) return fail(ctx, GGCAT_GPU_EHIP, "cycle walk left the cycle set");
                const u128_t y = (nx & 1) == 0 ? jt->second.first : rc(jt->second.first);
                ctx->cyc_seq.push_back((uint8_t)B[(uint32_t)(y >> (2 * (k - 1))) & 3u]);
                ctx->cyc_pos.push_back(q);
                cur = nx;
            }
            ctx->cyc_seq.push_back('\n');
            ctx->cyc_sub.push_back(joined ? 0u : sub);
            ctx->cyc_flags.push_back(4);
        }
    }
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, ctx->ev[11], ctx->ev[12]));
    ctx->ms[7] = t;
    ctx->u_n = tot_n;
    ctx->u_bytes = tot_bytes;
    ctx->u_npos = n_pos;
    ctx->have_unitigs = true;
    ctx->u_joined = joined;
    if (n_unitigs) *n_unitigs = tot_n + ctx->cyc_sub.size();
    if (n_bytes) *n_bytes = tot_bytes + ctx->cyc_seq.size();
    return GGCAT_GPU_OK;
}

}  // namespace

int ggcat_gpu_unitigs(ggcat_gpu_ctx* ctx, uint64_t* n_unitigs, uint64_t* n_bytes) {
    if (!ctx) return GGCAT_GPU_EINVAL;
    return build_unitigs(ctx, false, n_unitigs, n_bytes);
}

int ggcat_gpu_maximal_unitigs(ggcat_gpu_ctx* ctx, uint64_t* n_unitigs, uint64_t* n_bytes) {
    if (!ctx) return GGCAT_GPU_EINVAL;
    return build_unitigs(ctx, true, n_unitigs, n_bytes);
}

namespace {

// Color subset ids of the solid table (colored builds): every distinct color
// set gets one id, ids ordered by the sets' bitmaps so that they do not depend
// on the table's layout (the reference assigns them first-come,
// crates/colors/src/colors_memmap_writer.rs:50-59, so its ids vary between
// runs; the sets do not).
static int ensure_subsets(ggcat_gpu_ctx* ctx) {
    if (ctx->have_subsets) return GGCAT_GPU_OK;
    if (!ctx->have_solid) return fail(ctx, GGCAT_GPU_ESTATE, "no merge result");
    if (!ctx->colored) return fail(ctx, GGCAT_GPU_ESTATE, "the last build was not colored");
    HIPCHK(hipSetDevice(ctx->device));
    const uint32_t W = ctx->color_words;
    const uint64_t n_pos = (uint64_t)ctx->n_chunks_used * OUT_CHUNK + ctx->n_tail;
    std::vector<uint32_t> bm(n_pos * W);
    if (n_pos) HIPCHK(hipMemcpy(bm.data(), ctx->out_colors.p, n_pos * W * 4, hipMemcpyDeviceToHost));
    auto filled = [&](uint64_t p) {
        const uint64_t c = p / OUT_CHUNK;
        return c < ctx->n_chunks_used ? (p - c * OUT_CHUNK) < ctx->chunk_fill[c]
                                      : p - (uint64_t)ctx->n_chunks_used * OUT_CHUNK < ctx->n_tail;
    };
    auto less = [&](uint64_t a, uint64_t b) {
        return std::lexicographical_compare(&bm[a * W], &bm[a * W] + W, &bm[b * W], &bm[b * W] + W);
    };
    std::vector<std::pair<uint64_t, uint32_t>> hp;  // (hash of the bitmap, position)
    hp.reserve(n_pos);
    for (uint64_t p = 0; p < n_pos; p++) {
        if (!filled(p)) continue;
        uint64_t h = 0x9E3779B97F4A7C15ULL;
        for (uint32_t w = 0; w < W; w++) h = (h ^ bm[p * W + w]) * 0xbf58476d1ce4e5b9ULL;
        hp.emplace_back(h ^ (h >> 31), (uint32_t)p);
    }
    std::sort(hp.begin(), hp.end(), [&](const std::pair<uint64_t, uint32_t>& a, const std::pair<uint64_t, uint32_t>& b) {
        return a.first != b.first ? a.first < b.first : less(a.second, b.second);
    });
    // distinct bitmaps (equal bitmaps are adjacent), then ordered by bitmap
    std::vector<uint32_t> rep;
    std::vector<uint32_t> group(hp.size());
    for (size_t i = 0; i < hp.size(); i++) {
        const bool same = i > 0 && hp[i].first == hp[i - 1].first &&
                          std::equal(&bm[(uint64_t)hp[i].second * W], &bm[(uint64_t)hp[i].second * W] + W,
                                     &bm[(uint64_t)hp[i - 1].second * W]);
        if (!same) rep.push_back(hp[i].second);
        group[i] = (uint32_t)rep.size() - 1;
    }
    std::vector<uint32_t> order(rep.size());
    for (uint32_t i = 0; i < order.size(); i++) order[i] = i;
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return less(rep[a], rep[b]); });
    std::vector<uint32_t> id_of(rep.size());
    ctx->subsets.assign(rep.size(), {});
    for (uint32_t i = 0; i < order.size(); i++) {
        id_of[order[i]] = i;
        const uint64_t p = rep[order[i]];
        for (uint32_t w = 0; w < W; w++)
            for (uint32_t bits = bm[p * W + w]; bits; bits &= bits - 1)
                ctx->subsets[i].push_back(w * 32 + (uint32_t)__builtin_ctz(bits));
        if (ctx->subsets[i].empty()) return fail(ctx, GGCAT_GPU_EHIP, "a solid k-mer without colors");
    }
    ctx->pos_subset.assign(n_pos, ~0u);
    for (size_t i = 0; i < hp.size(); i++) ctx->pos_subset[hp[i].second] = id_of[group[i]];
    ctx->have_subsets = true;
    return GGCAT_GPU_OK;
}

}  // namespace

int ggcat_gpu_color_subsets(ggcat_gpu_ctx* ctx, uint64_t* n_subsets, uint32_t* subset_of_entry) {
    if (!ctx) return GGCAT_GPU_EINVAL;
    const int r = ensure_subsets(ctx);
    if (r != GGCAT_GPU_OK) return r;
    if (n_subsets) *n_subsets = ctx->subsets.size();
    if (subset_of_entry) {
        uint64_t o = 0;
        for (uint64_t p = 0; p < ctx->pos_subset.size(); p++)
            if (ctx->pos_subset[p] != ~0u) subset_of_entry[o++] = ctx->pos_subset[p];
    }
    return GGCAT_GPU_OK;
}

int ggcat_gpu_colormap_write(ggcat_gpu_ctx* ctx, const char* path, const char* const* color_names, uint32_t n_names,
                             uint64_t* n_subsets) {
    if (!ctx || !path || (!color_names && n_names)) return GGCAT_GPU_EINVAL;
    int r = ensure_subsets(ctx);
    if (r != GGCAT_GPU_OK) return r;
    if (n_names < ctx->n_color_ids)
        return fail(ctx, GGCAT_GPU_EINVAL, "every color id of the build needs a name");
    std::vector<std::string> names;
    for (uint32_t i = 0; i < n_names; i++) {
        if (!color_names[i]) return fail(ctx, GGCAT_GPU_EINVAL, "null color name");
        names.emplace_back(color_names[i]);
    }
    std::string err;
    if (!gg_host::colormap_write(path, names, ctx->subsets, &err)) return fail(ctx, GGCAT_GPU_EINVAL, err);
    if (n_subsets) *n_subsets = ctx->subsets.size();
    return GGCAT_GPU_OK;
}

int ggcat_gpu_colormap_write_sets(const char* path, const char* const* color_names, uint32_t n_names,
                                  const uint64_t* subset_off, const uint32_t* colors, uint64_t n_subsets) {
    if (!path || (!color_names && n_names) || !subset_off || (!colors && subset_off[n_subsets]))
        return fail(nullptr, GGCAT_GPU_EINVAL, "null argument");
    std::vector<std::string> names;
    for (uint32_t i = 0; i < n_names; i++) {
        if (!color_names[i]) return fail(nullptr, GGCAT_GPU_EINVAL, "null color name");
        names.emplace_back(color_names[i]);
    }
    std::vector<std::vector<uint32_t>> sets(n_subsets);
    for (uint64_t s = 0; s < n_subsets; s++) {
        if (subset_off[s + 1] <= subset_off[s]) return fail(nullptr, GGCAT_GPU_EINVAL, "empty color subset");
        sets[s].assign(colors + subset_off[s], colors + subset_off[s + 1]);
        for (size_t i = 1; i < sets[s].size(); i++)
            if (sets[s][i] <= sets[s][i - 1])
                return fail(nullptr, GGCAT_GPU_EINVAL, "color subsets must be strictly increasing");
        if (sets[s].back() >= n_names) return fail(nullptr, GGCAT_GPU_EINVAL, "a color without a name");
    }
    std::string err;
    if (!gg_host::colormap_write(path, names, sets, &err)) return fail(nullptr, GGCAT_GPU_EINVAL, err);
    return GGCAT_GPU_OK;
}

struct ggcat_gpu_colormap {
    gg_host::ColorMap m;
};

int ggcat_gpu_colormap_open(const char* path, ggcat_gpu_colormap** out) {
    if (!path || !out) return fail(nullptr, GGCAT_GPU_EINVAL, "null argument");
    *out = nullptr;
    std::unique_ptr<ggcat_gpu_colormap> c(new ggcat_gpu_colormap());
    std::string err;
    if (!gg_host::colormap_read(path, c->m, &err)) return fail(nullptr, GGCAT_GPU_EINVAL, err);
    *out = c.release();
    return GGCAT_GPU_OK;
}

int ggcat_gpu_colormap_info(const ggcat_gpu_colormap* c, uint32_t* n_colors, uint64_t* n_subsets) {
    if (!c) return GGCAT_GPU_EINVAL;
    if (n_colors) *n_colors = (uint32_t)c->m.names.size();
    if (n_subsets) *n_subsets = c->m.subset_off.size() - 1;
    return GGCAT_GPU_OK;
}

const char* ggcat_gpu_colormap_color_name(const ggcat_gpu_colormap* c, uint32_t color) {
    if (!c || color >= c->m.names.size()) return nullptr;
    return c->m.names[color].c_str();
}

int ggcat_gpu_colormap_subset(const ggcat_gpu_colormap* c, uint64_t subset, uint32_t* colors, uint32_t cap,
                              uint32_t* n) {
    if (!c || !n || subset + 1 >= c->m.subset_off.size()) return GGCAT_GPU_EINVAL;
    const uint64_t a = c->m.subset_off[subset], b = c->m.subset_off[subset + 1];
    *n = (uint32_t)(b - a);
    if (colors) {
        if (cap < b - a) return GGCAT_GPU_EINVAL;
        memcpy(colors, c->m.colors.data() + a, (b - a) * 4);
    }
    return GGCAT_GPU_OK;
}

void ggcat_gpu_colormap_close(ggcat_gpu_colormap* c) { delete c; }

int ggcat_gpu_unitigs_write_fasta(ggcat_gpu_ctx* ctx, const char* path, uint64_t* n_written) {
    if (!ctx || !path) return GGCAT_GPU_EINVAL;
    if (!ctx->have_unitigs) return fail(ctx, GGCAT_GPU_ESTATE, "no unitig result");
    HIPCHK(hipSetDevice(ctx->device));
    const uint64_t total = ctx->u_bytes + ctx->cyc_seq.size();
    std::vector<uint8_t> seq(total);
    if (ctx->u_bytes) HIPCHK(hipMemcpy(seq.data(), ctx->u_seq.p, ctx->u_bytes, hipMemcpyDeviceToHost));
    if (!ctx->cyc_seq.empty()) memcpy(seq.data() + ctx->u_bytes, ctx->cyc_seq.data(), ctx->cyc_seq.size());
    // colored builds: each unitig's k-mers' color subset ids as runs
    // " C:{id hex}:{count}" (crates/colors/src/managers/multiple.rs:707-740)
    std::vector<uint32_t> path_pos;
    const uint32_t k = ctx->P.k;
    if (ctx->colored) {
        const int r = ensure_subsets(ctx);
        if (r != GGCAT_GPU_OK) return r;
        const uint64_t n_path = ctx->u_bytes - (uint64_t)k * ctx->u_n;
        path_pos.resize(n_path + ctx->cyc_pos.size());
        if (n_path) HIPCHK(hipMemcpy(path_pos.data(), ctx->u_path.p, n_path * 4, hipMemcpyDeviceToHost));
        std::copy(ctx->cyc_pos.begin(), ctx->cyc_pos.end(), path_pos.begin() + n_path);
    }
    FILE* f = fopen(path, "wb");
    if (!f) return fail(ctx, GGCAT_GPU_EINVAL, std::string("cannot create ") + path);
    std::vector<char> buf(1 << 20);
    setvbuf(f, buf.data(), _IOFBF, buf.size());
    uint64_t id = 0, kp = 0;
    bool ok = true;
    for (uint64_t o = 0; o < total && ok;) {
        const uint8_t* nl = static_cast<const uint8_t*>(memchr(seq.data() + o, '\n', total - o));
        const uint64_t len = nl ? (uint64_t)(nl - (seq.data() + o)) : total - o;
        ok = fprintf(f, ">%llu LN:i:%llu", (unsigned long long)id, (unsigned long long)len) > 0;
        if (ctx->colored && ok) {
            const uint64_t nk = len - k + 1;
            if (kp + nk > path_pos.size()) {
                fclose(f);
                return fail(ctx, GGCAT_GPU_EHIP, "unitig k-mer positions do not cover the unitigs");
            }
            uint32_t run_id = ~0u;
            uint64_t run_n = 0;
            for (uint64_t i = 0; i < nk && ok; i++) {
                const uint32_t s = ctx->pos_subset[path_pos[kp + i]];
                if (s == run_id) {
                    run_n++;
                    continue;
                }
                if (run_n) ok = fprintf(f, " C:%x:%llu", run_id, (unsigned long long)run_n) > 0;
                run_id = s;
                run_n = 1;
            }
            if (run_n && ok) ok = fprintf(f, " C:%x:%llu", run_id, (unsigned long long)run_n) > 0;
            kp += nk;
        }
        ok = ok && fputc('\n', f) != EOF && fwrite(seq.data() + o, 1, len, f) == len && fputc('\n', f) != EOF;
        o += len + 1;
        id++;
    }
    ok = (fclose(f) == 0) && ok;
    if (!ok) return fail(ctx, GGCAT_GPU_EINVAL, std::string("write failed: ") + path);
    if (n_written) *n_written = id;
    return GGCAT_GPU_OK;
}

namespace {

// CanonicalSeqHash of a k-mer of ASCII bases in sequence orientation:
// (fw, rc) with base i at bits 2i of fw (crates/hashes/src/base/cn_seqhash_base.rs:22-72).
static void seqhash_pair(const uint8_t* s, uint32_t k, u128_t* fw, u128_t* rc) {
    u128_t f = 0, r = 0;
    for (uint32_t i = 0; i < k; i++) {
        const u128_t c = (s[i] >> 1) & 3;
        f |= c << (2 * i);
        r |= (c ^ 2) << (2 * (k - 1 - i));
    }
    *fw = f;
    *rc = r;
}

// CanonicalSeqHashFactory::get_bucket(0, log, hash)
// (crates/hashes/src/base/cn_seqhash_base.rs:140-150, constants
// crates/hashes/src/cn_seqhash.rs:15-27): hash * MULTIPLIER + BASIS, rotated
// right by 3, low `log` bits; 64-bit arithmetic for k <= 32, 128-bit above.
static uint32_t seqhash_bucket(u128_t h, uint32_t k, uint32_t log) {
    if (k <= 32) {
        uint64_t x = (uint64_t)h * 0x00000100000001b3ULL + 0xcbf29ce484222325ULL;
        x = (x >> 3) | (x << 61);
        return (uint32_t)(x & ((1ull << log) - 1));
    }
    const u128_t mul = ((u128_t)0x0000000001000000ULL << 64) | 0x000000000000013bULL;
    const u128_t basis = ((u128_t)0x6c62272e07bb0142ULL << 64) | 0x62b821756295c58dULL;
    u128_t x = h * mul + basis;
    x = (x >> 3) | (x << 125);
    return (uint32_t)((uint64_t)x & ((1ull << log) - 1));
}

}  // namespace

int ggcat_gpu_partial_unitigs(ggcat_gpu_ctx* ctx, uint32_t buckets_log, uint8_t* kind, uint16_t* bucket,
                              uint8_t* rc, uint8_t* last_align, uint8_t* flags) {
    if (!ctx) return GGCAT_GPU_EINVAL;
    if (!ctx->have_unitigs || ctx->u_joined) return fail(ctx, GGCAT_GPU_ESTATE, "no per-sub-bucket unitig result");
    if (buckets_log > 16) return fail(ctx, GGCAT_GPU_EINVAL, "buckets_log must be <= 16");
    HIPCHK(hipSetDevice(ctx->device));
    const uint32_t k = ctx->P.k;
    const uint64_t n = ctx->u_n + ctx->cyc_sub.size();
    std::vector<uint8_t> seq(ctx->u_bytes + ctx->cyc_seq.size()), fl(n);
    if (ctx->u_bytes) HIPCHK(hipMemcpy(seq.data(), ctx->u_seq.p, ctx->u_bytes, hipMemcpyDeviceToHost));
    memcpy(seq.data() + ctx->u_bytes, ctx->cyc_seq.data(), ctx->cyc_seq.size());
    if (ctx->u_n) HIPCHK(hipMemcpy(fl.data(), ctx->u_flags.p, ctx->u_n, hipMemcpyDeviceToHost));
    memcpy(fl.data() + ctx->u_n, ctx->cyc_flags.data(), ctx->cyc_flags.size());
    // ParallelKmersMergeFinalExecutor::output_sequence
    // (crates/assembler_kmers_merge/src/final_executor.rs:103-246)
    uint64_t o = 0;
    for (uint64_t i = 0; i < n; i++) {
        const uint8_t* s = seq.data() + o;
        const uint8_t* nl = static_cast<const uint8_t*>(memchr(s, '\n', seq.size() - o));
        const uint64_t len = nl ? (uint64_t)(nl - s) : seq.size() - o;
        o += len + 1;
        if (len < k) return fail(ctx, GGCAT_GPU_EHIP, "unitig shorter than k");
        const bool back = fl[i] & 1, fwd = (fl[i] & 2) != 0, circ = (fl[i] & 4) != 0;
        uint8_t kd = 0, r = 0, la = 0, f = 0;
        uint16_t bk = 0xFFFF;
        if (circ) {
            kd = 1;  // lonely circular unitig
        } else if (back || fwd) {
            kd = 2;
            uint32_t lb = 0xFFFF, rb = 0xFFFF;
            bool lrc = false, rrc = true;
            u128_t hf, hr;
            if (back) {  // get_extremal_hash(out_seq, k, true): the first k-mer
                seqhash_pair(s, k, &hf, &hr);
                lb = seqhash_bucket(hf < hr ? hf : hr, k, buckets_log);
                lrc = !(hf < hr);
            }
            if (fwd) {  // the last k-mer
                seqhash_pair(s + len - k, k, &hf, &hr);
                rb = seqhash_bucket(hf < hr ? hf : hr, k, buckets_log);
                rrc = !(hf < hr);
            }
            // always the smaller bucket
            const bool hash_beginning = lb <= rb;
            const bool should_rc = hash_beginning ? lrc : rrc;
            bk = (uint16_t)(hash_beginning ? lb : rb);
            r = should_rc;
            la = (hash_beginning ^ should_rc) ? 0 : (uint8_t)((len - k) % 4);
            f = (uint8_t)((((!hash_beginning) ^ should_rc) ? 2 : 0) | (back && fwd ? 4 : 0));
        }
        if (kind) kind[i] = kd;
        if (bucket) bucket[i] = bk;
        if (rc) rc[i] = r;
        if (last_align) last_align[i] = la;
        if (flags) flags[i] = f;
    }
    return GGCAT_GPU_OK;
}

int ggcat_gpu_unitigs_copy(ggcat_gpu_ctx* ctx, uint8_t* seq, uint32_t* sub, uint8_t* flags) {
    if (!ctx) return GGCAT_GPU_EINVAL;
    if (!ctx->have_unitigs) return fail(ctx, GGCAT_GPU_ESTATE, "no unitig result");
    HIPCHK(hipSetDevice(ctx->device));
    if (seq) {
        if (ctx->u_bytes) HIPCHK(hipMemcpy(seq, ctx->u_seq.p, ctx->u_bytes, hipMemcpyDeviceToHost));
        memcpy(seq + ctx->u_bytes, ctx->cyc_seq.data(), ctx->cyc_seq.size());
    }
    if (sub) {
        if (ctx->u_n) HIPCHK(hipMemcpy(sub, ctx->u_sub.p, ctx->u_n * 4, hipMemcpyDeviceToHost));
        memcpy(sub + ctx->u_n, ctx->cyc_sub.data(), ctx->cyc_sub.size() * 4);
    }
    if (flags) {
        if (ctx->u_n) HIPCHK(hipMemcpy(flags, ctx->u_flags.p, ctx->u_n, hipMemcpyDeviceToHost));
        memcpy(flags + ctx->u_n, ctx->cyc_flags.data(), ctx->cyc_flags.size());
    }
    return GGCAT_GPU_OK;
}

int ggcat_gpu_extension_copy(ggcat_gpu_ctx* ctx, uint8_t* ext) {
    if (!ctx || !ext) return GGCAT_GPU_EINVAL;
    if (!ctx->have_unitigs) return fail(ctx, GGCAT_GPU_ESTATE, "no unitig result");
    HIPCHK(hipSetDevice(ctx->device));
    uint64_t o = 0;
    for (uint32_t c = 0; c < ctx->n_chunks_used; c++) {
        if (!ctx->chunk_fill[c]) continue;
        HIPCHK(hipMemcpy(ext + o, P_<uint8_t>(ctx->u_ext) + (uint64_t)c * OUT_CHUNK, ctx->chunk_fill[c],
                         hipMemcpyDeviceToHost));
        o += ctx->chunk_fill[c];
    }
    if (ctx->n_tail)
        HIPCHK(hipMemcpy(ext + o, P_<uint8_t>(ctx->u_ext) + (uint64_t)ctx->n_chunks_used * OUT_CHUNK, ctx->n_tail,
                         hipMemcpyDeviceToHost));
    return GGCAT_GPU_OK;
}

}  // extern "C"

#include "noalloc.hip"
#include "query_graph.cpp"
