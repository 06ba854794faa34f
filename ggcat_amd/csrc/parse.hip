// parse.hip -- FASTA / FASTQ text to the joined layout on the device.
//
// The host reader (fasta.cpp) restates crates/io/src/sequences_reader.rs:
// 106-240 line by line over read blocks of `block` bytes (4 MiB in the
// reference).  For plain files the same records come out of a data-parallel
// pass over the raw file in HBM:
//   FASTA  a line starting with '>' is a header; every other line is sequence,
//          concatenated within a record with its '\n' (and a '\r' right before
//          it) removed; each header starts a new record (one '\n' separator).
//   FASTQ  line 4i + 1 is the sequence of record i; the other lines are
//          dropped; the sequence line's '\n' separates records.
// A '\r' is dropped where the host reader drops it: before a '\n', as the
// last byte of a read block, and as the file's last byte.  Empty records
// leave a bare '\n', which the bucketing reads exactly like no record.
//
// What falls back to the host reader (flags bit 0): a FASTA line starting with
// ';' (the reference's comment state swallows following lines depending on its
// read blocks, :127-128), and a FASTA line that is a lone '\r' at a block end
// (the reference sees no piece there and starts a new line after it).
//
// FASTA records whose sequence reaches the flush size F = max(block, 2(k-1))
// are cut by the reference into records that overlap by k-1 bases (:119-173):
// after each appended piece (a sequence line, or its part up to a block end)
// the pending sequence, once F bases or longer, is emitted and its last k-1
// bases start the next one.  Flags bit 1 says a record may be that long (two
// record starts, or a start and a file end, F raw bytes or more apart); the
// write pass then also marks every piece end in a bitmap over output
// positions, and flush_split_kernel / flush_expand_kernel insert the cuts.
//
// The file is cut in chunks of PARSE_CHUNK bytes.  Pass 1 gathers per chunk
// its newline count, its last newline and its first / last record start; a
// one-workgroup scan turns them into each chunk's carries (line number at
// its start, first character of the line its first byte belongs to, the
// previous record start); pass 2 counts the bytes each chunk keeps, a scan
// places them, pass 3 writes them after the output of the files before it
// (a device-side running offset, so files queue without host round trips).
#include "common.h"

namespace gg {

constexpr uint32_t PARSE_CHUNK = 4096;  // 256 threads x 16 bytes
constexpr uint64_t NO_POS64 = ~0ULL;

struct ChunkStats {
    uint64_t last_nl;  // position of the chunk's last '\n', NO_POS64 if none
    uint64_t first_rec, last_rec;  // FASTA record starts ('>' at a line start), NO_POS64 if none
    uint32_t n_nl;  // newlines in the chunk
    uint32_t host;  // needs the host reader (see the flags)
};

struct ChunkCarry {
    uint64_t line;  // newlines before the chunk
    uint64_t prev_rec;  // last record start before the chunk (NO_POS64: none)
    uint32_t first_char;  // first character of the line holding the chunk's first byte
    uint32_t pad;
};

struct ParseArgs {
    const uint8_t* raw;
    uint64_t n;
    uint32_t fastq;
    uint64_t n_chunks;
    uint64_t block;  // read block of the host reader
    uint64_t flush;  // FASTA flush size F
    ChunkStats* stats;
    ChunkCarry* carry;
    uint32_t* keep;  // kept bytes per chunk (count pass)
    const uint64_t* keep_off;  // their exclusive scan (write pass)
    uint8_t* out;  // joined output of every file
    uint64_t* base;  // output offset of this file (read by the write pass, advanced by parse_finish_kernel)
    unsigned long long* bits;  // piece ends, one bit per output position (marked when flags bit 1 is set)
    unsigned int* flags;  // this file's: bit 0 host reader needed, bit 1 flush possible
};

GG_DEV bool line_start(const ParseArgs& A, uint64_t i) { return i == 0 || A.raw[i - 1] == '\n'; }

GG_DEV bool drop_cr(const ParseArgs& A, uint64_t i) {
    return A.raw[i] == '\r' && (i + 1 == A.n || A.raw[i + 1] == '\n' || (i + 1) % A.block == 0);
}

__global__ __launch_bounds__(256) void parse_stats_kernel(ParseArgs A) {
    __shared__ unsigned long long s_last_nl, s_first_rec, s_last_rec;
    __shared__ unsigned int s_nl, s_host;
    for (uint64_t c = blockIdx.x; c < A.n_chunks; c += gridDim.x) {
        if (threadIdx.x == 0) {
            s_last_nl = 0;  // position + 1 (0: none)
            s_first_rec = ~0ULL;
            s_last_rec = 0;  // position + 1
            s_nl = 0;
            s_host = 0;
        }
        __syncthreads();
        const uint64_t p0 = c * PARSE_CHUNK + threadIdx.x * 16;
        uint32_t nl = 0, host = 0;
        unsigned long long last_nl = 0, first_rec = ~0ULL, last_rec = 0;
        for (uint32_t j = 0; j < 16; j++) {
            const uint64_t i = p0 + j;
            if (i >= A.n) break;
            const uint8_t ch = A.raw[i];
            if (ch == '\n') {
                nl++;
                last_nl = i + 1;
            }
            if (!A.fastq && line_start(A, i)) {
                if (ch == '>') {
                    if (first_rec == ~0ULL) first_rec = i;
                    last_rec = i + 1;
                }
                if (ch == ';') host = 1;
                if (ch == '\r' && (i + 1) % A.block == 0 && i + 1 < A.n && A.raw[i + 1] != '\n') host = 1;
            }
        }
        if (nl) atomicAdd(&s_nl, nl);
        if (host) atomicOr(&s_host, 1u);
        if (last_nl) atomicMax(&s_last_nl, last_nl);
        if (first_rec != ~0ULL) atomicMin(&s_first_rec, first_rec);
        if (last_rec) atomicMax(&s_last_rec, last_rec);
        __syncthreads();
        if (threadIdx.x == 0) {
            ChunkStats s;
            s.last_nl = s_last_nl ? s_last_nl - 1 : NO_POS64;
            s.first_rec = s_first_rec == ~0ULL ? NO_POS64 : s_first_rec;
            s.last_rec = s_last_rec ? s_last_rec - 1 : NO_POS64;
            s.n_nl = s_nl;
            s.host = s_host;
            A.stats[c] = s;
        }
        __syncthreads();
    }
}

// One workgroup of 1024 threads: each thread folds a contiguous run of
// chunks, the runs are combined in order, then every chunk gets its carry.
// The three carries are each an ordered fold: a sum (lines), "the last
// record start seen" and "the first character after the last newline seen".
__global__ __launch_bounds__(1024) void parse_carry_kernel(ParseArgs A) {
    __shared__ unsigned long long s_line[1024], s_rec[1024];
    __shared__ unsigned int s_char[1024];
    __shared__ unsigned int s_flag;
    const uint32_t t = threadIdx.x;
    const uint64_t per = (A.n_chunks + 1023) / 1024;
    const uint64_t c0 = t * per < A.n_chunks ? t * per : A.n_chunks;
    const uint64_t c1 = c0 + per < A.n_chunks ? c0 + per : A.n_chunks;
    // the run's own fold (identity: no newline / no record start)
    uint64_t lines = 0, rec = NO_POS64;
    uint32_t chr = 0x100;  // 0x100: no newline in the run
    for (uint64_t c = c0; c < c1; c++) {
        const ChunkStats s = A.stats[c];
        lines += s.n_nl;
        if (s.last_rec != NO_POS64) rec = s.last_rec;
        if (s.last_nl != NO_POS64) chr = s.last_nl + 1 < A.n ? A.raw[s.last_nl + 1] : 0u;
    }
    s_line[t] = lines;
    s_rec[t] = rec;
    s_char[t] = chr;
    if (t == 0) s_flag = 0;
    __syncthreads();
    // inclusive log-step scan of the run folds
    for (uint32_t d = 1; d < 1024; d <<= 1) {
        unsigned long long l = 0, r = NO_POS64;
        unsigned int ch = 0x100;
        if (t >= d) {
            l = s_line[t - d];
            r = s_rec[t - d];
            ch = s_char[t - d];
        }
        __syncthreads();
        if (t >= d) {
            s_line[t] += l;
            if (s_rec[t] == NO_POS64) s_rec[t] = r;
            if (s_char[t] == 0x100) s_char[t] = ch;
        }
        __syncthreads();
    }
    // carry into this run: the inclusive values of the previous run; the
    // file starts on a fresh line whose first character is raw[0]
    uint64_t line_in = t ? s_line[t - 1] : 0;
    uint64_t rec_in = t ? s_rec[t - 1] : NO_POS64;
    uint32_t chr_in = t ? s_char[t - 1] : 0x100;
    if (chr_in == 0x100) chr_in = A.n ? A.raw[0] : 0u;
    bool long_rec = false;
    for (uint64_t c = c0; c < c1; c++) {
        const ChunkStats s = A.stats[c];
        A.carry[c] = ChunkCarry{line_in, rec_in, chr_in, 0};
        if (!A.fastq) {
            // record starts F or more apart (the file's start counts as one)
            const uint64_t prev = rec_in == NO_POS64 ? 0 : rec_in;
            if (s.first_rec != NO_POS64 && s.first_rec - prev >= A.flush) long_rec = true;
        }
        line_in += s.n_nl;
        if (s.last_rec != NO_POS64) rec_in = s.last_rec;
        if (s.last_nl != NO_POS64) chr_in = s.last_nl + 1 < A.n ? A.raw[s.last_nl + 1] : 0u;
        if (s.host) atomicOr(&s_flag, 1u);
    }
    if (t == 1023 && !A.fastq) {
        const uint64_t prev = s_rec[1023] == NO_POS64 ? 0 : s_rec[1023];
        if (A.n - prev >= A.flush) long_rec = true;
    }
    if (long_rec) atomicOr(&s_flag, 2u);
    __syncthreads();
    if (t == 0 && s_flag) atomicOr(A.flags, s_flag);
}

// Output of byte i given its line's first character and line number:
// returns the byte to write, or -1 for nothing.
GG_DEV int parse_emit(const ParseArgs& A, uint64_t i, uint32_t first_char, uint64_t line) {
    const uint8_t ch = A.raw[i];
    if (A.fastq) {
        if ((line & 3) != 1) return -1;
        if (ch == '\n') return '\n';
        return drop_cr(A, i) ? -1 : ch;
    }
    if (first_char == '>') return line_start(A, i) ? '\n' : -1;  // a header line: one separator
    if (ch == '\n') return -1;
    return drop_cr(A, i) ? -1 : ch;
}

// Each thread walks its 16 bytes in order, tracking its line's first
// character and line number from the chunk carry and a block scan of the
// newline positions before it.
template <bool WRITE>
__global__ __launch_bounds__(256) void parse_emit_kernel(ParseArgs A) {
    __shared__ unsigned int s_cnt[256];
    __shared__ unsigned long long s_nl[256], s_lastnl[256];
    const bool mark = WRITE && !A.fastq && (*A.flags & 2u);
    const uint64_t file_base = WRITE ? *A.base : 0;
    for (uint64_t c = blockIdx.x; c < A.n_chunks; c += gridDim.x) {
        const uint64_t p0 = c * PARSE_CHUNK + threadIdx.x * 16;
        uint32_t nl = 0;
        unsigned long long lastnl = 0;  // position + 1
        for (uint32_t j = 0; j < 16; j++) {
            const uint64_t i = p0 + j;
            if (i < A.n && A.raw[i] == '\n') {
                nl++;
                lastnl = i + 1;
            }
        }
        s_nl[threadIdx.x] = nl;
        s_lastnl[threadIdx.x] = lastnl;
        __syncthreads();
        // exclusive: newlines and the last newline before this thread's bytes
        for (uint32_t d = 1; d < 256; d <<= 1) {
            unsigned long long a = 0, b = 0;
            if (threadIdx.x >= d) {
                a = s_nl[threadIdx.x - d];
                b = s_lastnl[threadIdx.x - d];
            }
            __syncthreads();
            if (threadIdx.x >= d) {
                s_nl[threadIdx.x] += a;
                if (b > s_lastnl[threadIdx.x]) s_lastnl[threadIdx.x] = b;
            }
            __syncthreads();
        }
        const ChunkCarry cy = A.carry[c];
        uint64_t line = cy.line + (threadIdx.x ? s_nl[threadIdx.x - 1] : 0);
        const unsigned long long before = threadIdx.x ? s_lastnl[threadIdx.x - 1] : 0;
        uint32_t first_char = before ? (before < A.n ? A.raw[before] : 0u) : cy.first_char;
        uint32_t o = 0;
        uint32_t ends = 0;  // bit j: a sequence piece ends after j kept bytes
        uint8_t bytes[16];
        for (uint32_t j = 0; j < 16; j++) {
            const uint64_t i = p0 + j;
            if (i >= A.n) break;
            const int e = parse_emit(A, i, first_char, line);
            if (e >= 0) bytes[o++] = (uint8_t)e;
            // a sequence line's pieces end at its '\n' and at block ends
            if (mark && first_char != '>' && (A.raw[i] == '\n' || (i + 1) % A.block == 0)) ends |= 1u << o;
            if (A.raw[i] == '\n') {
                line++;
                first_char = i + 1 < A.n ? A.raw[i + 1] : 0u;
            }
        }
        __syncthreads();
        if (!WRITE) {
            // kept bytes of the chunk
            s_cnt[threadIdx.x] = o;
            __syncthreads();
            for (uint32_t d = 128; d; d >>= 1) {
                if (threadIdx.x < d) s_cnt[threadIdx.x] += s_cnt[threadIdx.x + d];
                __syncthreads();
            }
            if (threadIdx.x == 0) A.keep[c] = s_cnt[0];
        } else {
            s_cnt[threadIdx.x] = o;
            __syncthreads();
            for (uint32_t d = 1; d < 256; d <<= 1) {
                const uint32_t v = threadIdx.x >= d ? s_cnt[threadIdx.x - d] : 0u;
                __syncthreads();
                s_cnt[threadIdx.x] += v;
                __syncthreads();
            }
            const uint64_t base = file_base + A.keep_off[c] + s_cnt[threadIdx.x] - o;
            for (uint32_t j = 0; j < o; j++) A.out[base + j] = bytes[j];
            while (ends) {
                const uint64_t q = base + (uint32_t)__builtin_ctz(ends);
                ends &= ends - 1;
                atomicOr(&A.bits[q >> 6], 1ULL << (q & 63));
            }
        }
        __syncthreads();
    }
}

// After a file's write pass: records where the file's output started, ends
// its last record with '\n' and advances the running offset.
__global__ void parse_finish_kernel(ParseArgs A, uint64_t* file_start) {
    if (threadIdx.x == 0) {
        const uint64_t s = *A.base;
        const uint64_t kept = A.n_chunks ? A.keep_off[A.n_chunks] : 0;
        *file_start = s;
        A.out[s + kept] = '\n';
        *A.base = s + kept + 1;
    }
}

// Zero fill of 64-bit words (the piece-end bitmap, the parse counters).
__global__ void parse_zero_kernel(unsigned long long* p, uint64_t n_words) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_words;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// First and last '\n' (position + 1, 0: none) of each `window` bytes of the
// joined output.  Records inside one window are shorter than the window; the
// host measures the others exactly between consecutive windows that hold a
// '\n', so with window <= F it finds every record of F bytes or more.
__global__ __launch_bounds__(256) void newline_summary_kernel(const uint8_t* out, uint64_t n, uint64_t window,
                                                              unsigned long long* first, unsigned long long* last) {
    __shared__ unsigned long long s_first, s_last;
    const uint64_t w0 = (uint64_t)blockIdx.x * window;
    if (threadIdx.x == 0) {
        s_first = ~0ULL;
        s_last = 0;
    }
    __syncthreads();
    unsigned long long f = ~0ULL, l = 0;
    for (uint64_t i = w0 + threadIdx.x; i < w0 + window && i < n; i += blockDim.x) {
        if (out[i] == '\n') {
            if (f == ~0ULL) f = i + 1;
            l = i + 1;
        }
    }
    if (f != ~0ULL) atomicMin(&s_first, f);
    if (l) atomicMax(&s_last, l);
    __syncthreads();
    if (threadIdx.x == 0) {
        first[blockIdx.x] = s_first == ~0ULL ? 0 : s_first;
        last[blockIdx.x] = s_last;
    }
}

// One wavefront per long record [a, a + L) of the joined output walks the
// reference's flush points.  With the pending sequence starting at s (record
// relative), the next cut is at the first piece end p >= s + F (the record's
// end is one); the output gets a record break there and the next pending
// sequence starts at p - (k - 1).  Cuts go to `cuts` (absolute positions, in
// no particular order).
__global__ __launch_bounds__(64) void flush_split_kernel(const uint64_t* rec_start, const uint64_t* rec_len,
                                                         uint32_t n_rec, const unsigned long long* bits,
                                                         uint64_t flush, uint64_t copyback, uint64_t* cuts,
                                                         unsigned int* n_cuts, uint32_t cap) {
    const uint32_t lane = threadIdx.x;
    for (uint32_t r = blockIdx.x; r < n_rec; r += gridDim.x) {
        const uint64_t a = rec_start[r], L = rec_len[r], end = a + L;
        uint64_t s = 0;
        while (s + flush <= L) {
            const uint64_t q = a + s + flush;
            uint64_t p = end;
            // first marked position in [q, end), 64 words per step
            for (uint64_t w0 = q >> 6; (w0 << 6) < end; w0 += 64) {
                const uint64_t w = w0 + lane;
                unsigned long long word = (w << 6) < end ? bits[w] : 0ULL;
                if (w == (q >> 6)) word &= ~0ULL << (q & 63);
                if ((w << 6) < end && end - (w << 6) < 64) word &= (1ULL << (end - (w << 6))) - 1;
                const unsigned long long hit = __ballot(word != 0ULL);
                if (hit) {
                    const uint32_t l = (uint32_t)__builtin_ctzll(hit);
                    const unsigned long long wl = __shfl(word, (int)l);
                    p = ((w0 + l) << 6) + (uint64_t)__builtin_ctzll(wl);
                    break;
                }
            }
            if (lane == 0) {
                const unsigned int slot = atomicAdd(n_cuts, 1u);
                if (slot < cap) cuts[slot] = p;
            }
            s = p - a - copyback;
        }
    }
}

// Copy of the joined output with every cut applied: a byte at p moves by
// (1 + copyback) times the cuts at or before it; cuts must be sorted.
__global__ __launch_bounds__(256) void flush_expand_kernel(const uint8_t* src, uint64_t n, const uint64_t* cuts,
                                                           uint32_t n_cuts, uint64_t copyback, uint8_t* dst) {
    for (uint64_t p0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 16; p0 < n;
         p0 += (uint64_t)gridDim.x * blockDim.x * 16) {
        uint32_t lo = 0, hi = n_cuts;  // cuts at or before p0
        while (lo < hi) {
            const uint32_t mid = (lo + hi) / 2;
            if (cuts[mid] <= p0)
                lo = mid + 1;
            else
                hi = mid;
        }
        uint32_t j = lo;
        for (uint64_t p = p0; p < p0 + 16 && p < n; p++) {
            while (j < n_cuts && cuts[j] <= p) j++;
            dst[p + (uint64_t)j * (1 + copyback)] = src[p];
        }
    }
}

// The inserted bytes: cut j (sorted) leaves a gap of 1 + copyback bytes at
// cuts[j] + j(1 + copyback), filled with '\n' and the copyback bases before
// the cut.
__global__ __launch_bounds__(256) void flush_insert_kernel(const uint8_t* src, const uint64_t* cuts, uint32_t n_cuts,
                                                           uint64_t copyback, uint8_t* dst) {
    for (uint32_t j = blockIdx.x; j < n_cuts; j += gridDim.x) {
        const uint64_t at = cuts[j] + (uint64_t)j * (1 + copyback);
        for (uint64_t t = threadIdx.x; t <= copyback; t += blockDim.x)
            dst[at + t] = t ? src[cuts[j] - copyback + (t - 1)] : (uint8_t)'\n';
    }
}

}  // namespace gg
