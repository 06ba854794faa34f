// merge.hip -- per-sub-bucket canonical k-mer counting on gfx950.
//
// Restates, per sub-bucket (bucket x second bucket, the unit of one
// kmers_transform job, crates/kmers_transform/src/lib.rs:277-386):
//   crates/assembler_kmers_merge/src/unitigs_extender/hashmap.rs:361-431
//     add_sequence: per k-mer occurrence, count += multiplicity and
//     flags |= (begin_ignored << !fw) | (end_ignored << fw)
//   crates/structs/src/map_entry.rs:79-83
//     multiplicity = count >> (flags == INCL_BEGIN|INCL_END)
//   hashmap.rs:134-137 solid filter: multiplicity >= min_multiplicity
// k-mer keys: crates/hashes/src/base/cn_seqhash_base.rs:22-72 (canonical
// 2-bit seq-hash, u64 for k <= 32, u128 for k <= 64) or
// crates/hashes/src/base/cn_rkhash_base.rs:65-110 (rabin-karp128).
//
// Work items are runs of physical units of one sub-bucket.  A k-mer shared by
// two groups of the same sub-bucket stored in different work items is seen
// half in each: both halves carry the same counter (every occurrence of such
// a k-mer is a group boundary, so it appears once on each side) and one flag
// bit each.  Those entries carry a cross mark and are combined afterwards
// (flags ORed), which yields exactly the sub-bucket's MapEntry.
//
// Table entry value: bits 0..28 counter, bit 29 cross mark, bits 30..31 the
// two MapEntry flags; one u32 atomicAdd + atomicOr update an entry.
#include "common.h"

namespace gg {

constexpr uint64_t EMPTY64 = ~0ULL;
constexpr uint32_t CNT_MASK = 0x1FFFFFFFu;
constexpr uint32_t VAL_CROSS = 1u << 29;

struct WorkItem {
    uint32_t u0;  // first physical unit
    uint32_t dup;  // 1: a unit of the item may still hold identical records (not deduplicated by the compaction)
    uint32_t occ;  // k-mer occurrences
    uint32_t sub;  // sub-bucket
    uint64_t r0, r1;  // its slots [r0, r1)
};

// One pass over a work item restricted to the k-mers whose partition hash
// falls in `part` of 2^log_parts.  Items whose distinct k-mers do not fit one
// LDS table are processed as several such passes.
struct PartWork {
    uint32_t item;
    uint32_t part;
    uint32_t log_parts;
    uint32_t occ;  // the item's k-mer occurrences (set on redo entries)
};
constexpr uint32_t MAX_LOG_PARTS = 16;

struct MergeArgs {
    const uint32_t* slots;
    uint32_t slot_words;
    const uint32_t* pool;
    uint32_t k;
    uint32_t min_mult;
    uint32_t kind;  // 0 = canonical seq-hash, 1 = canonical rabin-karp128
    uint32_t key_words;  // 1 (u64 keys) or 2 (u128 keys)
    uint32_t lds_cap;  // table entries in LDS (power of two)
    uint32_t lds_try;  // first pass: items with more occurrences go straight to partitioned passes
    uint32_t table_pct;  // table slots per 100 occurrences of a whole-item pass (power-of-two rounded)
    uint32_t dedup;  // identical records of an entry are inserted once with their count
    uint32_t tag_kmers;  // dedup: every k-mer position of the distinct records is a thread's work item
                         // (seq-hash, k <= 32); otherwise each distinct record's k-mers are rolled by one thread
    uint32_t rec_lds_off;  // byte offset of the record table in dynamic LDS
    uint32_t rec_cap;  // record table slots
    uint32_t rec_direct;  // items without dup: records placed in the record table by position, not hashed
    uint32_t claim;  // work entries claimed per atomic on the work cursor
    const WorkItem* items;
    const PartWork* parts;  // null: first pass, one work entry per item
    uint32_t n_work;  // entries of this pass
    unsigned int* work;  // work-stealing cursor
    PartWork* redo;  // work entries whose table overflowed (at most n_work)
    unsigned int* redo_count;
    uint64_t* out_key;  // key_words per entry, chunked (see OUT_CHUNK)
    uint32_t* out_val;  // multiplicity | flags << 30
    uint32_t* out_sub;
    unsigned int* chunk_cursor;
    uint32_t* chunk_fill;
    uint32_t n_chunks;
    uint64_t* fix_key;  // cross-marked entries awaiting their other half
    uint32_t* fix_val;
    uint32_t* fix_sub;
    unsigned long long* fix_cursor;
    uint64_t fix_cap;
    uint64_t pool_words;  // valid words in the long-record pool
    uint32_t* error_flag;  // bit 1: malformed slot
    // colored builds: per-slot color, color bitmap words per entry (0 when
    // uncolored), and the bitmaps of the output and fixup entries
    const uint32_t* slot_color;
    uint32_t color_words;
    uint32_t* out_colors;
    uint32_t* fix_colors;
    // rabin-karp128 constants (crates/hashes/src/cn_rkhash.rs:65-78)
    u128_t rk_mult, rk_mult_inv, rk_rmmult;
    u128_t rk_fwd[4], rk_bkw[4];
    u128_t rk_fwd_rm[4], rk_bkw_rm[4];  // rk_fwd[c] * rk_rmmult, rk_bkw[c] * rk_rmmult
    // per base position i < k - 1 of a k-mer and base c: entry 2(4i + c) is
    // rk_fwd[c] * rk_mult^(k-2-i), entry 2(4i + c) + 1 is rk_bkw[c] * rk_mult^(i+1)
    const u128_t* rk_pos;
};

GG_DEV uint32_t hash128(uint64_t lo, uint64_t hi) { return tab_hash64(lo ^ (hi * 0x9E3779B97F4A7C15ULL)); }

// Kernel variants: MODE 0 reads every switch from the arguments; MODE 1 is
// the seq-hash k <= 32 path (64-bit keys, 4-word slots, record dedup with one
// k-mer position per thread, no colors) and MODE 2 the uncolored
// rabin-karp128 path (128-bit keys, 8-word slots, record dedup with each
// distinct record rolled by one thread), with those switches compiled in,
// which keeps the other paths' code out of their register allocation.
template <int MODE>
struct Mode {
    static constexpr bool S64 = MODE == 1;  // seq-hash k <= 32, 4-word slots, uncolored
    static GG_DEV uint32_t key_words(const MergeArgs& A) { return S64 ? 1u : MODE == 2 ? 2u : A.key_words; }
    static GG_DEV uint32_t color_words(const MergeArgs& A) { return MODE != 0 ? 0u : A.color_words; }
    static GG_DEV bool seq64(const MergeArgs& A) { return S64 || (MODE == 0 && A.kind == 0 && A.k <= 32); }
    static GG_DEV bool seq128(const MergeArgs& A) { return MODE == 0 && A.kind == 0; }
    static GG_DEV bool dedup(const MergeArgs& A) { return S64 || A.dedup; }
    static GG_DEV bool tag_kmers(const MergeArgs& A) { return S64 || (MODE == 0 && A.tag_kmers); }
    static GG_DEV uint32_t slot_words(const MergeArgs& A) { return S64 ? 4u : MODE == 2 ? 8u : A.slot_words; }
    static constexpr uint32_t MAX_SW = S64 ? 4u : 8u;  // register copies of a slot
};

// Insert one occurrence.  Returns 1 when this call claimed a new slot, 0 when
// the key was present, 2 when no slot was found within min(mask+1,
// probe_limit) probes.
template <class KP, class VP>
GG_DEV uint32_t tab_insert64(KP keys, VP vals, uint32_t mask, uint64_t key, uint32_t bits, uint32_t probe_limit,
                             uint32_t& slot) {
    uint32_t i = tab_hash64(key) & mask;
    uint32_t claimed = 0;
    for (uint32_t probes = 0;; probes++) {
        if (probes > mask || probes >= probe_limit) return 2;
        uint64_t cur = keys[i];
        if (cur == key) break;
        if (cur == EMPTY64) {
            uint64_t old = atomicCAS((unsigned long long*)&keys[i], (unsigned long long)EMPTY64,
                                     (unsigned long long)key);
            if (old == EMPTY64) {
                claimed = 1;
                break;
            }
            if (old == key) break;
        }
        i = (i + 1) & mask;
    }
    atomicAdd((uint32_t*)&vals[i], bits & CNT_MASK);
    if (bits & ~CNT_MASK) atomicOr((uint32_t*)&vals[i], bits & ~CNT_MASK);
    slot = i;
    return claimed;
}

// 128-bit keys: a slot is claimed through its state word (0 free, 1 being
// written, 2 ready).  The loop body runs the writer and the waiters in the
// same iteration, so lanes of one wavefront cannot wait on each other; the
// wait on a slot in state 1 is bounded.
template <class KP, class SP, class VP>
GG_DEV uint32_t tab_insert128(KP klo, KP khi, SP st, VP vals, uint32_t mask, uint64_t lo, uint64_t hi,
                              uint32_t bits, uint32_t probe_limit, uint32_t& slot) {
    uint32_t i = hash128(lo, hi) & mask;
    uint32_t claimed = 0, probes = 0, spins = 0;
    bool done = false;
    while (!done) {
        uint32_t s = __hip_atomic_load((uint32_t*)&st[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (s == 0) {
            if (atomicCAS((uint32_t*)&st[i], 0u, 1u) == 0u) {
                klo[i] = lo;
                khi[i] = hi;
                __threadfence_block();  // the table is in LDS: workgroup scope
                atomicExch((uint32_t*)&st[i], 2u);
                claimed = 1;
                done = true;
            }
        } else if (s == 2) {
            uint64_t a = __hip_atomic_load((uint64_t*)&klo[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            uint64_t b = __hip_atomic_load((uint64_t*)&khi[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (a == lo && b == hi) {
                done = true;
            } else {
                i = (i + 1) & mask;
                if (++probes > mask || probes >= probe_limit) return 2;
            }
        } else if (++spins > (1u << 20)) {
            return 2;
        }
    }
    atomicAdd((uint32_t*)&vals[i], bits & CNT_MASK);
    if (bits & ~CNT_MASK) atomicOr((uint32_t*)&vals[i], bits & ~CNT_MASK);
    slot = i;
    return claimed;
}

// Rabin-karp128 hashes (cn_rkhash_base.rs:65-110) of the k-mers with index
// [i0, i1) of packed bases bw; cb(lo, hi, is_forward, idx).  The state over
// the first k - 1 bases of k-mer i0, fh = sum fwd[c] M^(k-2-i) and
// bk = sum bkw[c] M^(i+1) over its base offsets i (what the Horner loops of
// the reference's init compute), is summed from the per-position table:
// independent loads and 128-bit adds instead of two chains of k - 1
// dependent 128-bit multiplies; then each k-mer is one rolling step.
template <class CB>
GG_DEV void rk_segment(const MergeArgs& A, const uint32_t* bw, uint32_t i0, uint32_t i1, CB&& cb) {
    const uint32_t k = A.k;
    auto code = [&](uint32_t i) -> uint32_t { return (bw[i >> 4] >> (2 * (i & 15))) & 3; };
    u128_t fh = 0, bk = 0;
    for (uint32_t i = 0; i + 1 < k; i++) {
        const u128_t* e = A.rk_pos + 2 * (4 * i + code(i0 + i));
        fh += e[0];
        bk += e[1];
    }
    for (uint32_t x = i0; x < i1; x++) {
        const uint32_t in = code(x + k - 1), out = code(x);
        const u128_t cur_f = fh * A.rk_mult + A.rk_fwd[in];
        fh = cur_f - A.rk_fwd_rm[out];
        const u128_t cur_b = bk * A.rk_mult_inv + A.rk_bkw_rm[in];
        bk = cur_b - A.rk_bkw[out];
        const u128_t key = cur_f < cur_b ? cur_f : cur_b;
        cb((uint64_t)key, (uint64_t)(key >> 64), cur_f < cur_b, x);
    }
}

// k-mers per segment when the k-mers of rabin-karp128 records are spread
// over the threads of a workgroup (each segment starts from the table)
constexpr uint32_t RK_SEG = 16;

// Iterate the k-mers of one slot; cb(lo, hi, is_forward, idx, last_idx).
template <int MODE, class CB>
GG_DEV void for_each_kmer(const MergeArgs& A, const uint32_t* slot, CB&& cb) {
    const uint32_t meta = slot[0];
    const uint32_t* bw;
    uint32_t len;
    if (meta & META_LONG) {
        len = slot[1];
        uint64_t off = (uint64_t)slot[2] | ((uint64_t)slot[3] << 32);
        if (off + ((uint64_t)len + 15) / 16 > A.pool_words) {
            atomicOr(A.error_flag, 2u);
            return;
        }
        bw = A.pool + off;
    } else {
        len = meta_len(meta);
        if (len > (Mode<MODE>::slot_words(A) - 1) * 16) {
            atomicOr(A.error_flag, 2u);
            return;
        }
        bw = slot + 1;
    }
    const uint32_t k = A.k;
    if (len < k) return;
    const uint32_t last = len - k;
    if (Mode<MODE>::seq64(A)) {
        const uint64_t mask = (k == 32) ? ~0ULL : ((1ULL << (2 * k)) - 1);
        uint64_t fw = 0, rc = 0;
        uint32_t word = 0;
        for (uint32_t i = 0; i < len; i++) {
            if ((i & 15) == 0) word = bw[i >> 4];
            uint64_t c = (word >> (2 * (i & 15))) & 3;
            fw = (fw >> 2) | (c << (2 * (k - 1)));
            rc = ((rc << 2) | (c ^ 2)) & mask;
            if (i + 1 >= k) cb(fw < rc ? fw : rc, 0ULL, fw < rc, i + 1 - k, last);
        }
    } else if (Mode<MODE>::seq128(A)) {
        const u128_t mask = (k == 64) ? ~(u128_t)0 : (((u128_t)1 << (2 * k)) - 1);
        u128_t fw = 0, rc = 0;
        uint32_t word = 0;
        for (uint32_t i = 0; i < len; i++) {
            if ((i & 15) == 0) word = bw[i >> 4];
            u128_t c = (word >> (2 * (i & 15))) & 3;
            fw = (fw >> 2) | (c << (2 * (k - 1)));
            rc = ((rc << 2) | (c ^ 2)) & mask;
            if (i + 1 >= k) {
                u128_t key = fw < rc ? fw : rc;
                cb((uint64_t)key, (uint64_t)(key >> 64), fw < rc, i + 1 - k, last);
            }
        }
    } else {
        rk_segment(A, bw, 0, last + 1,
                   [&](uint64_t lo, uint64_t hi, bool fwd, uint32_t x) { cb(lo, hi, fwd, x, last); });
    }
}

// Value bits contributed by one occurrence (count 1, flags, cross mark).
GG_DEV uint32_t occurrence_bits(uint32_t meta, bool is_fwd, uint32_t idx, uint32_t last) {
    const uint32_t rflags = meta_flags(meta);
    bool begin_ignored = (rflags & 1) == 0 && idx == 0;
    bool end_ignored = (rflags & 2) == 0 && idx == last;
    uint32_t fl = ((begin_ignored ? 1u : 0u) << (is_fwd ? 0 : 1)) | ((end_ignored ? 1u : 0u) << (is_fwd ? 1 : 0));
    bool cross = (idx == 0 && (meta & META_XBEGIN)) || (idx == last && (meta & META_XEND));
    return 1u | (cross ? VAL_CROSS : 0u) | (fl << 30);
}

// Output is written in chunks of OUT_CHUNK entries.  Each wavefront owns one
// open chunk at a time (reserved with one atomic), fills it without any
// workgroup barrier and records its fill count when it moves on, so solid
// entries cost no global atomics.  Entries of chunk c live at
// [c * OUT_CHUNK, c * OUT_CHUNK + fill[c]).
constexpr uint32_t OUT_CHUNK = 4096;

// A wavefront's open chunk; every lane holds the same values.
struct EmitState {
    uint32_t chunk;  // open chunk, ~0u when none
    uint32_t fill;
};

GG_DEV void close_chunk(const MergeArgs& A, EmitState& es) {
    if ((threadIdx.x & 63) == 0 && es.chunk != ~0u && es.chunk < A.n_chunks) A.chunk_fill[es.chunk] = es.fill;
    es.chunk = ~0u;
    es.fill = 0;
}

// Emit solid entries among the table slots claimed by the entry (`claimed`,
// n of them); cross-marked entries holding only one flag bit go to the
// fixup list.  Wavefronts take 64 claimed slots at a time.
template <int MODE, class KP, class SP, class VP>
GG_DEV void emit_solid(const MergeArgs& A, uint32_t sub, KP klo, KP khi, SP st, VP vals, const uint32_t* cbits,
                       const uint16_t* claimed, uint32_t n, EmitState& es) {
    const uint32_t W = Mode<MODE>::color_words(A), KW = Mode<MODE>::key_words(A);
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t below = (1ULL << lane) - 1;
    for (uint32_t i0 = wave * 64; i0 < n; i0 += blockDim.x) {
        const uint32_t j = i0 + lane;
        const uint32_t i = j < n ? claimed[j] : 0u;
        bool solid = false, fix = false;
        uint32_t outv = 0;
        if (j < n) {
            bool occupied = (KW == 1) ? (klo[i] != EMPTY64) : (st[i] == 2);
            if (occupied) {
                uint32_t v = vals[i];
                uint32_t cnt = v & CNT_MASK, fl = v >> 30;
                uint32_t mult = cnt >> (fl == 3 ? 1 : 0);
                solid = mult >= A.min_mult;
                fix = solid && (v & VAL_CROSS) && (fl == 1 || fl == 2);
                outv = mult | (fl << 30);
            }
        }
        const bool main = solid && !fix;
        const uint64_t bal = __ballot(main), balf = __ballot(fix);
        const uint32_t cnt = (uint32_t)__popcll(bal);
        if (cnt) {
            if (es.chunk == ~0u || es.fill + cnt > OUT_CHUNK) {
                close_chunk(A, es);
                uint32_t c = 0;
                if (lane == 0) c = atomicAdd(A.chunk_cursor, 1u);
                es.chunk = __shfl(c, 0);
            }
            if (main && es.chunk < A.n_chunks) {
                const uint64_t pos = (uint64_t)es.chunk * OUT_CHUNK + es.fill + __popcll(bal & below);
                if (KW == 1) {
                    A.out_key[pos] = klo[i];
                } else {
                    A.out_key[2 * pos] = klo[i];
                    A.out_key[2 * pos + 1] = khi[i];
                }
                A.out_val[pos] = outv;
                A.out_sub[pos] = sub;
                for (uint32_t q = 0; q < W; q++) A.out_colors[pos * W + q] = cbits[i * W + q];
            }
            es.fill += cnt;
        }
        if (balf) {
            unsigned long long basef = 0;
            if (lane == 0) basef = atomicAdd(A.fix_cursor, (unsigned long long)__popcll(balf));
            basef = __shfl(basef, 0);
            if (fix) {
                uint64_t pos = basef + __popcll(balf & below);
                if (pos < A.fix_cap) {
                    if (KW == 1) {
                        A.fix_key[pos] = klo[i];
                    } else {
                        A.fix_key[2 * pos] = klo[i];
                        A.fix_key[2 * pos + 1] = khi[i];
                    }
                    A.fix_val[pos] = outv;
                    A.fix_sub[pos] = sub;
                    for (uint32_t q = 0; q < W; q++) A.fix_colors[pos * W + q] = cbits[i * W + q];
                }
            }
        }
    }
}

// Inserts every occurrence of the item's records whose k-mer lies in the
// work entry's partition (bits 16.. of the table hash; the slot index uses
// the low bits).  An insert that needs more than `probe_limit` probes sets
// *overflow and insertion stops (the entry is then redone as finer parts).
template <int MODE, class KP, class SP, class VP>
GG_DEV void insert_item(const MergeArgs& A, uint64_t r0, uint64_t r1, const PartWork& w, KP klo, KP khi, SP st,
                        VP vals, uint32_t* cbits, uint32_t mask, uint32_t probe_limit, uint32_t* overflow,
                        uint16_t* claimed, uint32_t* n_claimed) {
    const uint32_t W = Mode<MODE>::color_words(A), KW = Mode<MODE>::key_words(A);
    const uint32_t SW = A.slot_words;
    const uint32_t part_mask = (1u << w.log_parts) - 1;
    for (uint64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
        const uint32_t* slot = A.slots + r * SW;
        const uint32_t meta = slot[0];
        const uint32_t color = W ? A.slot_color[r] : 0u;
        const uint32_t extra = meta_mult(meta) - 1;  // identical copies stored once
        for_each_kmer<MODE>(A, slot, [&](uint64_t lo, uint64_t hi, bool is_fwd, uint32_t idx, uint32_t last) {
            if (*(volatile uint32_t*)overflow) return;
            if (part_mask) {
                const uint32_t h = (KW == 1) ? tab_hash64(lo) : hash128(lo, hi);
                if (((h >> 16) & part_mask) != w.part) return;
            }
            uint32_t bits = occurrence_bits(meta, is_fwd, idx, last) + extra, at = 0;
            uint32_t res = (KW == 1)
                               ? tab_insert64(klo, vals, mask, lo, bits, probe_limit, at)
                               : tab_insert128(klo, khi, st, vals, mask, lo, hi, bits, probe_limit, at);
            if (res == 2) {
                atomicOr(overflow, 1u);
                return;
            }
            if (res == 1) claimed[atomicAdd(n_claimed, 1u)] = (uint16_t)at;
            if (W) atomicOr(&cbits[at * W + (color >> 5)], 1u << (color & 31));
        });
    }
}

// ---- identical-record path (seq-hash keys, k <= 32) -----------------------
// The reference merges identical super-k-mers of a bucket before counting
// (crates/minimizer_bucketing/src/compactor.rs:128-196, multiplicities
// summed, :159-163); counting each distinct record once with its multiplicity
// gives the same MapEntries.  Records are keyed by their whole slot (meta
// word with length, flags, minimizer position and cross marks, then the
// packed bases) plus their color; a record the table cannot take is counted
// on its own.  After deduplication every thread takes k-mer positions of the
// distinct records, so the k-mer work is spread evenly over the workgroup.
// Record table slots (a power of two, 16..512); chunks of rec_cap / 2
// records are deduplicated at a time.
constexpr uint32_t REC_CAP_DEFAULT = 512;

struct RecTable {
    uint32_t* key;  // rec_cap * SW
    uint32_t* cnt;
    uint32_t* state;  // 0 free, 1 being written, 2 ready
    uint32_t* col;
    uint32_t* n;  // [1]: k-mers of the chunk, [2..5]: wave totals of the scan
    uint16_t* rec_of;  // each k-mer position of the chunk: record slot << 7 | k-mer index in the record
};

// k-mers of one inline record at most.
__host__ __device__ inline uint32_t rec_max_kmers(uint32_t slot_words, uint32_t k) {
    return (slot_words - 1) * 16 >= k ? (slot_words - 1) * 16 - k + 1 : 0;
}

GG_DEV RecTable rec_table(const MergeArgs& A, uint8_t* lds) {
    RecTable R;
    uint32_t* b = reinterpret_cast<uint32_t*>(lds + A.rec_lds_off);
    R.key = b;
    b += A.rec_cap * A.slot_words;
    R.cnt = b;
    b += A.rec_cap;
    R.state = b;
    b += A.rec_cap;
    R.col = b;
    b += A.rec_cap;
    R.n = b;
    b += 6;
    R.rec_of = reinterpret_cast<uint16_t*>(b);
    return R;
}

// Bytes of dynamic LDS the record table needs (k = 0xFFFF: no per-k-mer tags,
// the rolling mode of 128-bit keys).
// With rk_seg the tags are per segment of RK_SEG k-mers, for up to rec_cap
// records (the uncolored rabin-karp128 variant).
__host__ __device__ inline uint32_t rec_table_bytes(uint32_t slot_words, uint32_t k, uint32_t rec_cap,
                                                    bool rk_seg = false) {
    const uint32_t tags = rk_seg ? rec_cap * ((rec_max_kmers(slot_words, k) + RK_SEG - 1) / RK_SEG)
                                 : (rec_cap / 2) * rec_max_kmers(slot_words, k);
    return (rec_cap * slot_words + 3 * rec_cap + 6) * 4 + tags * 2 + 4;
}

GG_DEV uint32_t rec_hash(const uint32_t* w, uint32_t sw, uint32_t color) {
    uint64_t h = 0x9E3779B97F4A7C15ULL ^ color;
    for (uint32_t i = 0; i < sw; i++) h = (h ^ w[i]) * 0xbf58476d1ce4e5b9ULL;
    h ^= h >> 31;
    return (uint32_t)h;
}

// Canonical 2-bit k-mer (k <= 32) at base offset i of packed words (16 bases
// per word, base j at bits 2(j % 16)); *fwd when the forward strand is the
// smaller (the rolling path's fw < rc).
GG_DEV uint64_t kmer_at(const uint32_t* bw, uint32_t i, uint32_t k, bool* fwd) {
    const uint32_t wi = i >> 4, sh = (i & 15) * 2;
    const uint64_t lo = (uint64_t)bw[wi] | ((uint64_t)bw[wi + 1] << 32);
    uint64_t fw = sh ? (lo >> sh) | ((uint64_t)bw[wi + 2] << (64 - sh)) : lo;
    const uint64_t mask = (k == 32) ? ~0ULL : ((1ULL << (2 * k)) - 1);
    fw &= mask;
    // reverse complement: complement every base, reverse the 2-bit groups
    uint64_t x = fw ^ 0xAAAAAAAAAAAAAAAAULL;
    x = ((uint64_t)__brev((uint32_t)x) << 32) | __brev((uint32_t)(x >> 32));
    x = ((x >> 1) & 0x5555555555555555ULL) | ((x & 0x5555555555555555ULL) << 1);
    const uint64_t rc = x >> (64 - 2 * k);
    *fwd = fw < rc;
    return fw < rc ? fw : rc;
}

template <int MODE, class KP, class SP, class VP>
GG_DEV void insert_item_dedup(const MergeArgs& A, uint64_t r0, uint64_t r1, const PartWork& w, KP klo, KP khi, SP st,
                              VP vals, uint32_t* cbits, uint32_t mask, uint32_t probe_limit, uint32_t* overflow,
                              uint16_t* claimed, uint32_t* n_claimed, const RecTable& R, const uint32_t* pre,
                              uint32_t pre_color, bool hashed) {
    const uint32_t W = Mode<MODE>::color_words(A), KW = Mode<MODE>::key_words(A);
    const uint32_t SW = Mode<MODE>::slot_words(A), k = A.k, RC = A.rec_cap;
    const uint32_t part_mask = (1u << w.log_parts) - 1;
    // one occurrence (or cnt identical ones) of a k-mer of either key width
    auto insert_key = [&](uint64_t lo, uint64_t hi, uint32_t bits, uint32_t color) {
        if (*(volatile uint32_t*)overflow) return;
        if (part_mask) {
            const uint32_t h = (KW == 1) ? tab_hash64(lo) : hash128(lo, hi);
            if (((h >> 16) & part_mask) != w.part) return;
        }
        uint32_t at = 0;
        const uint32_t res = (KW == 1) ? tab_insert64(klo, vals, mask, lo, bits, probe_limit, at)
                                                : tab_insert128(klo, khi, st, vals, mask, lo, hi, bits, probe_limit, at);
        if (res == 2) {
            atomicOr(overflow, 1u);
            return;
        }
        if (res == 1) claimed[atomicAdd(n_claimed, 1u)] = (uint16_t)at;
        if (W) atomicOr(&cbits[at * W + (color >> 5)], 1u << (color & 31));
    };
    // the record table's states are all free here (cleared at kernel start
    // and by phase B of the previous chunk)
    // phase A: distinct records (long records and table misses go direct);
    // thread t holds record r0 + t, loaded by the caller (pre, pre_color)
    // while the previous chunk was processed
    {
        const uint64_t r = r0 + threadIdx.x;
        if (r >= r1) goto phase_a_done;
        const uint32_t* slot = A.slots + r * SW;
        const uint32_t meta = pre[0];
        const uint32_t color = pre_color;
        bool direct = (meta & META_LONG) != 0;
        if (!direct && !hashed) {
            // the item's records are distinct: thread t's record takes the
            // first of the table slots thread t reads in phase B (t < chunk
            // <= min(rec_cap, blockDim); both powers of two)
            const uint32_t i = RC > blockDim.x ? threadIdx.x * (RC / blockDim.x) : threadIdx.x;
            for (uint32_t q = 0; q < SW; q++) R.key[i * SW + q] = pre[q];
            R.col[i] = color;
            R.cnt[i] = meta_mult(meta);
            R.state[i] = 2;
        } else if (!direct) {
            uint32_t wv[Mode<MODE>::MAX_SW];
            for (uint32_t q = 0; q < SW; q++) wv[q] = pre[q];
            uint32_t i = rec_hash(wv, SW, color) & (RC - 1);
            uint32_t probes = 0, spins = 0;
            bool done = false;
            while (!done) {
                const uint32_t sv = __hip_atomic_load(&R.state[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (sv == 0) {
                    if (atomicCAS(&R.state[i], 0u, 1u) == 0u) {
                        for (uint32_t q = 0; q < SW; q++) R.key[i * SW + q] = wv[q];
                        R.col[i] = color;
                        R.cnt[i] = meta_mult(meta);
                        __threadfence_block();
                        atomicExch(&R.state[i], 2u);
                        done = true;
                    }
                } else if (sv == 2) {
                    bool eq = R.col[i] == color;
                    for (uint32_t q = 0; q < SW && eq; q++) eq = R.key[i * SW + q] == wv[q];
                    if (eq) {
                        atomicAdd(&R.cnt[i], meta_mult(meta));
                        done = true;
                    } else if (++probes >= 32) {
                        direct = true;
                        done = true;
                    } else {
                        i = (i + 1) & (RC - 1);
                    }
                } else if (++spins > (1u << 16)) {
                    direct = true;
                    done = true;
                }
            }
        }
        if (direct) {
            const uint32_t extra = meta_mult(meta) - 1;
            for_each_kmer<MODE>(A, slot, [&](uint64_t lo, uint64_t hi, bool is_fwd, uint32_t idx, uint32_t last) {
                insert_key(lo, hi, occurrence_bits(meta, is_fwd, idx, last) + extra, color);
            });
        }
    }
phase_a_done:
    __syncthreads();
    if (!Mode<MODE>::tag_kmers(A) && MODE != 2) {
        // 128-bit keys: each thread rolls the k-mers of the distinct records
        // in its table slots (rec_cap / blockDim of them) and frees the slots
        const uint32_t per = RC > blockDim.x ? RC / blockDim.x : 1u;
        const uint32_t s0 = threadIdx.x * per;
        for (uint32_t q = 0; q < per; q++) {
            const uint32_t si = s0 + q;
            if (si >= RC || R.state[si] != 2) continue;
            const uint32_t* rw = R.key + si * SW;
            const uint32_t cnt = R.cnt[si], color = W ? R.col[si] : 0u, meta = rw[0];
            for_each_kmer<MODE>(A, rw, [&](uint64_t lo, uint64_t hi, bool is_fwd, uint32_t idx, uint32_t last) {
                insert_key(lo, hi, occurrence_bits(meta, is_fwd, idx, last) - 1u + cnt, color);
            });
            R.state[si] = 0;
        }
        return;
    }
    // phase B: one exclusive scan over the table slots of their k-mer counts
    // places the distinct records' k-mers (each position tagged with its
    // record's slot and k-mer index; the rabin-karp128 variant tags segments
    // of RK_SEG k-mers with their index instead); each thread owns
    // rec_cap / blockDim consecutive slots (at least one) and frees them for
    // the next chunk once read.
    {
        const uint32_t per = RC > blockDim.x ? RC / blockDim.x : 1u;
        const uint32_t s0 = threadIdx.x * per;
        const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        uint32_t pk[2] = {0u, 0u};  // per <= 2 (rec_cap <= 512, 256 threads)
        uint32_t mine = 0;
        for (uint32_t q = 0; q < per; q++) {
            const uint32_t si = s0 + q;
            if (si < RC && R.state[si] == 2) {
                const uint32_t len = meta_len(R.key[si * SW]);
                const uint32_t nk = len >= k ? len - k + 1 : 0u;
                pk[q] = MODE == 2 ? (nk + RK_SEG - 1) / RK_SEG : nk;
                R.state[si] = 0;
            }
            mine += pk[q];
        }
        uint32_t inc = mine;
        for (uint32_t d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(inc, d);
            if (lane >= d) inc += y;
        }
        if (lane == 63) R.n[2 + wave] = inc;  // wave totals
        __syncthreads();
        uint32_t base = 0, all = 0;
        const uint32_t nw = blockDim.x >> 6;
        for (uint32_t q = 0; q < nw; q++) {
            const uint32_t t = R.n[2 + q];
            if (q < wave) base += t;
            all += t;
        }
        uint32_t run = base + inc - mine;
        for (uint32_t q = 0; q < per; q++) {
            if (pk[q]) {
                const uint32_t o = run, c = pk[q], tag = (s0 + q) << 7;
                for (uint32_t e = 0; e < c; e++) R.rec_of[o + e] = (uint16_t)(tag | e);
            }
            run += pk[q];
        }
        if (threadIdx.x == 0) R.n[1] = all;
        __syncthreads();
    }
    // phase B2: every k-mer of the distinct records, one per thread (the
    // rabin-karp128 variant: one segment of RK_SEG k-mers per thread)
    const uint32_t total = R.n[1];
    for (uint32_t j0 = 0; j0 < total; j0 += blockDim.x) {
        const uint32_t j = j0 + threadIdx.x;
        if (MODE == 2) {
            if (j < total) {
                const uint32_t v = R.rec_of[j];
                const uint32_t si = v >> 7, i0 = (v & 127u) * RK_SEG;
                const uint32_t* rw = R.key + si * SW;
                const uint32_t cnt = R.cnt[si], meta = rw[0], last = meta_len(meta) - k;
                const uint32_t i1 = i0 + RK_SEG < last + 1 ? i0 + RK_SEG : last + 1;
                rk_segment(A, rw + 1, i0, i1, [&](uint64_t lo, uint64_t hi, bool fwd, uint32_t x) {
                    insert_key(lo, hi, occurrence_bits(meta, fwd, x, last) - 1u + cnt, 0u);
                });
            }
            continue;
        }
        bool cl = false;
        uint32_t at = 0;
        if (j < total) {
            const uint32_t v = R.rec_of[j];
            const uint32_t si = v >> 7, i = v & 127u;
            const uint32_t* rw = R.key + si * SW;
            const uint32_t cnt = R.cnt[si], color = W ? R.col[si] : 0u;
            const uint32_t meta = rw[0];
            const uint32_t len = meta_len(meta);
            bool fwd;
            const uint64_t key = kmer_at(rw + 1, i, k, &fwd);
            const uint32_t bits = occurrence_bits(meta, fwd, i, len - k) - 1u + cnt;
            if (!part_mask || ((tab_hash64(key) >> 16) & part_mask) == w.part) {
                const uint32_t res = tab_insert64(klo, vals, mask, key, bits, probe_limit, at);
                if (res == 2) {
                    atomicOr(overflow, 1u);
                } else {
                    cl = res == 1;
                    if (W) atomicOr(&cbits[at * W + (color >> 5)], 1u << (color & 31));
                }
            }
        }
        // one LDS atomic per wavefront for the slots it claimed
        const uint64_t bal = __ballot(cl);
        if (bal) {
            const uint32_t lane = threadIdx.x & 63, first = (uint32_t)__builtin_ctzll(bal);
            uint32_t base = 0;
            if (lane == first) base = atomicAdd(n_claimed, (uint32_t)__popcll(bal));
            base = __shfl(base, (int)first);
            if (cl) claimed[base + (uint32_t)__popcll(bal & ((1ULL << lane) - 1))] = (uint16_t)at;
        }
    }
}

// Work entry with the metadata a pass needs.
struct Fetched {
    uint32_t e;
    PartWork w;
    WorkItem it;  // with its slot range [it.r0, it.r1)
};

// Thread 0's pipeline of work entries: the next entry's metadata is loaded
// in two dependent steps (part entry, work item with its slot range) spread over
// the current entry's phases, and the entry after it is already claimed, so
// no barrier waits on a chain of global loads.  Entries are claimed from the
// global cursor `claim` at a time: one atomic on that single address per
// batch instead of per entry (the address serialises its atomics).
struct Pipe {
    Fetched n1;  // next entry
    uint32_t e2;  // the entry after it (claimed)
    uint32_t bat, bat_end;  // claimed batch [bat, bat_end)
};
GG_DEV uint32_t pipe_claim(const MergeArgs& A, Pipe& p) {
    if (p.bat >= p.bat_end) {
        p.bat = atomicAdd(A.work, A.claim);
        p.bat_end = p.bat + A.claim;
    }
    return p.bat < A.n_work ? p.bat++ : A.n_work;
}
GG_DEV void pipe_part(const MergeArgs& A, Pipe& p) {
    if (p.n1.e < A.n_work) p.n1.w = A.parts ? A.parts[p.n1.e] : PartWork{p.n1.e, 0u, 0u, 0u};
}
GG_DEV void pipe_item(const MergeArgs& A, Pipe& p) {
    if (p.n1.e < A.n_work) p.n1.it = A.items[p.n1.w.item];
}

// Persistent work-stealing merge: each workgroup takes work entries from a
// global cursor, thread 0 keeping the next entry's metadata in flight.
template <int MODE>
GG_DEV void merge_lds_body(const MergeArgs& A) {
    extern __shared__ __align__(16) uint8_t lds[];
    __shared__ uint32_t s_overflow;
    __shared__ Fetched s_cur;
    EmitState es{~0u, 0u};
    const uint32_t cap_max = A.lds_cap;  // the host allocates this many entries plus a margin
    uint64_t* klo = reinterpret_cast<uint64_t*>(lds);
    uint64_t* khi = klo + cap_max;  // only used for 128-bit keys
    const uint32_t KW = Mode<MODE>::key_words(A);
    uint32_t* vals = reinterpret_cast<uint32_t*>(lds + (size_t)cap_max * 8 * KW);
    uint32_t* st = vals + cap_max;  // only used for 128-bit keys
    uint32_t* cbits = vals + (size_t)cap_max * KW;  // color_words per entry (colored builds)
    const uint32_t W = Mode<MODE>::color_words(A);
    uint16_t* claimed = reinterpret_cast<uint16_t*>(cbits + (size_t)cap_max * W);  // slots claimed by the entry
    __shared__ uint32_t s_claimed;
    const bool leader = threadIdx.x == 0;
    // the table starts empty; afterwards only claimed slots are reset
    for (uint32_t i = threadIdx.x; i < cap_max; i += blockDim.x) {
        klo[i] = EMPTY64;
        vals[i] = 0;
        if (KW == 2) st[i] = 0;
    }
    for (uint32_t i = threadIdx.x; i < cap_max * W; i += blockDim.x) cbits[i] = 0;
    if (Mode<MODE>::dedup(A)) {
        // record table slots start free; each chunk's phase B frees them again
        const RecTable R = rec_table(A, lds);
        for (uint32_t i = threadIdx.x; i < A.rec_cap; i += blockDim.x) R.state[i] = 0;
    }
    // each thread's record of the current chunk (dedup path)
    uint32_t chunk = A.rec_cap / 2;
    const uint32_t SW = Mode<MODE>::slot_words(A);
    constexpr uint32_t MSW = Mode<MODE>::MAX_SW;
    auto load_rec = [&](uint64_t r, uint64_t r_end, uint32_t* wv, uint32_t& col) {
        if (threadIdx.x < chunk && r < r_end) {
            const uint32_t* sp = A.slots + r * SW;
            if (SW == 4) {
                const uint4 v = *reinterpret_cast<const uint4*>(sp);
                wv[0] = v.x;
                wv[1] = v.y;
                wv[2] = v.z;
                wv[3] = v.w;
            } else {
                for (uint32_t q = 0; q < SW; q++) wv[q] = sp[q];
            }
            col = W ? A.slot_color[r] : 0u;
        }
    };
    uint32_t cw[MSW] = {}, ccol = 0;
    Pipe pp{};
    if (leader) {
        s_claimed = 0;
        pp.bat = pp.bat_end = 0;
        pp.n1.e = pipe_claim(A, pp);
        pipe_part(A, pp);
        pipe_item(A, pp);
        pp.e2 = pp.n1.e < A.n_work ? pipe_claim(A, pp) : A.n_work;
    }
    while (true) {
        if (leader) {
            s_cur = pp.n1;
            s_overflow = 0;
            pp.n1.e = pp.e2;
            pp.e2 = (pp.e2 < A.n_work) ? pipe_claim(A, pp) : A.n_work;
            pipe_part(A, pp);
        }
        __syncthreads();
        const Fetched cur = s_cur;
        __syncthreads();
        if (cur.e >= A.n_work) break;
        const PartWork w = cur.w;
        const WorkItem it = cur.it;
        if (!A.parts && it.occ > A.lds_try) {
            if (leader) {
                A.redo[atomicAdd(A.redo_count, 1u)] = PartWork{w.item, w.part, w.log_parts, it.occ};
                pipe_item(A, pp);
            }
            continue;
        }
        uint32_t cap = 256;
        const uint64_t want = ((uint64_t)it.occ * A.table_pct + 99) / 100;
        while (cap < want && cap < cap_max) cap <<= 1;
        if (w.log_parts) cap = cap_max;
        const uint32_t mask = cap - 1;
        if (Mode<MODE>::dedup(A)) {
            const bool hashed = !A.rec_direct || it.dup;
            // records in chunks of rec_cap / 2 <= 256 (the record table's load
            // factor stays at most one half), one record per thread; identical
            // records of one chunk are counted once, the k-mer table
            // accumulates over the chunks.  Each thread's record of the next
            // chunk is loaded while the current chunk is processed.  Records
            // placed by position without per-k-mer tags (each distinct record
            // rolled by one thread) fill the whole record table: one record
            // per thread of the workgroup.
            chunk = (hashed || Mode<MODE>::tag_kmers(A)) ? A.rec_cap / 2
                                                         : (A.rec_cap < blockDim.x ? A.rec_cap : blockDim.x);
            load_rec(it.r0 + threadIdx.x, it.r1, cw, ccol);
            for (uint64_t c0 = it.r0; c0 < it.r1; c0 += chunk) {
                const uint64_t c1 = c0 + chunk < it.r1 ? c0 + chunk : it.r1;
                uint32_t nw[MSW] = {}, ncol = 0;
                load_rec(c1 + threadIdx.x, it.r1, nw, ncol);
                insert_item_dedup<MODE>(A, c0, c1, w, klo, khi, st, vals, cbits, mask, 96u, &s_overflow, claimed,
                                  &s_claimed, rec_table(A, lds), cw, ccol, hashed);
#pragma unroll
                for (uint32_t q = 0; q < MSW; q++) cw[q] = nw[q];
                ccol = ncol;
                __syncthreads();  // the next chunk reuses the record table
            }
        } else {
            insert_item<MODE>(A, it.r0, it.r1, w, klo, khi, st, vals, cbits, mask, 96u, &s_overflow, claimed, &s_claimed);
        }
        if (leader) pipe_item(A, pp);
        __syncthreads();
        const uint32_t n_cl = s_claimed;
        if (s_overflow) {
            // more distinct k-mers than the table holds: redone as finer parts
            if (leader) A.redo[atomicAdd(A.redo_count, 1u)] = PartWork{w.item, w.part, w.log_parts, it.occ};
        } else {
            emit_solid<MODE>(A, it.sub, klo, khi, st, vals, cbits, claimed, n_cl, es);
        }
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < n_cl; j += blockDim.x) {
            const uint32_t i = claimed[j];
            klo[i] = EMPTY64;
            vals[i] = 0;
            if (KW == 2) st[i] = 0;
            for (uint32_t q = 0; q < W; q++) cbits[i * W + q] = 0;
        }
        if (leader) s_claimed = 0;
        __syncthreads();
    }
    close_chunk(A, es);
}

__global__ __launch_bounds__(256) void merge_lds_kernel(MergeArgs A) { merge_lds_body<0>(A); }

// The seq-hash k <= 32 variant is held to 64 VGPRs: eight waves per SIMD when
// the LDS tables allow that many workgroups.
__global__ __launch_bounds__(256, 8) void merge_lds_seq64_kernel(MergeArgs A) { merge_lds_body<1>(A); }

__global__ __launch_bounds__(256) void merge_lds_rk128_kernel(MergeArgs A) { merge_lds_body<2>(A); }

// One thread per sub-bucket: runs of its units become work items of at most
// `group_occ` occurrences (a single unit above that is an item of its own),
// each with its slot range.  An item is marked dup when one of its units
// stores more than `dd_max` records: the compaction deduplicates units of up
// to dd_max records, so only such units can still hold identical records.
__global__ void make_items_kernel(const uint32_t* unit_kmers, const uint64_t* unit_off, uint32_t n_sub, uint32_t xb,
                                  uint32_t group_occ, uint32_t dd_max, WorkItem* items, unsigned int* n_items) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_sub) return;
    const uint32_t u_begin = s << xb, u_end = (s + 1) << xb;
    uint32_t start = u_begin, acc = 0, dup = 0;
    for (uint32_t u = u_begin; u < u_end; u++) {
        const uint32_t o = unit_kmers[u];
        if (acc == 0 && o == 0) {
            start = u + 1;
            continue;
        }
        if (acc > 0 && acc + o > group_occ) {
            items[atomicAdd(n_items, 1u)] = WorkItem{start, dup, acc, s, unit_off[start], unit_off[u]};
            start = u;
            acc = 0;
            dup = 0;
        }
        acc += o;
        if (unit_off[u + 1] - unit_off[u] > dd_max) dup = 1;
    }
    if (acc > 0) items[atomicAdd(n_items, 1u)] = WorkItem{start, dup, acc, s, unit_off[start], unit_off[u_end]};
}

}  // namespace gg
