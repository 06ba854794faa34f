// resplit.hip -- outlier sub-buckets split again with a longer minimizer
// (crates/kmers_transform/src/resplitter.rs:47-183).
//
// Every stored record of an outlier sub-bucket goes through process_sequence
// (crates/assembler_minimizer_bucketing/src/lib.rs:152-270) once more, in its
// stored orientation: m' = k - 10 (min(m, 2) for k <= 11,
// crates/assembler_kmers_merge/src/lib.rs:236-246), windows of k - m'
// m'-mers, 2^L normal buckets plus a duplicates bucket (index 2^L), no second
// bucket, and the record's READ_FLAG_INCL_BEGIN / READ_FLAG_INCL_END as the
// include flags of its ends (reprocess_sequence): an end that is not included
// is left out of the window minima.  Each piece keeps the record's
// multiplicity and color.  The resplit buckets of outlier i become extra
// sub-buckets base_i .. base_i + 2^L_i (duplicates bucket last), one physical
// unit each, appended after the bucketing's units.
//
// One thread walks one record (records are at most a few hundred bases and
// outliers are rare): rolling ntHash over the m'-mers, the window minimum of
// the last k - m' <= 10 m'-mers folded with mm_combine, a split where the
// minimum changes.  A count pass sizes the extra sub-buckets; a write pass
// stores the pieces into the rebuilt arena.
#pragma once
#include "common.h"

namespace gg {

constexpr uint32_t RS_RING = 16;  // window m'-mers held per thread (k - m' <= 10)

struct ResplitArgs {
    const uint32_t* slots;  // the bucketing's arena
    const uint32_t* slot_color;  // nullptr when uncolored
    const uint32_t* pool;
    uint64_t pool_words;
    uint32_t SW, k, m2, w2;
    uint32_t n_out;
    const uint64_t* rec_prefix;  // [n_out + 1]: records of the outliers before outlier i
    const uint64_t* rec_first;  // [n_out]: arena slot of outlier i's first record
    const uint32_t* log_buckets;  // [n_out]
    const uint32_t* ext_base;  // [n_out]: first extra sub-bucket of outlier i, relative to the first extra
    // count pass
    uint32_t* ext_records;  // [n_extra]
    uint32_t* ext_kmers;  // [n_extra]
    unsigned long long* pool_need;  // pool words of the pieces too long for a slot
    unsigned long long* sums;  // [4]: copies and k-mer occurrences of the records, then of the pieces
    // write pass
    uint32_t* ext_cursor;  // [n_extra]
    const uint64_t* unit_off;  // rebuilt unit offsets
    uint32_t first_extra_unit, xb;
    uint32_t* out_slots;
    uint32_t* out_color;
    uint32_t* out_pool;
    unsigned long long* pool_cursor;  // starts at the pool words already used
    uint64_t pool_cap;
    uint32_t* error_flag;
};

GG_DEV uint64_t rs_rotl(uint64_t x, uint32_t r) {
    r &= 63;
    return r ? (x << r) | (x >> (64 - r)) : x;
}
GG_DEV uint64_t rs_rotr(uint64_t x, uint32_t r) {
    r &= 63;
    return r ? (x >> r) | (x << (64 - r)) : x;
}
// ntHash seeds of a 2-bit code (A0 C1 T2 G3 is ASCII (c >> 1) & 3, and
// (ASCII & 6) == 2 * code): forward h = (2c + 1) * MULT, complement
// rc = ((2c ^ 4) + 1) * MULT (crates/hashes/src/nthash_base.rs:49-61).
GG_DEV uint64_t rs_h(uint32_t c) { return (uint64_t)(2 * c + 1) * NT_MULT; }
GG_DEV uint64_t rs_rc(uint32_t c) { return (uint64_t)(((2 * c) ^ 4) + 1) * NT_MULT; }

// Piece [start, end) of a record: its bucket and slot fields.
struct RsPiece {
    uint32_t start, end, bucket, minpos, flags;
    bool rc;
};

// Walks one record and calls emit(piece) for every piece, in order.
template <class EMIT>
GG_DEV void resplit_walk(const ResplitArgs& A, const uint32_t* bw, uint32_t len, uint32_t fl, uint32_t log_b,
                         EMIT&& emit) {
    const uint32_t k = A.k, m = A.m2, w = A.w2;
    if (len < k || len < m) return;
    const uint32_t n = len - m + 1;
    const bool inc_first = (fl & 1) != 0, inc_last = (fl & 2) != 0;
    const uint32_t sb = inc_first ? 0u : 1u, se = inc_last ? 0u : 1u;
    if (n < sb + se + w) return;  // no complete window once the ends are skipped
    auto code = [&](uint32_t i) -> uint32_t { return (bw[i >> 4] >> (2 * (i & 15))) & 3; };
    uint64_t rv[RS_RING];
    uint32_t rp[RS_RING];
    bool rf[RS_RING];
    const uint32_t C = 1u << log_b;
    auto piece = [&](uint32_t start, uint32_t end, uint64_t v, uint32_t pos, bool fwd, bool first, bool last) {
        RsPiece P;
        const bool uniq = (v & 1) != 0;
        P.rc = uniq && !fwd;
        P.bucket = uniq ? (uint32_t)((v >> 1) & (C - 1)) : C;
        P.minpos = uniq ? (P.rc ? end - pos - m : pos - start) : 0u;
        const uint32_t f0 = (first && inc_first) ? 1u : 0u, f1 = (last && inc_last) ? 1u : 0u;
        P.flags = (f0 << (P.rc ? 1 : 0)) | (f1 << (P.rc ? 0 : 1));
        P.start = start;
        P.end = end;
        emit(P);
    };
    // rolling ntHash (crates/hashes/src/cn_nthash.rs:19-58)
    uint64_t fh = 0, bk = 0;
    for (uint32_t i = 0; i + 1 < m; i++) {
        fh ^= rs_rotl(rs_h(code(i)), m - i - 2);
        bk ^= rs_rotl(rs_rc(code(i)), i);
    }
    bool have = false, first = true;
    uint64_t last_v = 0;
    uint32_t last_p = 0, start = 0;
    bool last_f = false;
    for (uint32_t q = 0; q < n; q++) {
        const uint32_t ci = code(q), ck = code(q + m - 1);
        const uint64_t f = rs_rotl(fh, 1) ^ rs_h(ck);
        fh = f ^ rs_rotl(rs_h(ci), m - 1);
        const uint64_t r = bk ^ rs_rotl(rs_rc(ck), m - 1);
        bk = rs_rotr(r ^ rs_rc(ci), 1);
        if (q < sb || q >= n - se) continue;
        const uint32_t slot = (q - sb) % RS_RING;
        rv[slot] = ((f < r ? f : r) << 1) | (f != r ? 1ULL : 0ULL);
        rp[slot] = q;
        rf[slot] = f < r;
        if (q < sb + w - 1) continue;
        // window of the m'-mers p..q, folded in order
        const uint32_t p = q + 1 - w;
        uint32_t s0 = (p - sb) % RS_RING;
        uint64_t v = rv[s0];
        uint32_t pos = rp[s0];
        bool fwd = rf[s0];
        for (uint32_t j = 1; j < w; j++) {
            const uint32_t sj = (p - sb + j) % RS_RING;
            uint64_t nv;
            uint32_t np;
            mm_combine(v, pos, rv[sj], rp[sj], nv, np);
            if (np != pos) fwd = rf[sj];
            v = nv;
            pos = np;
        }
        if (have && (v != last_v || ((v & 1) && pos != last_p))) {
            // the group before window p ends with the k-mer at p - 1
            piece(start, p + k - 1, last_v, last_p, last_f, first, false);
            start = p - 1;
            first = false;
        }
        have = true;
        last_v = v;
        last_p = pos;
        last_f = fwd;
    }
    if (have) piece(start, len, last_v, last_p, last_f, first, true);
}

// Record r of the flattened outlier records: its outlier, slot, bases.
GG_DEV bool resplit_record(const ResplitArgs& A, uint64_t t, uint32_t& oi, uint64_t& slot_i, const uint32_t*& bw,
                           uint32_t& len, uint32_t& meta) {
    uint32_t lo = 0, hi = A.n_out;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (A.rec_prefix[mid] <= t) lo = mid; else hi = mid;
    }
    oi = lo;
    slot_i = A.rec_first[lo] + (t - A.rec_prefix[lo]);
    const uint32_t* sl = A.slots + slot_i * A.SW;
    meta = sl[0];
    if (meta & META_LONG) {
        len = sl[1];
        const uint64_t off = (uint64_t)sl[2] | ((uint64_t)sl[3] << 32);
        if (off + ((uint64_t)len + 15) / 16 > A.pool_words) {
            atomicOr(A.error_flag, 2u);
            return false;
        }
        bw = A.pool + off;
    } else {
        len = meta_len(meta);
        if (len > (A.SW - 1) * 16) {
            atomicOr(A.error_flag, 2u);
            return false;
        }
        bw = sl + 1;
    }
    return true;
}

__global__ void resplit_count_kernel(ResplitArgs A) {
    const uint64_t total = A.rec_prefix[A.n_out];
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t oi, len, meta;
        uint64_t si;
        const uint32_t* bw;
        if (!resplit_record(A, t, oi, si, bw, len, meta)) continue;
        const uint32_t base = A.ext_base[oi];
        const unsigned long long mult = meta_mult(meta);
        unsigned long long pieces = 0, piece_kmers = 0;
        resplit_walk(A, bw, len, meta_flags(meta), A.log_buckets[oi], [&](const RsPiece& P) {
            const uint32_t L = P.end - P.start;
            atomicAdd(&A.ext_records[base + P.bucket], 1u);
            atomicAdd(&A.ext_kmers[base + P.bucket], L - A.k + 1);
            if (L > (A.SW - 1) * 16) atomicAdd(A.pool_need, (unsigned long long)((L + 15) / 16));
            pieces += mult;
            piece_kmers += (L - A.k + 1) * mult;
        });
        atomicAdd(&A.sums[0], mult);
        atomicAdd(&A.sums[1], pieces);
        atomicAdd(&A.sums[2], len >= A.k ? (len - A.k + 1) * mult : 0ull);
        atomicAdd(&A.sums[3], piece_kmers);
    }
}

__global__ void resplit_write_kernel(ResplitArgs A) {
    const uint64_t total = A.rec_prefix[A.n_out];
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t oi, len, meta;
        uint64_t si;
        const uint32_t* bw;
        if (!resplit_record(A, t, oi, si, bw, len, meta)) continue;
        const uint32_t base = A.ext_base[oi];
        const uint32_t mult_bits = meta & (0x3FFu << META_MULT_SHIFT);
        const uint32_t color = A.slot_color ? A.slot_color[si] : 0u;
        resplit_walk(A, bw, len, meta_flags(meta), A.log_buckets[oi], [&](const RsPiece& P) {
            const uint32_t e = base + P.bucket;
            const uint32_t L = P.end - P.start;
            const uint32_t rank = atomicAdd(&A.ext_cursor[e], 1u);
            const uint64_t unit = (uint64_t)(A.first_extra_unit + e) << A.xb;
            const uint64_t at = A.unit_off[unit] + rank;
            if (at >= A.unit_off[unit + 1]) {
                atomicOr(A.error_flag, 4u);
                return;
            }
            uint32_t* out = A.out_slots + at * A.SW;
            const uint32_t hdr = (P.flags << 8) | ((P.minpos & 0xFF) << 10) | (P.rc ? META_RC : 0u) | mult_bits;
            uint32_t* dst;
            if (L > (A.SW - 1) * 16) {
                const uint64_t words = (L + 15) / 16;
                const unsigned long long off = atomicAdd(A.pool_cursor, (unsigned long long)words);
                out[0] = hdr | META_LONG;
                out[1] = L;
                out[2] = (uint32_t)off;
                out[3] = (uint32_t)(off >> 32);
                for (uint32_t q = 4; q < A.SW; q++) out[q] = 0;
                if (off + words > A.pool_cap) {
                    atomicOr(A.error_flag, 4u);
                    return;
                }
                dst = A.out_pool + off;
            } else {
                out[0] = hdr | L;
                for (uint32_t q = 1; q < A.SW; q++) out[q] = 0;
                dst = out + 1;
            }
            uint32_t acc = 0;
            for (uint32_t i = 0; i < L; i++) {
                const uint32_t c = P.rc ? (((bw[(P.end - 1 - i) >> 4] >> (2 * ((P.end - 1 - i) & 15))) & 3) ^ 2u)
                                        : ((bw[(P.start + i) >> 4] >> (2 * ((P.start + i) & 15))) & 3);
                acc |= c << (2 * (i & 15));
                if ((i & 15) == 15 || i + 1 == L) {
                    dst[i >> 4] = acc;
                    acc = 0;
                }
            }
            if (A.out_color) A.out_color[at] = color;
        });
    }
}

// Per-unit arrays of the rebuilt arena: the outliers' units emptied, each
// extra sub-bucket's pieces in its first unit (k-mer occurrences and
// stored k-mer positions alike: pieces are not deduplicated).
__global__ void resplit_units_kernel(const uint32_t* dcount, const uint32_t* count, const uint32_t* kmers,
                                     const uint32_t* skmers, const uint8_t* is_outlier, uint32_t n_units,
                                     uint32_t n_units_all, uint32_t xb, uint32_t first_extra,
                                     const uint32_t* ext_records, const uint32_t* ext_kmers, uint32_t* dcount2,
                                     uint32_t* count2, uint32_t* kmers2, uint32_t* skmers2) {
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n_units_all; u += gridDim.x * blockDim.x) {
        uint32_t d = 0, c = 0, km = 0, skm = 0;
        if (u < n_units) {
            if (!is_outlier[u >> xb]) {
                d = dcount[u];
                c = count[u];
                km = kmers[u];
                skm = skmers[u];
            }
        } else if ((u & ((1u << xb) - 1)) == 0) {
            const uint32_t e = (u >> xb) - first_extra;
            d = c = ext_records[e];
            km = skm = ext_kmers[e];
        }
        dcount2[u] = d;
        count2[u] = c;
        kmers2[u] = km;
        skmers2[u] = skm;
    }
}

// The kept units' stored records to their place in the rebuilt arena, one
// wavefront per unit.
__global__ __launch_bounds__(256) void resplit_copy_kernel(const uint32_t* slots, const uint32_t* slot_color,
                                                          const uint64_t* off, const uint64_t* off2,
                                                          const uint8_t* is_outlier, uint32_t n_units, uint32_t xb,
                                                          uint32_t SW, uint32_t* slots2, uint32_t* color2) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    for (uint32_t u = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); u < n_units; u += waves) {
        if (is_outlier[u >> xb]) continue;
        const uint64_t a = off[u], n = off[u + 1] - a, b = off2[u];
        const uint32_t* src = slots + a * SW;
        uint32_t* dst = slots2 + b * SW;
        for (uint64_t i = lane; i < n * SW; i += 64) dst[i] = src[i];
        if (slot_color)
            for (uint64_t i = lane; i < n; i += 64) color2[b + i] = slot_color[a + i];
    }
}

}  // namespace gg
