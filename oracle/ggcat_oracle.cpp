// ggcat_oracle.cpp -- CPU restatement of GGCAT's minimizer bucketing and
// per-sub-bucket k-mer merge, used ONLY as a test checker and as the
// measured CPU baseline ("port").  Nothing in the product library links,
// loads or calls this file; the HIP path must stand on its own.
//
// Each function cites the reference file:line whose behaviour it restates
// (paths relative to the algbio/ggcat repository root).
//
// Parity status of each part (how it is pinned; see also tests/):
//   ntHash m-mer hash (nthash_m)          pinned: closed form + RC symmetry + roll==direct
//   BatchMinQueue restatement (minqueue)  pinned: brute-force closed-form window minimum
//   process_sequence / splits (bucket_run) pinned: super-k-mer cover property + brute force
//   outlier resplit (resplit_records)     pinned: independent pure-Python restatement (brute-force
//                                         windows with the record's include flags) + the solid
//                                         k-mers and multiplicities of an outlier equal before and
//                                         after its resplit
//   seq-hash u64/u128, rabin-karp128      pinned: invertibility / roll==direct / RC symmetry
//   MapEntry merge (merge_subbucket)      pinned: union of solid sets == exact k-mer counts
//   compaction (compact)                  pinned: multiset expansion equality
//   color sets (merge_subbucket, colored) pinned: brute-force colored k-mer counts
//   query counts (orc_query)              pinned: pure-Python brute force over the graph k-mer set
//   unitig chains, order-free (unitig_chains) pinned: definition check (maximal paths of mutually unique
//                                         unflagged neighbours) + every solid k-mer covered once
//   reference extension loop (unitig_sequential) pinned against unitig_chains on inputs where the
//                                         reference's visiting order cannot matter
//   extension flags (orc_extension)       pinned: equal to neighbour counts of the solid set (same test)
//   sorting extender (normal buckets,     pinned: independent pure-Python restatement (explicit
//     merge_subbucket_sorting)            supertigs) on small inputs; equal (k-mer, multiplicity) set
//                                         to the hash-map rule on every test shape; exact while no
//                                         centred k-mer holds both extra bases (counted, tests assert 0)
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

typedef unsigned __int128 u128;

namespace {

// ---------------------------------------------------------------- ntHash ---
// crates/hashes/src/nthash_base.rs:49-61  h(c)/rc(c) for ASCII input:
// mult = (c & 6) + 1 (forward), ((c & 6) ^ 4) + 1 (complement).
constexpr uint64_t NT_MULT = 0x397f178c6ae330f9ULL;
inline uint64_t rotl64(uint64_t x, unsigned r) {
    r &= 63;
    return r ? (x << r) | (x >> (64 - r)) : x;
}
inline uint64_t rotr64(uint64_t x, unsigned r) {
    r &= 63;
    return r ? (x >> r) | (x << (64 - r)) : x;
}
inline uint64_t nt_h(uint8_t c) { return (uint64_t)((c & 6) + 1) * NT_MULT; }
inline uint64_t nt_rc(uint8_t c) { return (uint64_t)(((c & 6) ^ 4) + 1) * NT_MULT; }

// crates/hashes/src/cn_nthash.rs:19-58 CanonicalNtHashIterator::{new,roll_hash}
// Produces, for every m-mer of `seq`, the (fw, rc) pair exactly as the
// rolling iterator yields it.
void nthash_m(const uint8_t* seq, size_t len, unsigned m, uint64_t* fw_out, uint64_t* rc_out) {
    if (len < m) return;
    uint64_t fh = 0, bw = 0;
    for (unsigned i = 0; i + 1 < m; i++) {
        fh ^= rotl64(nt_h(seq[i]), m - i - 2);
        bw ^= rotl64(nt_rc(seq[i]), i);
    }
    const unsigned km1 = m - 1;
    for (size_t i = 0; i + km1 < len; i++) {
        uint8_t bi = seq[i], bk = seq[i + km1];
        uint64_t res = rotl64(fh, 1) ^ nt_h(bk);
        fh = res ^ rotl64(nt_h(bi), km1);
        uint64_t res_rc = bw ^ rotl64(nt_rc(bk), km1);
        bw = rotr64(res_rc ^ nt_rc(bi), 1);
        fw_out[i] = res;
        rc_out[i] = res_rc;
    }
}

// ------------------------------------------------------- BatchMinQueue ---
// MinimizerExtraData (crates/assembler_minimizer_bucketing/src/lib.rs:31-35)
struct MinX {
    uint32_t index = 0;
    bool is_forward = false;
    bool operator!=(const MinX& o) const { return index != o.index || is_forward != o.is_forward; }
};
struct Item {
    uint64_t v = 0;
    MinX x;
};
// std::cmp::min_by_key: returns the first argument unless the second is
// strictly smaller.
inline Item min_by_key(const Item& a, const Item& b) { return (b.v < a.v) ? b : a; }

// crates/hashes/src/rolling/batch_minqueue.rs:8-189 (ENABLE_DUPLICATE_CHECKING = true)
struct BatchMinQueue {
    size_t size;
    std::vector<Item> backward;
    struct Split {
        Item value;
        size_t position;
    };
    std::vector<Split> splits;
    explicit BatchMinQueue(size_t w) : size(w), backward(w - 1), splits(w + 2) {}

    // batch_minqueue.rs:37-124
    template <class CB, class FL>
    void get_minimizers(const Item* it, size_t it_len, size_t skip_ending_count, CB&& cb, FL&& flush) {
        size_t pos = 0;
        size_t sz = it_len - skip_ending_count;
        if (sz < backward.size()) return;
        for (auto& b : backward) b = it[pos++];
        size_t offset = 0;
        sz -= backward.size();
        const size_t bl = backward.size();
        while (offset < sz) {
            size_t remaining = std::min(sz - offset, bl);
            for (size_t ii = bl - 1; ii-- > 0;) {  // (0..bl-1).rev()
                Item cur = backward[ii];
                Item nxt = backward[ii + 1];
                if (cur.v == nxt.v) cur.v &= ~1ULL;
                backward[ii] = min_by_key(cur, nxt);
            }
            Item new_item = it[pos++];
            Item first_min = min_by_key(new_item, backward[0]);
            if (new_item.v == backward[0].v) first_min.v &= ~1ULL;
            cb(first_min, offset);
            Item last_forward = new_item;
            backward[0] = new_item;
            for (size_t i = 1; i < remaining; i++) {
                Item ni = it[pos++];
                Item& cb_ = backward[i];
                if (ni.v == last_forward.v) last_forward.v &= ~1ULL;
                last_forward = min_by_key(last_forward, ni);
                Item cur_min = min_by_key(last_forward, cb_);
                if (last_forward.v == cb_.v) cur_min.v &= ~1ULL;
                cb(cur_min, offset + i);
                cb_ = ni;
            }
            offset += remaining;
            flush(offset == sz);
        }
    }

    // batch_minqueue.rs:127-189
    template <class SCB>
    void get_minimizer_splits(const Item* it, size_t it_len, size_t skip_beginning_count,
                              size_t skip_ending_count, SCB&& splits_cb) {
        Item last_value;
        bool is_first = true;
        size_t last_index = it_len - size;
        size_t ptr = 0;
        get_minimizers(
            it + skip_beginning_count, it_len - skip_beginning_count, skip_ending_count,
            [&](const Item& mm, size_t index) {
                bool has_diff = !is_first && (last_value.v != mm.v ||
                                              (last_value.x != mm.x && (mm.v & 1) == 1));
                splits[ptr].value = last_value;
                splits[ptr].position = index + skip_beginning_count;
                ptr += has_diff ? 1 : 0;
                is_first = false;
                last_value = mm;
            },
            [&](bool is_last) {
                if (is_last) {
                    splits[ptr].value = last_value;
                    splits[ptr].position = last_index;
                    ptr++;
                }
                for (size_t c = 0; c < ptr; c++) {
                    bool last = (c + 1 == ptr);
                    splits_cb(splits[c].position, splits[c].value, last && is_last);
                }
                ptr = 0;
            });
    }
};

// ------------------------------------------------------------ records ---
struct Params {
    uint32_t k, m, b, sb;
    int canonical;
};

struct RecordSet {
    // One entry per super-k-mer record; bases are ASCII (uppercase ACGT),
    // already reverse-complemented when rc was chosen.
    std::vector<uint32_t> sub;  // bucket * 2^sb + second_bucket
    std::vector<uint32_t> len;
    std::vector<uint8_t> flags;
    std::vector<uint16_t> minpos;
    std::vector<uint32_t> mult;
    std::vector<uint32_t> color;
    std::vector<uint64_t> off;  // into bases
    std::vector<uint8_t> bases;
    void clear() {
        sub.clear(); len.clear(); flags.clear(); minpos.clear(); mult.clear(); color.clear();
        off.clear(); bases.clear();
    }
    void append(const RecordSet& o) {
        uint64_t base = bases.size();
        sub.insert(sub.end(), o.sub.begin(), o.sub.end());
        len.insert(len.end(), o.len.begin(), o.len.end());
        flags.insert(flags.end(), o.flags.begin(), o.flags.end());
        minpos.insert(minpos.end(), o.minpos.begin(), o.minpos.end());
        mult.insert(mult.end(), o.mult.begin(), o.mult.end());
        color.insert(color.end(), o.color.begin(), o.color.end());
        for (auto x : o.off) off.push_back(base + x);
        bases.insert(bases.end(), o.bases.begin(), o.bases.end());
    }
};

inline uint8_t comp_ascii(uint8_t c) {
    switch (c) {
        case 'A': return 'T';
        case 'C': return 'G';
        case 'G': return 'C';
        case 'T': return 'A';
    }
    return 'N';
}

// crates/io/src/sequences_reader.rs:26-37 SEQ_LETTERS_MAPPING
inline uint8_t normalize(uint8_t c) {
    switch (c) {
        case 'A': case 'a': return 'A';
        case 'C': case 'c': return 'C';
        case 'G': case 'g': return 'G';
        case 'T': case 't': return 'T';
    }
    return 'N';
}

// crates/assembler_minimizer_bucketing/src/lib.rs:171-270 process_sequence
// with SEPARATE_DUPLICATES over one N-free sequence.  include_first /
// include_last come from preprocess_dna_sequence (:149-150, both true) for
// an input run, or from reprocess_sequence (:152-168: the record's
// READ_FLAG_INCL_BEGIN / READ_FLAG_INCL_END) when a record is split again;
// an end that is not included is skipped by the minimizer queue (:187-188).
// Every record emitted carries `mult`.
void split_sequence(const uint8_t* seq, size_t L, const Params& P, uint32_t color, uint32_t mult,
                    bool include_first0, bool include_last0, RecordSet& out, std::vector<uint64_t>& fw,
                    std::vector<uint64_t>& rcv, std::vector<Item>& items, BatchMinQueue& q) {
    const uint32_t k = P.k, m = P.m;
    if (L < m) return;
    size_t n = L - m + 1;
    fw.resize(n);
    rcv.resize(n);
    items.resize(n);
    nthash_m(seq, L, m, fw.data(), rcv.data());
    for (size_t i = 0; i < n; i++) {
        // to_unextendable() | !is_rc_symmetric()   (cn_nthash.rs:93-109, lib.rs:200-209)
        uint64_t mn = std::min(fw[i], rcv[i]);
        items[i].v = (mn << 1) | (fw[i] != rcv[i] ? 1ULL : 0ULL);
        items[i].x.index = (uint32_t)i;
        items[i].x.is_forward = fw[i] < rcv[i];
    }
    size_t last_index = 1;
    bool include_first = include_first0;
    const bool include_last_pre = include_last0;
    const uint32_t dup_bucket = 1u << P.b;
    const size_t skip_before = include_first0 ? 0 : 1, skip_after = include_last0 ? 0 : 1;
    q.get_minimizer_splits(items.data(), n, skip_before, skip_after, [&](size_t index, const Item& mh, bool is_last) {
        bool include_last = include_last_pre && is_last;
        uint32_t bucket;
        bool rc;
        uint16_t minimizer_pos;
        if ((mh.v & 1) == 0) {
            bucket = dup_bucket;
            rc = false;
            minimizer_pos = 0;
        } else {
            rc = P.canonical && !mh.x.is_forward;
            // MNHFactory::get_bucket(0, b, h) = (h >> 1) % 2^b   (cn_nthash.rs:135-142)
            bucket = (uint32_t)((mh.v >> 1) & ((1ULL << P.b) - 1));
            minimizer_pos = rc ? (uint16_t)(index + k - 1 - mh.x.index - m)
                               : (uint16_t)(mh.x.index - (last_index - 1));
        }
        uint32_t second = (uint32_t)((mh.v >> (P.b + 1)) & ((1ULL << P.sb) - 1));
        size_t s0 = last_index - 1, s1 = index + k - 1;
        uint8_t flags = (uint8_t)(((include_first ? 1 : 0) << (rc ? 1 : 0)) |
                                  ((include_last ? 1 : 0) << (rc ? 0 : 1)));
        out.sub.push_back((bucket << P.sb) | second);
        out.len.push_back((uint32_t)(s1 - s0));
        out.flags.push_back(flags);
        out.minpos.push_back(minimizer_pos);
        out.mult.push_back(mult);
        out.color.push_back(color);
        out.off.push_back(out.bases.size());
        if (rc) {
            for (size_t i = s1; i-- > s0;) out.bases.push_back(comp_ascii(seq[i]));
        } else {
            out.bases.insert(out.bases.end(), seq + s0, seq + s1);
        }
        last_index = index;
        include_first = false;
    });
}

// One N-free run of the input (include_first = include_last = true).
void bucket_run(const uint8_t* seq, size_t L, const Params& P, uint32_t color, RecordSet& out,
                std::vector<uint64_t>& fw, std::vector<uint64_t>& rcv, std::vector<Item>& items,
                BatchMinQueue& q) {
    split_sequence(seq, L, P, color, 1, true, true, out, fw, rcv, items, q);
}

// Resplit of an outlier sub-bucket (crates/kmers_transform/src/resplitter.rs:47-183):
// every record, in its stored orientation and with its multiplicity and
// color, goes through process_sequence again with the resplit minimizer
// length (crates/assembler_kmers_merge/src/lib.rs:236-246: m' = k - 10 when
// k > 11, else min(m, 2)), 2^log_buckets normal buckets and one duplicates
// bucket (index 2^log_buckets, new_with_duplicates at
// crates/assembler_kmers_merge/src/lib.rs:120-128), no second bucket, and
// include_first / include_last from the record's flags (reprocess_sequence).
// Output records carry sub = base + resplit bucket.
inline uint32_t resplit_m(uint32_t k, uint32_t m) { return k > 11 ? k - 10 : std::min<uint32_t>(m, 2); }

void resplit_records(const RecordSet& R, const std::vector<size_t>& idx, uint32_t k, uint32_t m,
                     uint32_t log_buckets, uint32_t base, RecordSet& out) {
    Params P{k, resplit_m(k, m), log_buckets, 0, 1};
    std::vector<uint64_t> fw, rcv;
    std::vector<Item> items;
    BatchMinQueue q(P.k - P.m);
    RecordSet part;
    for (size_t r : idx) {
        part.clear();
        const uint8_t fl = R.flags[r];
        split_sequence(R.bases.data() + R.off[r], R.len[r], P, R.color[r], R.mult[r], (fl & 1) != 0, (fl & 2) != 0,
                       part, fw, rcv, items, q);
        for (auto& s : part.sub) s += base;
        out.append(part);
    }
}

// crates/minimizer_bucketing/src/sequences_splitter.rs:15-40 (runs >= k of
// non-N bases) applied to normalised records, then bucket_run on each run.
void bucket_records(const uint8_t* seq, const uint64_t* rec_off, size_t r0, size_t r1,
                    const uint32_t* rec_color, const Params& P, RecordSet& out) {
    std::vector<uint64_t> fw, rcv;
    std::vector<Item> items;
    std::vector<uint8_t> norm;
    BatchMinQueue q(P.k - P.m);
    for (size_t r = r0; r < r1; r++) {
        const uint8_t* s = seq + rec_off[r];
        size_t len = rec_off[r + 1] - rec_off[r];
        norm.resize(len);
        for (size_t i = 0; i < len; i++) norm[i] = normalize(s[i]);
        size_t end = 0;
        uint32_t color = rec_color ? rec_color[r] : 0;
        while (end < len) {
            size_t start = end;
            while (start < len && ((norm[start] ^ 'N') & 7) == 0) start++;
            end = start;
            while (end < len && ((norm[end] ^ 'N') & 7) != 0) end++;
            if (end - start >= P.k) bucket_run(norm.data() + start, end - start, P, color, out, fw, rcv, items, q);
        }
    }
}

// ------------------------------------------------------- k-mer hashes ---
inline uint8_t cbase(uint8_t ascii) { return (ascii >> 1) & 3; }  // utils/src/lib.rs:44-47

// crates/hashes/src/base/cn_seqhash_base.rs:22-72 (u64 for k <= 32, u128 otherwise)
struct SeqHash {
    u128 fw, rc;
};
inline u128 seq_mask(unsigned k) { return (k >= 64) ? ~(u128)0 : (((u128)1 << (2 * k)) - 1); }

// crates/hashes/src/cn_rkhash.rs:61-78 (u128 constants), base/cn_rkhash_base.rs:10-110
inline u128 mk128(uint64_t hi, uint64_t lo) { return ((u128)hi << 64) | lo; }
const u128 RK_MULT = mk128(0x3eb9402f3e733993ULL, 0xadd64d3ca00e1b6bULL);
const u128 RK_MULT_INV = mk128(0x09cb6ff6f1b1a6d7ULL, 0x33e0952e899c3943ULL);
const u128 RK_A = mk128(0x4751137d01d863c5ULL, 0xb8c36de2b7d399dfULL);
const u128 RK_C = mk128(0x37ea3a13226503fbULL, 0x783f5cb69f4552bdULL);
const u128 RK_G = mk128(0x50796b285343f09aULL, 0x0c53113ae736572bULL);
const u128 RK_T = mk128(0x1e62d96a5e1f5adeULL, 0x2d4e68d8f88110b7ULL);
// FWD_LOOKUP / BKW_LOOKUP on compressed codes A=0,C=1,T=2,G=3
inline u128 rk_fwd(uint8_t c) {
    switch (c & 3) { case 0: return RK_A; case 1: return RK_C; case 2: return RK_T; default: return RK_G; }
}
inline u128 rk_bkw(uint8_t c) {
    switch (c & 3) { case 0: return RK_T; case 1: return RK_G; case 2: return RK_A; default: return RK_C; }
}
u128 pow128(u128 b, size_t e) {  // crates/hashes/src/lib.rs:169-183 fastexp
    u128 r = 1;
    while (e) {
        if (e & 1) r *= b;
        e >>= 1;
        b *= b;
    }
    return r;
}

// hash kind: 0 = canonical seq-hash, 1 = canonical rabin-karp128
// Iterates the k-mers of a 2-bit sequence calling cb(idx, key, is_forward).
template <class CB>
void kmer_hashes(const uint8_t* codes, size_t len, unsigned k, int kind, CB&& cb) {
    if (len < k) return;
    if (kind == 0) {
        const u128 mask = seq_mask(k);
        u128 fh = 0, bw = 0;
        for (unsigned i = 0; i + 1 < k; i++) {  // cn_seqhash_base.rs:22-51
            fh |= (u128)codes[i] << (2 * i);
            bw = (bw << 2) | (u128)(codes[i] ^ 2);
        }
        fh <<= 2;
        bw &= mask;
        for (size_t idx = k - 1; idx < len; idx++) {  // roll_hash :54-72
            u128 base = codes[idx];
            fh = (fh >> 2) | (base << ((k - 1) * 2));
            bw = ((bw << 2) | (base ^ 2)) & mask;
            cb(idx - (k - 1), fh < bw ? fh : bw, fh < bw);
        }
    } else {
        u128 fh = 0, bw = 0;  // cn_rkhash_base.rs:65-110
        for (unsigned i = 0; i + 1 < k; i++) fh = fh * RK_MULT + rk_fwd(codes[i]);
        for (unsigned i = k - 1; i-- > 0;) bw = bw * RK_MULT + rk_bkw(codes[i]);
        bw *= RK_MULT;
        const u128 rmmult = pow128(RK_MULT, k - 1);
        for (size_t idx = k - 1; idx < len; idx++) {
            uint8_t in = codes[idx], out = codes[idx - (k - 1)];
            u128 cur_fh = fh * RK_MULT + rk_fwd(in);
            fh = cur_fh - rk_fwd(out) * rmmult;
            u128 cur_bk = bw * RK_MULT_INV + rk_bkw(in) * rmmult;
            bw = cur_bk - rk_bkw(out);
            cb(idx - (k - 1), cur_fh < cur_bk ? cur_fh : cur_bk, cur_fh < cur_bk);
        }
    }
}

struct U128Hash {
    size_t operator()(const u128& x) const {
        uint64_t a = (uint64_t)x, b = (uint64_t)(x >> 64);
        uint64_t h = a * 0x9E3779B97F4A7C15ULL ^ (b + 0x632BE59BD9B4E019ULL + (a << 6) + (a >> 2));
        return (size_t)(h ^ (h >> 31));
    }
};

// structs/src/map_entry.rs:6-85 (count 61 bits | used | 2 flag bits)
struct Entry {
    uint64_t count = 0;
    uint8_t flags = 0;
    std::vector<uint32_t> cols;  // colors of the occurrences (colored builds)
};

// Sorted unique union of an entry's occurrence colors: the k-mer's color
// subset (crates/colors/src/managers/multiple.rs:174-208, process_colors:
// sort_unstable + dedup of the concatenated per-occurrence color lists).
inline void append_color_set(std::vector<uint32_t>& cols, std::vector<uint32_t>& out_off,
                             std::vector<uint32_t>& out) {
    std::sort(cols.begin(), cols.end());
    cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
    out.insert(out.end(), cols.begin(), cols.end());
    out_off.push_back((uint32_t)out.size());
}

struct SolidSet {
    std::vector<uint32_t> sub;
    std::vector<uint64_t> key_lo, key_hi;
    std::vector<uint32_t> count;  // get_kmer_multiplicity()
    std::vector<uint8_t> flags;
    std::vector<uint32_t> colors_off;  // color-set offsets (colored runs)
    std::vector<uint32_t> colors;
    void append(const SolidSet& o) {
        if (!o.colors_off.empty()) {
            if (colors_off.empty()) colors_off.push_back(0);
            const uint32_t base = (uint32_t)colors.size();
            for (size_t i = 1; i < o.colors_off.size(); i++) colors_off.push_back(base + o.colors_off[i]);
            colors.insert(colors.end(), o.colors.begin(), o.colors.end());
        }
        sub.insert(sub.end(), o.sub.begin(), o.sub.end());
        key_lo.insert(key_lo.end(), o.key_lo.begin(), o.key_lo.end());
        key_hi.insert(key_hi.end(), o.key_hi.begin(), o.key_hi.end());
        count.insert(count.end(), o.count.begin(), o.count.end());
        flags.insert(flags.end(), o.flags.begin(), o.flags.end());
    }
};

// crates/assembler_kmers_merge/src/unitigs_extender/hashmap.rs:361-431
// add_sequence over every record of one sub-bucket, then the solid filter
// of hashmap.rs:134-137 / map_entry.rs:79-83.
void merge_subbucket(const RecordSet& R, const std::vector<size_t>& idx, uint32_t sub, unsigned k,
                     int kind, uint32_t min_mult, bool colored, SolidSet& out,
                     std::unordered_map<u128, Entry, U128Hash>& map, std::vector<uint8_t>& codes) {
    map.clear();
    if (colored && out.colors_off.empty()) out.colors_off.push_back(0);
    for (size_t r : idx) {
        uint32_t L = R.len[r];
        const uint8_t* b = R.bases.data() + R.off[r];
        codes.resize(L);
        for (uint32_t i = 0; i < L; i++) codes[i] = cbase(b[i]);
        const uint8_t fl = R.flags[r];
        const uint32_t mult = R.mult[r];
        const size_t last = L - k;
        kmer_hashes(codes.data(), L, k, kind, [&](size_t i, u128 key, bool is_fwd) {
            bool begin_ignored = (fl & 1) == 0 && i == 0;
            bool end_ignored = (fl & 2) == 0 && i == last;
            Entry& e = map[key];
            e.flags |= (uint8_t)(((begin_ignored ? 1 : 0) << (is_fwd ? 0 : 1)) |
                                 ((end_ignored ? 1 : 0) << (is_fwd ? 1 : 0)));
            e.count += mult;
            if (colored && (e.cols.empty() || e.cols.back() != R.color[r])) e.cols.push_back(R.color[r]);
        });
    }
    for (auto& kv : map) {
        uint64_t mult = kv.second.count >> (kv.second.flags == 3 ? 1 : 0);
        if (mult < min_mult) continue;
        out.sub.push_back(sub);
        out.key_lo.push_back((uint64_t)kv.first);
        out.key_hi.push_back((uint64_t)(kv.first >> 64));
        out.count.push_back((uint32_t)mult);
        out.flags.push_back(kv.second.flags);
        if (colored) append_color_set(kv.second.cols, out.colors_off, out.colors);
    }
}

// ---------------------------------------------------- sorting extender ---
// Normal sub-buckets (neither the duplicates bucket nor an outlier) are not
// counted with the hash map in the reference.  add_read
// (crates/assembler_kmers_merge/src/map_processor.rs:127-166) groups the
// sub-bucket's records by the hash of their minimizer m-mer in a
// FuzzyHashmap (crates/utils/src/fuzzy_hashmap.rs:5-127); process_map
// (crates/assembler_kmers_merge/src/final_executor.rs:316-437) outputs a
// group of one record as it is when its multiplicity reaches the cutoff and
// runs SortingExtender::process_reads
// (crates/assembler_kmers_merge/src/unitigs_extender/sorting.rs:641-1050)
// on larger groups.  process_reads sorts the records by the bases after and
// before their minimizer (centred suffix and prefix) and walks runs of
// records that share at least k bases around the minimizer
// (sorting.rs:420-444): a k-mer is identified by its bases together with
// its offset from the minimizer position, and the multiplicity of a
// supertig is the sum of the multiplicities of the records holding it there
// (:431).  A supertig of exactly one k-mer that is the last k-mer of a record
// without READ_FLAG_INCL_END and the first k-mer of a record without
// READ_FLAG_INCL_BEGIN is a confirmed double-counted k-mer: its
// multiplicity is halved and both extra-base flags are cleared (:454-486).
// A supertig is kept iff its multiplicity >= min_multiplicity (:501).
//
// Restated here per centred k-mer: groups are keyed by the exact minimizer
// m-mer (records whose minimizer m-mers differ share no centred k-mer, since
// the k shared bases would cover the minimizer, so the fuzzy grouping only
// changes which records are sorted together, not what they share), and the
// halving is applied to every centred k-mer that has both extra bases.  The
// second step equals the reference whenever such a k-mer is alone in its
// supertig; `both_extra` counts those k-mers (a centred k-mer cannot be the
// last k-mer of one group and the first of another: both groups would have
// the same minimizer at the same position, which is no split,
// crates/hashes/src/rolling/batch_minqueue.rs:155-160), and the tests check
// that it stays zero.  Output flags are the extra bases in stored
// orientation: bit 0 backward (first k-mer, no INCL_BEGIN), bit 1 forward.
struct CentredKey {
    u128 key;  // canonical k-mer key
    uint64_t minimizer;  // 2-bit packed minimizer m-mer (stored orientation), m <= 32
    int32_t offset;  // k-mer start - minimizer position
    uint8_t fwd;  // orientation of the stored k-mer relative to its canonical key
    bool operator==(const CentredKey& o) const {
        return key == o.key && minimizer == o.minimizer && offset == o.offset && fwd == o.fwd;
    }
};
struct CentredHash {
    size_t operator()(const CentredKey& c) const {
        return U128Hash()(c.key ^ ((u128)c.minimizer << 64)) ^ ((size_t)(uint32_t)c.offset * 0x9E3779B97F4A7C15ULL) ^
               c.fwd;
    }
};
struct CentredEntry {
    uint64_t count = 0;
    bool fwd_extra = false, bwd_extra = false;
    std::vector<uint32_t> cols;
};

struct SortingStats {
    uint64_t both_extra = 0;  // centred k-mers holding both extra bases (halved)
    uint64_t groups = 0;  // minimizer groups with more than one record
    uint64_t centred = 0;  // centred k-mers seen
};

void merge_subbucket_sorting(const RecordSet& R, const std::vector<size_t>& idx, uint32_t sub, unsigned k,
                             unsigned m, int kind, uint32_t min_mult, bool colored, SolidSet& out,
                             SortingStats& st, std::vector<uint8_t>& codes) {
    if (colored && out.colors_off.empty()) out.colors_off.push_back(0);
    std::unordered_map<uint64_t, std::vector<size_t>> groups;
    for (size_t r : idx) {
        const uint8_t* b = R.bases.data() + R.off[r] + R.minpos[r];
        uint64_t mm = 0;
        for (unsigned j = 0; j < m; j++) mm |= (uint64_t)cbase(b[j]) << (2 * j);
        groups[mm].push_back(r);
    }
    std::unordered_map<CentredKey, CentredEntry, CentredHash> map;
    for (auto& g : groups) {
        map.clear();
        if (g.second.size() > 1) st.groups++;
        for (size_t r : g.second) {
            const uint32_t L = R.len[r];
            const uint8_t* b = R.bases.data() + R.off[r];
            codes.resize(L);
            for (uint32_t i = 0; i < L; i++) codes[i] = cbase(b[i]);
            const uint8_t fl = R.flags[r];
            const uint32_t mult = R.mult[r];
            const size_t last = L - k;
            const int32_t p = R.minpos[r];
            kmer_hashes(codes.data(), L, k, kind, [&](size_t i, u128 key, bool is_fwd) {
                CentredEntry& e = map[CentredKey{key, g.first, (int32_t)i - p, (uint8_t)is_fwd}];
                e.count += mult;
                if (i == last && (fl & 2) == 0) e.fwd_extra = true;
                if (i == 0 && (fl & 1) == 0) e.bwd_extra = true;
                if (colored && (e.cols.empty() || e.cols.back() != R.color[r])) e.cols.push_back(R.color[r]);
            });
        }
        for (auto& kv : map) {
            st.centred++;
            CentredEntry& e = kv.second;
            uint64_t mult = e.count;
            uint8_t flags = (uint8_t)((e.bwd_extra ? 1 : 0) | (e.fwd_extra ? 2 : 0));
            // a group of one record is output as it is (final_executor.rs:316-320)
            if (e.fwd_extra && e.bwd_extra && g.second.size() > 1) {
                st.both_extra++;
                mult /= 2;
                flags = 0;
            }
            if (mult < min_mult) continue;
            out.sub.push_back(sub);
            out.key_lo.push_back((uint64_t)kv.first.key);
            out.key_hi.push_back((uint64_t)(kv.first.key >> 64));
            out.count.push_back((uint32_t)mult);
            out.flags.push_back(flags);
            if (colored) append_color_set(e.cols, out.colors_off, out.colors);
        }
    }
}

// ----------------------------------------------------- unitig chains ----
// Intra-sub-bucket unitig extension over the solid seq-hash table
// (crates/assembler_kmers_merge/src/unitigs_extender/hashmap.rs:162-297,
// 442-601).  Two restatements:
//  * unitig_chains: the order-free rule the device implements.  In a k-mer's
//    canonical orientation side 0 is the predecessor side, side 1 the
//    successor side; deg = number of solid k-mers of the same sub-bucket one
//    base away on that side (manual_roll_forward / _reverse,
//    crates/hashes/src/base/cn_seqhash_base.rs:165-200).  Sides (p,s) and
//    (q,t) are linked when each is the other's only neighbour, p != q, and
//    neither side is the flagged side of a k-mer whose flags are exactly
//    INCL_BEGIN (side 0) or INCL_END (side 1): the reference stops after such
//    a k-mer (contig_break, hashmap.rs:287-292) and never extends from its
//    flagged side (begin_ignored / end_ignored, :483-494).  Unitigs are the
//    maximal linked paths; an end is open when its outer side is flagged.
//    Cycles start at their smallest key, forward.
//  * unitig_sequential: the reference's greedy loop itself, started from
//    every unused k-mer in a given order, with its used marks.  It agrees
//    with the order-free rule except where a flagged side has a neighbour in
//    the same sub-bucket, the case in which the reference's own result
//    depends on which k-mer it starts from.
struct UnitigOut {
    std::vector<uint8_t> seq;  // ASCII, each unitig followed by '\n'
    std::vector<uint32_t> sub;
    std::vector<uint8_t> flags;  // bit 0 open start, bit 1 open end, bit 2 circular
};

inline u128 kmer_mask(unsigned k) { return k >= 64 ? ~(u128)0 : (((u128)1 << (2 * k)) - 1); }
inline u128 kmer_rc(u128 x, unsigned k) {
    u128 r = 0;
    for (unsigned i = 0; i < k; i++) r |= (u128)(((x >> (2 * i)) & 3) ^ 2) << (2 * (k - 1 - i));
    return r;
}
inline void put_kmer(std::vector<uint8_t>& out, u128 x, unsigned k) {
    static const char B[4] = {'A', 'C', 'T', 'G'};
    for (unsigned i = 0; i < k; i++) out.push_back((uint8_t)B[(x >> (2 * i)) & 3]);
}

struct SubGraph {
    unsigned k;
    std::vector<u128> key;
    std::vector<uint8_t> flags;
    std::unordered_map<u128, uint32_t, U128Hash> at;
    // neighbour of side s of node p: (q, t) with the number of candidates
    uint32_t neighbours(uint32_t p, int s, uint32_t* q, int* t) const {
        const u128 x = key[p], mask = kmer_mask(k);
        uint32_t cnt = 0;
        for (uint32_t b = 0; b < 4; b++) {
            const u128 y = s ? ((x >> 2) | ((u128)b << (2 * (k - 1)))) : (((x << 2) | b) & mask);
            const u128 yr = kmer_rc(y, k);
            const u128 c = y < yr ? y : yr;
            auto it = at.find(c);
            if (it == at.end()) continue;
            cnt++;
            *q = it->second;
            // moving forward we enter q at its predecessor side when y is q's
            // canonical orientation; moving backward, at its successor side
            *t = (y == c) ? (s ? 0 : 1) : (s ? 1 : 0);
        }
        return cnt;
    }
    bool flagged(uint32_t p, int s) const { return (flags[p] == 1 && s == 0) || (flags[p] == 2 && s == 1); }
};

void unitig_chains(const SubGraph& G, uint32_t sub, UnitigOut& out) {
    const uint32_t n = (uint32_t)G.key.size();
    const unsigned k = G.k;
    std::vector<int64_t> link(2 * (size_t)n, -1);
    for (uint32_t p = 0; p < n; p++)
        for (int s = 0; s < 2; s++) {
            uint32_t q = 0;
            int t = 0;
            if (G.neighbours(p, s, &q, &t) != 1 || q == p || G.flagged(p, s) || G.flagged(q, t)) continue;
            uint32_t p2 = 0;
            int s2 = 0;
            if (G.neighbours(q, t, &p2, &s2) != 1 || p2 != p || s2 != s) continue;
            link[2 * p + s] = 2 * (int64_t)q + t;
        }
    std::vector<uint8_t> seen(n, 0);
    auto walk = [&](uint32_t p, int s_in, bool circular) {
        // enter p at side s_in; forward orientation when entering at side 0
        const size_t start = out.seq.size();
        u128 x = s_in == 0 ? G.key[p] : kmer_rc(G.key[p], k);
        put_kmer(out.seq, x, k);
        const bool open_start = G.flagged(p, s_in);
        seen[p] = 1;
        uint32_t cur = p;
        int side_out = s_in ^ 1;
        uint32_t steps = 1;
        while (link[2 * cur + side_out] >= 0) {
            const int64_t e = link[2 * cur + side_out];
            const uint32_t q = (uint32_t)(e >> 1);
            const int t = (int)(e & 1);
            if (seen[q]) break;  // back at the start of a cycle
            seen[q] = 1;
            const u128 y = t == 0 ? G.key[q] : kmer_rc(G.key[q], k);
            static const char B[4] = {'A', 'C', 'T', 'G'};
            out.seq.push_back((uint8_t)B[(y >> (2 * (k - 1))) & 3]);
            cur = q;
            side_out = t ^ 1;
            steps++;
        }
        (void)start;
        const bool open_end = !circular && G.flagged(cur, side_out);
        out.seq.push_back('\n');
        out.sub.push_back(sub);
        out.flags.push_back((uint8_t)((open_start && !circular ? 1 : 0) | (open_end ? 2 : 0) | (circular ? 4 : 0)));
    };
    // linear chains from the end with the smaller (key, side)
    for (uint32_t p = 0; p < n; p++) {
        if (seen[p]) continue;
        const bool end0 = link[2 * p] < 0, end1 = link[2 * p + 1] < 0;
        if (!end0 && !end1) continue;
        // find the other end of this chain
        int s_in = end0 ? 0 : 1;
        uint32_t cur = p;
        int side_out = s_in ^ 1;
        while (link[2 * cur + side_out] >= 0) {
            const int64_t e = link[2 * cur + side_out];
            cur = (uint32_t)(e >> 1);
            side_out = (int)(e & 1) ^ 1;
        }
        // chain ends: (p, s_in) and (cur, side_out); start from the smaller
        const u128 kp = G.key[p], kc = G.key[cur];
        if (kc < kp || (kc == kp && side_out < s_in))
            walk(cur, side_out, false);
        else
            walk(p, s_in, false);
    }
    // cycles: smallest key, forward
    std::vector<uint32_t> rest;
    for (uint32_t p = 0; p < n; p++)
        if (!seen[p]) rest.push_back(p);
    std::sort(rest.begin(), rest.end(), [&](uint32_t a, uint32_t b) { return G.key[a] < G.key[b]; });
    for (uint32_t p : rest)
        if (!seen[p]) walk(p, 0, true);
}

void unitig_sequential(const SubGraph& G, uint32_t sub, const std::vector<uint32_t>& order, UnitigOut& out) {
    const unsigned k = G.k;
    std::vector<uint8_t> used(G.key.size(), 0);
    static const char B[4] = {'A', 'C', 'T', 'G'};
    for (uint32_t p : order) {
        if (used[p]) continue;
        used[p] = 1;
        // extend from side s of p (forward string grows at its end)
        auto extend = [&](int s, std::vector<uint8_t>& bases) -> bool {
            uint32_t cur = p;
            int side_out = s;
            while (true) {
                uint32_t q = 0;
                int t = 0;
                if (G.neighbours(cur, side_out, &q, &t) != 1) return false;
                uint32_t back = 0;
                int bs = 0;
                if (G.neighbours(q, t, &back, &bs) != 1) return false;
                if (used[q]) return false;
                used[q] = 1;
                const u128 y = t == 0 ? G.key[q] : kmer_rc(G.key[q], k);
                bases.push_back((uint8_t)B[(y >> (2 * (k - 1))) & 3]);
                cur = q;
                side_out = t ^ 1;
                if (G.flags[q] == 1 || G.flags[q] == 2) return true;
            }
        };
        std::vector<uint8_t> fw, bw;
        const bool fw_open = G.flagged(p, 1) ? true : extend(1, fw);
        // backward: extend the reverse complement forward
        bool bw_open;
        if (G.flagged(p, 0)) {
            bw_open = true;
        } else {
            bw_open = extend(0, bw);  // bases of the reverse strand
        }
        // sequence: rc(bw bases) + kmer + fw bases
        for (size_t i = bw.size(); i-- > 0;) {
            const uint8_t c = bw[i];
            out.seq.push_back(c == 'A' ? 'T' : c == 'C' ? 'G' : c == 'G' ? 'C' : 'A');
        }
        put_kmer(out.seq, G.key[p], k);
        out.seq.insert(out.seq.end(), fw.begin(), fw.end());
        out.seq.push_back('\n');
        out.sub.push_back(sub);
        out.flags.push_back((uint8_t)((bw_open ? 1 : 0) | (fw_open ? 2 : 0)));
    }
}

struct Handle {
    RecordSet rec;
    SolidSet solid;
    SolidSet solid_sorting;  // normal sub-buckets under the sorting extender's rule
    SortingStats sorting_stats;
    UnitigOut unitigs;
    bool colored = false;
    double t_bucket = 0, t_merge = 0;
};

std::vector<std::vector<size_t>> group_by_sub(const RecordSet& R, uint32_t n_sub) {
    std::vector<std::vector<size_t>> g(n_sub);
    for (size_t i = 0; i < R.sub.size(); i++) g[R.sub[i]].push_back(i);
    return g;
}

}  // namespace

// ------------------------------------------------------------- C API ----
extern "C" {

// ntHash over an ASCII sequence: fw/rc arrays of len-m+1 entries.
void orc_nthash(const uint8_t* seq, size_t len, unsigned m, uint64_t* fw, uint64_t* rc) {
    nthash_m(seq, len, m, fw, rc);
}

// BatchMinQueue::get_minimizers over explicit (v, index, is_forward) items.
// Writes one (v, index, is_forward) per window; returns the window count.
size_t orc_minqueue(const uint64_t* v, const uint32_t* index, const uint8_t* fwd, size_t n, size_t w,
                    uint64_t* out_v, uint32_t* out_index, uint8_t* out_fwd) {
    std::vector<Item> it(n);
    for (size_t i = 0; i < n; i++) {
        it[i].v = v[i];
        it[i].x.index = index[i];
        it[i].x.is_forward = fwd[i] != 0;
    }
    BatchMinQueue q(w);
    size_t cnt = 0;
    q.get_minimizers(it.data(), n, 0, [&](const Item& mm, size_t off) {
        out_v[off] = mm.v;
        out_index[off] = mm.x.index;
        out_fwd[off] = mm.x.is_forward;
        cnt = std::max(cnt, off + 1);
    }, [](bool) {});
    return cnt;
}

// Full bucketing of a record collection with `threads` workers.
void* orc_bucketing(const uint8_t* seq, const uint64_t* rec_off, size_t n_rec, const uint32_t* rec_color,
                    uint32_t k, uint32_t m, uint32_t b, uint32_t sb, int canonical, int threads) {
    Handle* h = new Handle();
    Params P{k, m, b, sb, canonical};
    if (threads < 1) threads = 1;
    auto t0 = std::chrono::steady_clock::now();
    std::vector<RecordSet> parts(threads);
    std::vector<std::thread> th;
    for (int t = 0; t < threads; t++) {
        size_t r0 = n_rec * t / threads, r1 = n_rec * (t + 1) / threads;
        th.emplace_back([&, t, r0, r1] { bucket_records(seq, rec_off, r0, r1, rec_color, P, parts[t]); });
    }
    for (auto& x : th) x.join();
    for (auto& p : parts) h->rec.append(p);
    h->colored = rec_color != nullptr;
    h->t_bucket = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return h;
}

// crates/minimizer_bucketing/src/compactor.rs:128-196: identical records
// (same sub-bucket, length, packed bases and flags) merged, multiplicities
// summed.  Produces the fully compacted form of a record set.
void orc_compact(void* hv) {
    Handle* h = (Handle*)hv;
    RecordSet& R = h->rec;
    size_t n = R.sub.size();
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; i++) order[i] = i;
    auto cmp = [&](size_t a, size_t b) {
        if (R.sub[a] != R.sub[b]) return R.sub[a] < R.sub[b];
        if (R.len[a] != R.len[b]) return R.len[a] < R.len[b];
        if (R.flags[a] != R.flags[b]) return R.flags[a] < R.flags[b];
        int c = memcmp(R.bases.data() + R.off[a], R.bases.data() + R.off[b], R.len[a]);
        if (c != 0) return c < 0;
        return R.color[a] < R.color[b];
    };
    std::sort(order.begin(), order.end(), cmp);
    RecordSet O;
    for (size_t i = 0; i < n;) {
        size_t j = i;
        uint32_t mult = 0;
        size_t a = order[i];
        while (j < n) {
            size_t c = order[j];
            if (R.sub[c] != R.sub[a] || R.len[c] != R.len[a] || R.flags[c] != R.flags[a] ||
                memcmp(R.bases.data() + R.off[c], R.bases.data() + R.off[a], R.len[a]) != 0 ||
                (h->colored && R.color[c] != R.color[a]))
                break;
            mult += R.mult[c];
            j++;
        }
        O.sub.push_back(R.sub[a]);
        O.len.push_back(R.len[a]);
        O.flags.push_back(R.flags[a]);
        O.minpos.push_back(R.minpos[a]);
        O.mult.push_back(mult);
        O.color.push_back(R.color[a]);
        O.off.push_back(O.bases.size());
        O.bases.insert(O.bases.end(), R.bases.begin() + R.off[a], R.bases.begin() + R.off[a] + R.len[a]);
        i = j;
    }
    R = std::move(O);
}

size_t orc_rec_count(void* hv) { return ((Handle*)hv)->rec.sub.size(); }
size_t orc_rec_bases(void* hv) { return ((Handle*)hv)->rec.bases.size(); }
void orc_rec_get(void* hv, uint32_t* sub, uint32_t* len, uint8_t* flags, uint16_t* minpos, uint32_t* mult,
                 uint64_t* off, uint8_t* bases) {
    RecordSet& R = ((Handle*)hv)->rec;
    size_t n = R.sub.size();
    memcpy(sub, R.sub.data(), n * 4);
    memcpy(len, R.len.data(), n * 4);
    memcpy(flags, R.flags.data(), n);
    memcpy(minpos, R.minpos.data(), n * 2);
    memcpy(mult, R.mult.data(), n * 4);
    memcpy(off, R.off.data(), n * 8);
    memcpy(bases, R.bases.data(), R.bases.size());
}

void orc_rec_colors(void* hv, uint32_t* color) {
    RecordSet& R = ((Handle*)hv)->rec;
    memcpy(color, R.color.data(), R.color.size() * 4);
}

// Replaces the records of each listed outlier sub-bucket by its resplit
// records (resplit_records).  Outlier i is split into counts[i] (a power of
// two) normal buckets plus its duplicates bucket, which get the sub-bucket
// ids base_i .. base_i + counts[i] (the duplicates bucket last), with
// base_0 = first_extra and base_{i+1} = base_i + counts[i] + 1.
void orc_resplit(void* hv, uint32_t k, uint32_t m, const uint32_t* subs, const uint32_t* counts, size_t n,
                 uint32_t first_extra) {
    Handle* h = (Handle*)hv;
    RecordSet& R = h->rec;
    std::unordered_map<uint32_t, size_t> which;
    for (size_t i = 0; i < n; i++) which[subs[i]] = i;
    std::vector<std::vector<size_t>> idx(n);
    RecordSet kept;
    for (size_t r = 0; r < R.sub.size(); r++) {
        auto it = which.find(R.sub[r]);
        if (it != which.end()) {
            idx[it->second].push_back(r);
            continue;
        }
        kept.sub.push_back(R.sub[r]);
        kept.len.push_back(R.len[r]);
        kept.flags.push_back(R.flags[r]);
        kept.minpos.push_back(R.minpos[r]);
        kept.mult.push_back(R.mult[r]);
        kept.color.push_back(R.color[r]);
        kept.off.push_back(kept.bases.size());
        kept.bases.insert(kept.bases.end(), R.bases.begin() + R.off[r], R.bases.begin() + R.off[r] + R.len[r]);
    }
    uint32_t base = first_extra;
    for (size_t i = 0; i < n; i++) {
        uint32_t lg = 0;
        while ((1u << lg) < counts[i]) lg++;
        resplit_records(R, idx[i], k, m, lg, base, kept);
        base += (1u << lg) + 1;
    }
    R = std::move(kept);
}

// Per-sub-bucket MapEntry merge over the handle's records, `threads` workers
// (sub-bucket ids past the bucketing's, from a resplit, included).
void orc_merge(void* hv, uint32_t k, uint32_t b, uint32_t sb, int kind, uint32_t min_mult, int threads) {
    Handle* h = (Handle*)hv;
    auto t0 = std::chrono::steady_clock::now();
    uint32_t n_sub = ((1u << b) + 1) << sb;
    for (uint32_t s : h->rec.sub) n_sub = std::max(n_sub, s + 1);
    auto groups = group_by_sub(h->rec, n_sub);
    if (threads < 1) threads = 1;
    std::vector<SolidSet> parts(threads);
    std::atomic<uint32_t> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < threads; t++) {
        th.emplace_back([&, t] {
            std::unordered_map<u128, Entry, U128Hash> map;
            std::vector<uint8_t> codes;
            for (;;) {
                uint32_t s = next.fetch_add(1);
                if (s >= n_sub) break;
                if (groups[s].empty()) continue;
                merge_subbucket(h->rec, groups[s], s, k, kind, min_mult, h->colored, parts[t], map, codes);
            }
        });
    }
    for (auto& x : th) x.join();
    h->solid = SolidSet();
    for (auto& p : parts) h->solid.append(p);
    h->t_merge = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// The sorting extender's solid k-mers (see merge_subbucket_sorting) for the
// sub-buckets of the normal buckets (the duplicates bucket 2^b is counted
// with the hash map in the reference too); stats[0..3] = centred k-mers
// with both extra bases, minimizer groups of more than one record, centred
// k-mers.  Entries stay in the handle (orc_sorting_count / orc_sorting_get).
void orc_merge_sorting(void* hv, uint32_t k, uint32_t m, uint32_t b, uint32_t sb, int kind, uint32_t min_mult,
                       int threads, uint64_t* stats) {
    Handle* h = (Handle*)hv;
    const uint32_t n_normal = (1u << b) << sb;
    auto groups = group_by_sub(h->rec, ((1u << b) + 1) << sb);
    if (threads < 1) threads = 1;
    std::vector<SolidSet> parts(threads);
    std::vector<SortingStats> pst(threads);
    std::atomic<uint32_t> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < threads; t++) {
        th.emplace_back([&, t] {
            std::vector<uint8_t> codes;
            for (;;) {
                uint32_t s = next.fetch_add(1);
                if (s >= n_normal) break;
                if (groups[s].empty()) continue;
                merge_subbucket_sorting(h->rec, groups[s], s, k, m, kind, min_mult, h->colored, parts[t], pst[t],
                                        codes);
            }
        });
    }
    for (auto& x : th) x.join();
    h->solid_sorting = SolidSet();
    h->sorting_stats = SortingStats();
    for (int t = 0; t < threads; t++) {
        h->solid_sorting.append(parts[t]);
        h->sorting_stats.both_extra += pst[t].both_extra;
        h->sorting_stats.groups += pst[t].groups;
        h->sorting_stats.centred += pst[t].centred;
    }
    if (stats) {
        stats[0] = h->sorting_stats.both_extra;
        stats[1] = h->sorting_stats.groups;
        stats[2] = h->sorting_stats.centred;
    }
}
size_t orc_sorting_count(void* hv) { return ((Handle*)hv)->solid_sorting.sub.size(); }
void orc_sorting_get(void* hv, uint32_t* sub, uint64_t* key_lo, uint64_t* key_hi, uint32_t* count, uint8_t* flags) {
    SolidSet& S = ((Handle*)hv)->solid_sorting;
    size_t n = S.sub.size();
    memcpy(sub, S.sub.data(), n * 4);
    memcpy(key_lo, S.key_lo.data(), n * 8);
    memcpy(key_hi, S.key_hi.data(), n * 8);
    memcpy(count, S.count.data(), n * 4);
    memcpy(flags, S.flags.data(), n);
}

// Unitig chains over the merged solid table (seq-hash keys).  mode 0: the
// order-free rule; mode 1: the reference's greedy loop, starting k-mers in
// ascending key order per sub-bucket; mode 2: the maximal unitigs of the
// whole solid k-mer set (the output of build_final_unitigs,
// crates/assembler_pipeline/src/lib.rs:111, extend_unitigs.rs:348-993): one
// node per distinct k-mer over all sub-buckets, no side flagged, sub-bucket
// 0.  Returns the number of unitigs; the data stays in the handle
// (orc_unitigs_get).
size_t orc_unitigs(void* hv, uint32_t k, int mode, int threads) {
    Handle* h = (Handle*)hv;
    const SolidSet& S = h->solid;
    h->unitigs = UnitigOut();
    if (mode == 2) {
        SubGraph G;
        G.k = k;
        for (size_t i = 0; i < S.sub.size(); i++) {
            const u128 key = ((u128)S.key_hi[i] << 64) | S.key_lo[i];
            if (!G.at.emplace(key, (uint32_t)G.key.size()).second) continue;
            G.key.push_back(key);
            G.flags.push_back(0);
        }
        unitig_chains(G, 0, h->unitigs);
        return h->unitigs.sub.size();
    }
    std::map<uint32_t, std::vector<size_t>> by_sub;
    for (size_t i = 0; i < S.sub.size(); i++) by_sub[S.sub[i]].push_back(i);
    std::vector<const std::pair<const uint32_t, std::vector<size_t>>*> subs;
    for (auto& kv : by_sub) subs.push_back(&kv);
    // sub-buckets are independent: `threads` workers take contiguous runs of
    // them and their outputs are concatenated in sub-bucket order
    if (threads < 1) threads = 1;
    std::vector<UnitigOut> parts(threads);
    std::vector<std::thread> th;
    for (int t = 0; t < threads; t++) {
        th.emplace_back([&, t] {
            const size_t a = subs.size() * t / threads, b = subs.size() * (t + 1) / threads;
            for (size_t s = a; s < b; s++) {
                const auto& kv = *subs[s];
                SubGraph G;
                G.k = k;
                for (size_t i : kv.second) {
                    const u128 key = ((u128)S.key_hi[i] << 64) | S.key_lo[i];
                    G.at.emplace(key, (uint32_t)G.key.size());
                    G.key.push_back(key);
                    G.flags.push_back(S.flags[i]);
                }
                if (mode == 0) {
                    unitig_chains(G, kv.first, parts[t]);
                } else {
                    std::vector<uint32_t> order(G.key.size());
                    for (uint32_t i = 0; i < order.size(); i++) order[i] = i;
                    std::sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return G.key[x] < G.key[y]; });
                    unitig_sequential(G, kv.first, order, parts[t]);
                }
            }
        });
    }
    for (auto& x : th) x.join();
    for (auto& p : parts) {
        h->unitigs.seq.insert(h->unitigs.seq.end(), p.seq.begin(), p.seq.end());
        h->unitigs.sub.insert(h->unitigs.sub.end(), p.sub.begin(), p.sub.end());
        h->unitigs.flags.insert(h->unitigs.flags.end(), p.flags.begin(), p.flags.end());
    }
    return h->unitigs.sub.size();
}
// In/out extension flags of every solid entry (solid order): degree of side
// 0 in bits 0-1 and of side 1 in bits 2-3 (2 = two or more), counted among
// the solid k-mers of the same sub-bucket.
void orc_extension(void* hv, uint32_t k, uint8_t* ext) {
    Handle* h = (Handle*)hv;
    const SolidSet& S = h->solid;
    std::map<uint32_t, std::vector<size_t>> by_sub;
    for (size_t i = 0; i < S.sub.size(); i++) by_sub[S.sub[i]].push_back(i);
    for (auto& kv : by_sub) {
        SubGraph G;
        G.k = k;
        for (size_t i : kv.second) {
            const u128 key = ((u128)S.key_hi[i] << 64) | S.key_lo[i];
            G.at.emplace(key, (uint32_t)G.key.size());
            G.key.push_back(key);
            G.flags.push_back(S.flags[i]);
        }
        for (uint32_t j = 0; j < G.key.size(); j++) {
            uint32_t q = 0;
            int t = 0;
            const uint32_t d0 = G.neighbours(j, 0, &q, &t), d1 = G.neighbours(j, 1, &q, &t);
            ext[kv.second[j]] = (uint8_t)((d0 > 2 ? 2 : d0) | ((d1 > 2 ? 2 : d1) << 2));
        }
    }
}

size_t orc_unitigs_bytes(void* hv) { return ((Handle*)hv)->unitigs.seq.size(); }
void orc_unitigs_get(void* hv, uint8_t* seq, uint32_t* sub, uint8_t* flags) {
    UnitigOut& U = ((Handle*)hv)->unitigs;
    memcpy(seq, U.seq.data(), U.seq.size());
    memcpy(sub, U.sub.data(), U.sub.size() * 4);
    memcpy(flags, U.flags.data(), U.flags.size());
}

size_t orc_solid_count(void* hv) { return ((Handle*)hv)->solid.sub.size(); }
void orc_solid_get(void* hv, uint32_t* sub, uint64_t* key_lo, uint64_t* key_hi, uint32_t* count, uint8_t* flags) {
    SolidSet& S = ((Handle*)hv)->solid;
    size_t n = S.sub.size();
    memcpy(sub, S.sub.data(), n * 4);
    memcpy(key_lo, S.key_lo.data(), n * 8);
    memcpy(key_hi, S.key_hi.data(), n * 8);
    memcpy(count, S.count.data(), n * 4);
    memcpy(flags, S.flags.data(), n);
}
// Color subsets of the solid entries (colored handles): offsets [n+1] into
// the concatenated sorted color lists.
size_t orc_solid_colors_total(void* hv) { return ((Handle*)hv)->solid.colors.size(); }
void orc_solid_colors_get(void* hv, uint32_t* off, uint32_t* colors) {
    SolidSet& S = ((Handle*)hv)->solid;
    if (S.colors_off.empty()) {
        for (size_t i = 0; i <= S.sub.size(); i++) off[i] = 0;
        return;
    }
    memcpy(off, S.colors_off.data(), S.colors_off.size() * 4);
    memcpy(colors, S.colors.data(), S.colors.size() * 4);
}
void orc_times(void* hv, double* t_bucket, double* t_merge) {
    *t_bucket = ((Handle*)hv)->t_bucket;
    *t_merge = ((Handle*)hv)->t_merge;
}
void orc_free(void* hv) { delete (Handle*)hv; }

// Query against the handle's solid k-mers (the graph's k-mer set).
// crates/querier/src/lib.rs:138-150: query_kmers = len - (k-1), saturating,
// over the raw record length; crates/querier/src/pipeline/
// parallel_kmers_query.rs:372-402,437-471: every k-mer occurrence of an ACGT
// run of the query (runs split at other bytes, as in minimizer bucketing)
// whose canonical hash is in the graph counts once for the query and once
// for the graph k-mer's color subset; colored_query_output.rs:200-320 turns
// subset counts into per-color sums.  Queries: records of a joined buffer
// (each followed by '\n').  Outputs per query: query_kmers[q], matched[q];
// colored handles also fill color_counts[q * n_colors + c].
uint64_t orc_query(void* hv, const uint8_t* joined, uint64_t n_bytes, unsigned k, int kind, uint32_t n_colors,
                   uint64_t* query_kmers, uint64_t* matched, uint64_t* color_counts, uint64_t max_queries) {
    Handle* h = (Handle*)hv;
    const SolidSet& S = h->solid;
    std::unordered_map<u128, size_t, U128Hash> graph;
    graph.reserve(S.sub.size() * 2);
    for (size_t i = 0; i < S.sub.size(); i++) graph.emplace(((u128)S.key_hi[i] << 64) | S.key_lo[i], i);
    const bool colored = h->colored && !S.colors_off.empty();
    uint64_t q = 0, start = 0;
    std::vector<uint8_t> codes;
    for (uint64_t p = 0; p < n_bytes && q < max_queries; p++) {
        if (joined[p] != '\n') continue;
        const uint8_t* s = joined + start;
        const uint64_t len = p - start;
        query_kmers[q] = len >= k ? len - (k - 1) : 0;
        matched[q] = 0;
        if (colored)
            for (uint32_t c = 0; c < n_colors; c++) color_counts[q * n_colors + c] = 0;
        uint64_t i = 0;
        while (i < len) {
            while (i < len && normalize(s[i]) == 'N') i++;
            const uint64_t st = i;
            while (i < len && normalize(s[i]) != 'N') i++;
            if (i - st < k) continue;
            codes.resize(i - st);
            for (uint64_t j = st; j < i; j++) codes[j - st] = cbase(normalize(s[j]));
            kmer_hashes(codes.data(), i - st, k, kind, [&](size_t, u128 key, bool) {
                auto it = graph.find(key);
                if (it == graph.end()) return;
                matched[q]++;
                if (colored) {
                    const size_t e = it->second;
                    for (uint32_t o = S.colors_off[e]; o < S.colors_off[e + 1]; o++)
                        if (S.colors[o] < n_colors) color_counts[q * n_colors + S.colors[o]]++;
                }
            });
        }
        q++;
        start = p + 1;
    }
    return q;
}

// Canonical k-mer keys of one ASCII sequence (all k-mers, no N handling).
void orc_kmer_keys(const uint8_t* ascii, size_t len, unsigned k, int kind, uint64_t* key_lo, uint64_t* key_hi,
                   uint8_t* fwd) {
    std::vector<uint8_t> codes(len);
    for (size_t i = 0; i < len; i++) codes[i] = cbase(ascii[i]);
    kmer_hashes(codes.data(), len, k, kind, [&](size_t i, u128 key, bool f) {
        key_lo[i] = (uint64_t)key;
        key_hi[i] = (uint64_t)(key >> 64);
        fwd[i] = f;
    });
}

// Brute force: exact canonical (2-bit) k-mer occurrence counts over the
// N-free runs of every record; writes keys with count >= min_mult.
// With rec_color, each solid key also gets the sorted set of the colors of
// the records it occurs in.
void* orc_bruteforce_counts(const uint8_t* seq, const uint64_t* rec_off, size_t n_rec, const uint32_t* rec_color,
                            unsigned k, uint32_t min_mult) {
    Handle* h = new Handle();
    h->colored = rec_color != nullptr;
    std::unordered_map<u128, Entry, U128Hash> map;
    std::vector<uint8_t> codes;
    for (size_t r = 0; r < n_rec; r++) {
        const uint8_t* s = seq + rec_off[r];
        size_t len = rec_off[r + 1] - rec_off[r];
        size_t i = 0;
        while (i < len) {
            while (i < len && normalize(s[i]) == 'N') i++;
            size_t st = i;
            while (i < len && normalize(s[i]) != 'N') i++;
            if (i - st < k) continue;
            codes.resize(i - st);
            for (size_t j = st; j < i; j++) codes[j - st] = cbase(normalize(s[j]));
            kmer_hashes(codes.data(), i - st, k, 0, [&](size_t, u128 key, bool) {
                Entry& e = map[key];
                e.count++;
                if (rec_color) e.cols.push_back(rec_color[r]);
            });
        }
    }
    if (rec_color) h->solid.colors_off.push_back(0);
    for (auto& kv : map) {
        if (kv.second.count < min_mult) continue;
        h->solid.sub.push_back(0);
        h->solid.key_lo.push_back((uint64_t)kv.first);
        h->solid.key_hi.push_back((uint64_t)(kv.first >> 64));
        h->solid.count.push_back((uint32_t)kv.second.count);
        h->solid.flags.push_back(0);
        if (rec_color) append_color_set(kv.second.cols, h->solid.colors_off, h->solid.colors);
    }
    return h;
}

}  // extern "C"
