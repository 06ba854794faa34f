// query.hip -- colored / uncolored k-mer queries against the solid k-mer
// table of the last merge, kept resident in HBM.
//
// Semantics (crates/querier/src/pipeline/parallel_kmers_query.rs:372-471,
// counters_sorting.rs:106-273, colored_query_output.rs:200-320,
// crates/querier/src/lib.rs:138-150): every k-mer occurrence of an ACGT run
// of a query whose canonical hash is a graph k-mer counts once for the query
// and once for each color of that k-mer's color set.  The graph k-mer set is
// the solid table of the build (the k-mers of its unitigs).
//
// Index: the solid entries hashed into 2^L buckets (about 4 entries each) and
// stored bucket-major (key words + position of the entry in the solid
// arrays), built with a count / scan / scatter like the bucketing.  A lookup
// reads the bucket bounds and scans one short run of keys.
#include "common.h"

namespace gg {

GG_DEV uint64_t mix64(uint64_t x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ULL;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebULL;
    x ^= x >> 31;
    return x;
}

GG_DEV uint64_t index_hash(uint64_t lo, uint64_t hi) { return mix64(lo ^ mix64(hi + 0x9E3779B97F4A7C15ULL)); }

struct IndexArgs {
    const uint64_t* out_key;  // solid keys, key_words per position
    const uint32_t* chunk_fill;  // valid entries of each output chunk
    uint32_t n_chunks;  // chunks before the tail
    uint64_t n_tail;  // entries after the chunks
    uint32_t chunk;  // entries per chunk
    uint32_t key_words;
    uint32_t log_buckets;
    uint32_t* bcount;
    const uint64_t* boff;
    uint32_t* bcursor;
    uint64_t* idx_key;
    uint32_t* idx_pos;
};

GG_DEV bool solid_valid(const IndexArgs& A, uint64_t p) {
    const uint64_t c = p / A.chunk;
    if (c < A.n_chunks) return (p - c * A.chunk) < A.chunk_fill[c];
    return p - (uint64_t)A.n_chunks * A.chunk < A.n_tail;
}

GG_DEV uint32_t index_bucket(const IndexArgs& A, uint64_t p) {
    const uint64_t lo = A.out_key[p * A.key_words];
    const uint64_t hi = A.key_words == 2 ? A.out_key[p * 2 + 1] : 0;
    return A.log_buckets ? (uint32_t)(index_hash(lo, hi) >> (64 - A.log_buckets)) : 0u;
}

__global__ void index_count_kernel(IndexArgs A, uint64_t n_pos) {
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < n_pos; p += (uint64_t)gridDim.x * blockDim.x)
        if (solid_valid(A, p)) atomicAdd(&A.bcount[index_bucket(A, p)], 1u);
}

__global__ void index_fill_kernel(IndexArgs A, uint64_t n_pos) {
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < n_pos; p += (uint64_t)gridDim.x * blockDim.x) {
        if (!solid_valid(A, p)) continue;
        const uint32_t b = index_bucket(A, p);
        const uint64_t slot = A.boff[b] + atomicAdd(&A.bcursor[b], 1u);
        for (uint32_t q = 0; q < A.key_words; q++) A.idx_key[slot * A.key_words + q] = A.out_key[p * A.key_words + q];
        A.idx_pos[slot] = (uint32_t)p;
    }
}

struct QueryArgs {
    const uint8_t* seq;  // joined queries
    const uint64_t* qoff;  // [n_queries + 1] start of each query (query q ends at qoff[q+1] - 1, its '\n')
    uint32_t n_queries;
    uint32_t k;
    uint32_t kind;  // 0 seq-hash, 1 rabin-karp128
    uint32_t key_words;
    const uint64_t* idx_key;
    const uint32_t* idx_pos;
    const uint64_t* boff;
    uint32_t log_buckets;
    const uint32_t* out_colors;  // color_words per solid position (colored)
    uint32_t color_words;  // 0: uncolored
    uint32_t n_colors;  // per-query color counters written (<= 32 * color_words)
    uint64_t* matched;  // [n_queries]
    uint32_t* color_counts;  // [n_queries * n_colors]
    unsigned int* work;
    u128_t rk_mult, rk_mult_inv, rk_rmmult;
    u128_t rk_fwd[4], rk_bkw[4];
    u128_t rk_fwd_rm[4], rk_bkw_rm[4];  // rk_fwd[c] * rk_rmmult, rk_bkw[c] * rk_rmmult
};

constexpr uint32_t NO_POS = 0xFFFFFFFFu;

GG_DEV uint32_t index_lookup(const QueryArgs& A, uint64_t lo, uint64_t hi) {
    const uint32_t b = A.log_buckets ? (uint32_t)(index_hash(lo, hi) >> (64 - A.log_buckets)) : 0u;
    const uint64_t e0 = A.boff[b], e1 = A.boff[b + 1];
    for (uint64_t e = e0; e < e1; e++) {
        if (A.key_words == 1) {
            if (A.idx_key[e] == lo) return A.idx_pos[e];
        } else if (A.idx_key[2 * e] == lo && A.idx_key[2 * e + 1] == hi) {
            return A.idx_pos[e];
        }
    }
    return NO_POS;
}

GG_DEV bool colors_equal(const QueryArgs& A, uint32_t p0, uint32_t p1) {
    if (p0 == p1) return true;
    for (uint32_t w = 0; w < A.color_words; w++)
        if (A.out_colors[(uint64_t)p0 * A.color_words + w] != A.out_colors[(uint64_t)p1 * A.color_words + w])
            return false;
    return true;
}

GG_DEV void flush_run(const QueryArgs& A, uint32_t pos, uint32_t cnt, uint32_t* counts) {
    if (pos == NO_POS || cnt == 0) return;
    for (uint32_t w = 0; w < A.color_words; w++) {
        uint32_t bits = A.out_colors[(uint64_t)pos * A.color_words + w];
        while (bits) {
            const uint32_t c = w * 32 + (__ffs(bits) - 1);
            bits &= bits - 1;
            if (c < A.n_colors) atomicAdd(&counts[c], cnt);
        }
    }
}

GG_DEV uint32_t ascii_code(uint8_t c) { return (c >> 1) & 3u; }
GG_DEV bool ascii_acgt(uint8_t c) {
    const uint8_t u = c & 0xDF;
    return u == 'A' || u == 'C' || u == 'G' || u == 'T';
}

// One workgroup per query (work-stealing over the batch).  Each thread takes
// a contiguous range of k-mer start positions, rolls the canonical hash over
// it and looks every all-ACGT k-mer up; runs of hits with the same color set
// are added to the per-color LDS counters once per run.
__global__ __launch_bounds__(256) void query_kernel(QueryArgs A) {
    extern __shared__ __align__(16) uint8_t lds[];
    uint32_t* counts = reinterpret_cast<uint32_t*>(lds);
    __shared__ uint32_t s_q;
    __shared__ unsigned long long s_matched;
    const uint32_t k = A.k;
    while (true) {
        if (threadIdx.x == 0) {
            s_q = atomicAdd(A.work, 1u);
            s_matched = 0;
        }
        __syncthreads();
        const uint32_t q = s_q;
        if (q >= A.n_queries) break;
        for (uint32_t c = threadIdx.x; c < A.n_colors; c += blockDim.x) counts[c] = 0;
        __syncthreads();
        const uint8_t* s = A.seq + A.qoff[q];
        const uint64_t len = A.qoff[q + 1] - A.qoff[q] - 1;
        const uint64_t nk = len >= k ? len - k + 1 : 0;
        const uint64_t per = (nk + blockDim.x - 1) / blockDim.x;
        const uint64_t j0 = per * threadIdx.x, j1 = j0 + per < nk ? j0 + per : nk;
        uint64_t hits = 0;
        uint32_t run_pos = NO_POS, run_cnt = 0;
        if (j0 < j1) {
            int64_t last_bad = (int64_t)j0 - 1;  // last non-ACGT base seen
            auto on_key = [&](uint64_t lo, uint64_t hi) {
                const uint32_t pos = index_lookup(A, lo, hi);
                if (pos == NO_POS) return;
                hits++;
                if (A.color_words) {
                    if (run_pos != NO_POS && colors_equal(A, run_pos, pos)) {
                        run_cnt++;
                    } else {
                        flush_run(A, run_pos, run_cnt, counts);
                        run_pos = pos;
                        run_cnt = 1;
                    }
                }
            };
            if (A.kind == 0 && k <= 32) {
                const uint64_t mask = (k == 32) ? ~0ULL : ((1ULL << (2 * k)) - 1);
                uint64_t fw = 0, rc = 0;
                for (uint64_t i = j0; i < j1 + k - 1; i++) {
                    const uint8_t ch = s[i];
                    if (!ascii_acgt(ch)) last_bad = (int64_t)i;
                    const uint64_t c = ascii_code(ch);
                    fw = (fw >> 2) | (c << (2 * (k - 1)));
                    rc = ((rc << 2) | (c ^ 2)) & mask;
                    if (i + 1 >= j0 + k && last_bad < (int64_t)(i + 1 - k)) on_key(fw < rc ? fw : rc, 0);
                }
            } else if (A.kind == 0) {
                const u128_t mask = (k == 64) ? ~(u128_t)0 : (((u128_t)1 << (2 * k)) - 1);
                u128_t fw = 0, rc = 0;
                for (uint64_t i = j0; i < j1 + k - 1; i++) {
                    const uint8_t ch = s[i];
                    if (!ascii_acgt(ch)) last_bad = (int64_t)i;
                    const u128_t c = ascii_code(ch);
                    fw = (fw >> 2) | (c << (2 * (k - 1)));
                    rc = ((rc << 2) | (c ^ 2)) & mask;
                    if (i + 1 >= j0 + k && last_bad < (int64_t)(i + 1 - k)) {
                        const u128_t key = fw < rc ? fw : rc;
                        on_key((uint64_t)key, (uint64_t)(key >> 64));
                    }
                }
            } else {
                // rabin-karp128 rolling (crates/hashes/src/base/cn_rkhash_base.rs:65-110)
                u128_t fh = 0, bk = 0;
                for (uint64_t i = j0; i + 1 < j0 + k; i++) {
                    if (!ascii_acgt(s[i])) last_bad = (int64_t)i;
                    fh = fh * A.rk_mult + A.rk_fwd[ascii_code(s[i])];
                }
                for (int64_t i = (int64_t)(j0 + k) - 2; i >= (int64_t)j0; i--) bk = bk * A.rk_mult + A.rk_bkw[ascii_code(s[i])];
                bk = bk * A.rk_mult;
                for (uint64_t idx = j0 + k - 1; idx < j1 + k - 1; idx++) {
                    if (!ascii_acgt(s[idx])) last_bad = (int64_t)idx;
                    const uint32_t in = ascii_code(s[idx]), out = ascii_code(s[idx - (k - 1)]);
                    const u128_t cur_f = fh * A.rk_mult + A.rk_fwd[in];
                    fh = cur_f - A.rk_fwd_rm[out];
                    const u128_t cur_b = bk * A.rk_mult_inv + A.rk_bkw_rm[in];
                    bk = cur_b - A.rk_bkw[out];
                    if (last_bad < (int64_t)(idx + 1 - k)) {
                        const u128_t key = cur_f < cur_b ? cur_f : cur_b;
                        on_key((uint64_t)key, (uint64_t)(key >> 64));
                    }
                }
            }
            if (A.color_words) flush_run(A, run_pos, run_cnt, counts);
        }
        if (hits) atomicAdd(&s_matched, (unsigned long long)hits);
        __syncthreads();
        if (threadIdx.x == 0) A.matched[q] = s_matched;
        for (uint32_t c = threadIdx.x; c < A.n_colors; c += blockDim.x)
            A.color_counts[(uint64_t)q * A.n_colors + c] = counts[c];
        __syncthreads();
    }
}

}  // namespace gg
