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
// Table entry value: bits 0..29 counter, bits 30..31 the two MapEntry flags,
// so one u32 atomicAdd + one atomicOr update an entry.
#include "common.h"

namespace gg {

constexpr uint64_t EMPTY64 = ~0ULL;

struct MergeArgs {
    const uint32_t* slots;
    uint32_t slot_words;
    const uint32_t* pool;
    const uint64_t* sub_off;
    const uint32_t* sub_kmers;
    uint32_t n_sub;
    uint32_t k;
    uint32_t min_mult;
    uint32_t kind;  // 0 = canonical seq-hash, 1 = canonical rabin-karp128
    uint32_t key_words;  // 1 (u64 keys) or 2 (u128 keys)
    uint32_t lds_cap;  // max table entries in LDS (power of two)
    uint32_t big_thresh;  // occurrences above which the global path is used
    unsigned int* work;
    uint32_t* big_list;
    unsigned int* big_count;
    uint64_t* out_key;  // key_words per entry
    uint32_t* out_val;  // multiplicity | flags << 30
    uint32_t* out_sub;
    unsigned long long* out_cursor;
    uint64_t out_cap;
    // global-memory tables for oversize sub-buckets
    uint64_t* g_klo;
    uint64_t* g_khi;
    uint32_t* g_st;
    uint32_t* g_val;
    const uint64_t* g_off;  // per big-list entry: table offset
    const uint32_t* g_cap;  // per big-list entry: table capacity (power of two)
    // rabin-karp128 constants (crates/hashes/src/cn_rkhash.rs:65-78)
    u128_t rk_mult, rk_mult_inv, rk_rmmult;
    u128_t rk_fwd[4], rk_bkw[4];
};

GG_DEV uint32_t hash128(uint64_t lo, uint64_t hi) { return tab_hash64(lo ^ (hi * 0x9E3779B97F4A7C15ULL)); }

template <class KP, class VP>
GG_DEV void tab_insert64(KP keys, VP vals, uint32_t mask, uint64_t key, uint32_t add, uint32_t fl) {
    uint32_t i = tab_hash64(key) & mask;
    while (true) {
        uint64_t cur = keys[i];
        if (cur == key) break;
        if (cur == EMPTY64) {
            uint64_t old = atomicCAS((unsigned long long*)&keys[i], (unsigned long long)EMPTY64,
                                     (unsigned long long)key);
            if (old == EMPTY64 || old == key) break;
        }
        i = (i + 1) & mask;
    }
    atomicAdd((uint32_t*)&vals[i], add);
    if (fl) atomicOr((uint32_t*)&vals[i], fl << 30);
}

// 128-bit keys: a slot is claimed through its state word (0 free, 1 being
// written, 2 ready).  The loop body runs the writer and the waiters in the
// same iteration, so lanes of one wavefront cannot wait on each other.
template <class KP, class SP, class VP>
GG_DEV void tab_insert128(KP klo, KP khi, SP st, VP vals, uint32_t mask, uint64_t lo, uint64_t hi, uint32_t add,
                          uint32_t fl) {
    uint32_t i = hash128(lo, hi) & mask;
    bool done = false;
    while (!done) {
        uint32_t s = __hip_atomic_load((uint32_t*)&st[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (s == 0) {
            if (atomicCAS((uint32_t*)&st[i], 0u, 1u) == 0u) {
                klo[i] = lo;
                khi[i] = hi;
                __threadfence();
                atomicExch((uint32_t*)&st[i], 2u);
                done = true;
            }
        } else if (s == 2) {
            uint64_t a = __hip_atomic_load((uint64_t*)&klo[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            uint64_t b = __hip_atomic_load((uint64_t*)&khi[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (a == lo && b == hi) done = true;
            else i = (i + 1) & mask;
        }
    }
    atomicAdd((uint32_t*)&vals[i], add);
    if (fl) atomicOr((uint32_t*)&vals[i], fl << 30);
}

// Iterate the k-mers of one slot; cb(lo, hi, is_forward, idx, last_idx).
template <class CB>
GG_DEV void for_each_kmer(const MergeArgs& A, const uint32_t* slot, CB&& cb) {
    const uint32_t meta = slot[0];
    const uint32_t* bw;
    uint32_t len;
    if (meta & META_LONG) {
        len = slot[1];
        uint64_t off = (uint64_t)slot[2] | ((uint64_t)slot[3] << 32);
        bw = A.pool + off;
    } else {
        len = meta_len(meta);
        bw = slot + 1;
    }
    const uint32_t k = A.k;
    if (len < k) return;
    const uint32_t last = len - k;
    if (A.kind == 0 && k <= 32) {
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
    } else if (A.kind == 0) {
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
        // rabin-karp128 rolling (cn_rkhash_base.rs:65-110)
        u128_t fh = 0, bk = 0;
        auto code = [&](uint32_t i) -> uint32_t { return (bw[i >> 4] >> (2 * (i & 15))) & 3; };
        for (uint32_t i = 0; i + 1 < k; i++) fh = fh * A.rk_mult + A.rk_fwd[code(i)];
        for (int i = (int)k - 2; i >= 0; i--) bk = bk * A.rk_mult + A.rk_bkw[code(i)];
        bk = bk * A.rk_mult;
        for (uint32_t idx = k - 1; idx < len; idx++) {
            uint32_t in = code(idx), out = code(idx - (k - 1));
            u128_t cur_f = fh * A.rk_mult + A.rk_fwd[in];
            fh = cur_f - A.rk_fwd[out] * A.rk_rmmult;
            u128_t cur_b = bk * A.rk_mult_inv + A.rk_bkw[in] * A.rk_rmmult;
            bk = cur_b - A.rk_bkw[out];
            u128_t key = cur_f < cur_b ? cur_f : cur_b;
            cb((uint64_t)key, (uint64_t)(key >> 64), cur_f < cur_b, idx + 1 - k, last);
        }
    }
}

GG_DEV uint32_t occurrence_flags(uint32_t rflags, bool is_fwd, uint32_t idx, uint32_t last) {
    bool begin_ignored = (rflags & 1) == 0 && idx == 0;
    bool end_ignored = (rflags & 2) == 0 && idx == last;
    return ((begin_ignored ? 1u : 0u) << (is_fwd ? 0 : 1)) | ((end_ignored ? 1u : 0u) << (is_fwd ? 1 : 0));
}

// Emit solid entries of table positions [i0, i1) (stride 256 per thread).
template <class KP, class SP, class VP>
GG_DEV void emit_solid(const MergeArgs& A, uint32_t sub, KP klo, KP khi, SP st, VP vals, uint32_t cap) {
    const uint32_t lane = threadIdx.x & 63;
    for (uint32_t i0 = 0; i0 < cap; i0 += blockDim.x) {
        uint32_t i = i0 + threadIdx.x;
        bool solid = false;
        uint32_t outv = 0;
        if (i < cap) {
            bool occupied = (A.key_words == 1) ? (klo[i] != EMPTY64) : (st[i] == 2);
            if (occupied) {
                uint32_t v = vals[i];
                uint32_t cnt = v & 0x3FFFFFFFu, fl = v >> 30;
                uint32_t mult = cnt >> (fl == 3 ? 1 : 0);
                solid = mult >= A.min_mult;
                outv = mult | (fl << 30);
            }
        }
        uint64_t bal = __ballot(solid);
        if (bal == 0) continue;
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(A.out_cursor, (unsigned long long)__popcll(bal));
        base = __shfl(base, 0);
        if (solid) {
            uint64_t pos = base + __popcll(bal & ((1ULL << lane) - 1));
            if (pos < A.out_cap) {
                if (A.key_words == 1) {
                    A.out_key[pos] = klo[i];
                } else {
                    A.out_key[2 * pos] = klo[i];
                    A.out_key[2 * pos + 1] = khi[i];
                }
                A.out_val[pos] = outv;
                A.out_sub[pos] = sub;
            }
        }
    }
}

template <class KP, class SP, class VP>
GG_DEV void insert_records(const MergeArgs& A, uint32_t sub, KP klo, KP khi, SP st, VP vals, uint32_t mask) {
    const uint64_t r0 = A.sub_off[sub], r1 = A.sub_off[sub + 1];
    const uint32_t SW = A.slot_words;
    for (uint64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
        const uint32_t* slot = A.slots + r * SW;
        const uint32_t rflags = meta_flags(slot[0]);
        for_each_kmer(A, slot, [&](uint64_t lo, uint64_t hi, bool is_fwd, uint32_t idx, uint32_t last) {
            uint32_t fl = occurrence_flags(rflags, is_fwd, idx, last);
            if (A.key_words == 1) tab_insert64(klo, vals, mask, lo, 1u, fl);
            else tab_insert128(klo, khi, st, vals, mask, lo, hi, 1u, fl);
        });
    }
}

__global__ __launch_bounds__(256) void merge_lds_kernel(MergeArgs A) {
    extern __shared__ __align__(16) uint8_t lds[];
    __shared__ uint32_t s_item;
    const uint32_t cap_max = A.lds_cap;
    uint64_t* klo = reinterpret_cast<uint64_t*>(lds);
    uint64_t* khi = klo + cap_max;  // only used for 128-bit keys
    uint32_t* vals = reinterpret_cast<uint32_t*>(lds + (size_t)cap_max * 8 * A.key_words);
    uint32_t* st = vals + cap_max;
    while (true) {
        if (threadIdx.x == 0) s_item = atomicAdd(A.work, 1u);
        __syncthreads();
        const uint32_t sub = s_item;
        __syncthreads();
        if (sub >= A.n_sub) break;
        const uint32_t occ = A.sub_kmers[sub];
        if (occ == 0) continue;
        if (occ > A.big_thresh) {
            if (threadIdx.x == 0) A.big_list[atomicAdd(A.big_count, 1u)] = sub;
            continue;
        }
        uint32_t cap = 256;
        while (cap < 2 * occ && cap < cap_max) cap <<= 1;
        const uint32_t mask = cap - 1;
        for (uint32_t i = threadIdx.x; i < cap; i += blockDim.x) {
            klo[i] = EMPTY64;
            vals[i] = 0;
            if (A.key_words == 2) st[i] = 0;
        }
        __syncthreads();
        insert_records(A, sub, klo, khi, st, vals, mask);
        __syncthreads();
        emit_solid(A, sub, klo, khi, st, vals, cap);
        __syncthreads();
    }
}

// Oversize sub-buckets: same algorithm with the table in HBM (one workgroup
// per sub-bucket, tables pre-cleared by the host).
__global__ __launch_bounds__(256) void merge_global_kernel(MergeArgs A, uint32_t n_big) {
    for (uint32_t e = blockIdx.x; e < n_big; e += gridDim.x) {
        const uint32_t sub = A.big_list[e];
        const uint64_t off = A.g_off[e];
        const uint32_t cap = A.g_cap[e];
        uint64_t* klo = A.g_klo + off;
        uint64_t* khi = A.g_khi ? A.g_khi + off : nullptr;
        uint32_t* st = A.g_st ? A.g_st + off : nullptr;
        uint32_t* vals = A.g_val + off;
        insert_records(A, sub, klo, khi, st, vals, cap - 1);
        __syncthreads();
        __threadfence();
        emit_solid(A, sub, klo, khi, st, vals, cap);
        __syncthreads();
    }
}

__global__ void fill_u64_kernel(uint64_t* p, uint64_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

}  // namespace gg
