// unitigs.hip -- in/out extension flags and intra-sub-bucket unitig chains
// over the solid seq-hash table of the last merge.
//
// Restates the extension of crates/assembler_kmers_merge/src/unitigs_extender/
// hashmap.rs:162-297,442-601 as an order-free rule (the reference's greedy
// loop gives the same chains except where a k-mer's flagged side has a
// neighbour in its own sub-bucket, where its own result depends on the order
// in which it visits k-mers):
//   * in a k-mer's canonical orientation, side 0 is the predecessor side and
//     side 1 the successor side; its degree on a side is the number of solid
//     k-mers of the same sub-bucket one base away (manual_roll_forward /
//     manual_roll_reverse, crates/hashes/src/base/cn_seqhash_base.rs:165-200);
//   * a k-mer whose flags are exactly INCL_BEGIN (or INCL_END) has its side 0
//     (or side 1) flagged: its continuation is in another bucket
//     (contig_break, hashmap.rs:287-292; begin/end_ignored, :483-494);
//   * sides (p,s) and (q,t) are linked when each is the other's only
//     neighbour, p != q, and neither side is flagged;
//   * a unitig is a maximal linked path, written from its end with the
//     smaller (key, side); an end is open when its outer side is flagged.
// Cycles (paths without ends) are rare and are walked on the host.
//
// Joined mode (global != 0) builds the maximal unitigs of the whole solid
// set, i.e. the partial unitigs of every sub-bucket joined across buckets
// (what the links / reorganize / build_final_unitigs stages of
// crates/assembler_pipeline/src/extend_unitigs.rs:348-993 produce): a k-mer
// held by several sub-buckets (a boundary k-mer, counted on both sides) is
// one node, represented by its smallest table position; neighbours are
// looked up over the whole table and no side is flagged.
#include "common.h"

namespace gg {

constexpr uint32_t NO_LINK = 0xFFFFFFFFu;

struct UnitigArgs {
    const uint64_t* out_key;
    const uint32_t* out_val;
    const uint32_t* out_sub;
    const uint32_t* chunk_fill;
    uint32_t n_chunks;
    uint64_t n_tail;
    uint32_t chunk;
    uint32_t key_words;
    uint32_t k;
    uint64_t* tab;  // [tab_mask + 1] open addressing: hash bits 32-63 << 32 | (position + 1); 0 = empty
    uint64_t tab_mask;
    uint32_t* nb;  // [2 n_pos] only neighbour (q << 1 | t) or NO_LINK
    uint8_t* ext;  // [n_pos] degree side 0 (0, 1, 2 = more) | degree side 1 << 2
    uint32_t* link;  // [2 n_pos]
    uint32_t* head_bytes;  // [2 n_pos] bytes of the unitig written from this end (0 if not its start)
    uint32_t* head_count;  // [2 n_pos] 1 when the unitig starts here
    const uint64_t* head_off;  // [2 n_pos + 1] exclusive scan of head_bytes
    const uint64_t* head_idx;  // [2 n_pos + 1] exclusive scan of head_count
    uint8_t* seq;
    uint32_t* uni_sub;
    uint8_t* uni_flags;
    uint8_t* visited;  // [n_pos]
    uint32_t global;  // joined mode: one graph over all sub-buckets
    uint8_t* rep;  // [n_pos] joined mode: 1 where p is its key's smallest position
    unsigned int* n_nodes;  // joined mode: number of representatives
    uint32_t mark_visited;  // write pass: mark path nodes (needed only when cycles exist)
    uint32_t* path_pos;  // write pass, optional: the table position of every k-mer of every unitig, in
                         // written order (unitig i's k-mers start at head_off - i * k)
};

// Position p holds a solid entry (chunks are filled from their start).
GG_DEV bool slot_filled(const UnitigArgs& A, uint64_t p) {
    const uint64_t c = p / A.chunk;
    if (c < A.n_chunks) return (p - c * A.chunk) < A.chunk_fill[c];
    return p - (uint64_t)A.n_chunks * A.chunk < A.n_tail;
}

// Position p is a node of the graph being compacted.
GG_DEV bool unitig_valid(const UnitigArgs& A, uint64_t p) {
    return slot_filled(A, p) && (!A.global || A.rep[p]);
}

GG_DEV u128_t key_at(const UnitigArgs& A, uint64_t p) {
    const uint64_t lo = A.out_key[p * A.key_words];
    const uint64_t hi = A.key_words == 2 ? A.out_key[p * 2 + 1] : 0;
    return ((u128_t)hi << 64) | lo;
}

// Reverse complement of a k-mer (base j at bits 2j, A0 C1 T2 G3, so the
// complement of a base is code ^ 2): complement all 64 positions, reverse
// the 128 bits, swap the bits inside each 2-bit group back, and drop the
// 64 - k positions above the k-mer, which the reversal put at the bottom.
GG_DEV u128_t kmer_rc128(u128_t x, uint32_t k) {
    constexpr uint64_t AA = 0xAAAAAAAAAAAAAAAAULL, M5 = 0x5555555555555555ULL;
    uint64_t lo = __brevll((uint64_t)(x >> 64) ^ AA), hi = __brevll((uint64_t)x ^ AA);
    lo = ((lo >> 1) & M5) | ((lo & M5) << 1);
    hi = ((hi >> 1) & M5) | ((hi & M5) << 1);
    const u128_t r = ((u128_t)hi << 64) | lo;
    return r >> (128 - 2 * k);
}

GG_DEV bool side_flagged(const UnitigArgs& A, uint64_t p, uint32_t s) {
    if (A.global) return false;
    const uint32_t fl = A.out_val[p] >> 30;
    return (fl == 1 && s == 0) || (fl == 2 && s == 1);
}

// Every filled position goes into the lookup table (linear probing; the
// table has at least twice as many slots as positions, so probing ends).
__global__ void utab_insert_kernel(UnitigArgs A, uint64_t n_pos) {
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < n_pos;
         p += (uint64_t)gridDim.x * blockDim.x) {
        if (!slot_filled(A, p)) continue;
        const u128_t c = key_at(A, p);
        const uint64_t h = index_hash((uint64_t)c, (uint64_t)(c >> 64));
        const uint64_t v = (h & 0xFFFFFFFF00000000ULL) | (p + 1);
        for (uint64_t i = h & A.tab_mask;; i = (i + 1) & A.tab_mask)
            if (atomicCAS(reinterpret_cast<unsigned long long*>(A.tab + i), 0ULL, (unsigned long long)v) == 0ULL) break;
    }
}

// Position of key `c` inside sub-bucket `sub`, or NO_LINK.  In joined mode
// the sub-bucket is ignored and the key's smallest position is returned.
GG_DEV uint32_t lookup_in_sub(const UnitigArgs& A, u128_t c, uint32_t sub) {
    const uint64_t h = index_hash((uint64_t)c, (uint64_t)(c >> 64));
    const uint32_t fp = (uint32_t)(h >> 32);
    uint32_t best = NO_LINK;
    for (uint64_t i = h & A.tab_mask;; i = (i + 1) & A.tab_mask) {
        const uint64_t v = A.tab[i];
        if (v == 0) break;
        if ((uint32_t)(v >> 32) != fp) continue;
        const uint32_t q = (uint32_t)v - 1u;
        if (key_at(A, q) != c) continue;
        if (!A.global) {
            if (A.out_sub[q] == sub) return q;
        } else if (q < best) {
            best = q;
        }
    }
    return best;
}

// Joined mode: marks each filled position that is its key's smallest one.
// Counts them into *n_nodes (blocks of 256 threads: one atomic per block).
__global__ __launch_bounds__(256) void rep_kernel(UnitigArgs A, uint64_t n_pos) {
    __shared__ unsigned int part[4];
    unsigned int cnt = 0;
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < n_pos; p += (uint64_t)gridDim.x * blockDim.x) {
        const bool r = slot_filled(A, p) && lookup_in_sub(A, key_at(A, p), 0) == (uint32_t)p;
        A.rep[p] = r;
        cnt += r ? 1u : 0u;
    }
    for (uint32_t d = 32; d; d >>= 1) cnt += __shfl_down(cnt, d);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned int t = part[0] + part[1] + part[2] + part[3];
        if (t) atomicAdd(A.n_nodes, t);
    }
}

__global__ void ext_kernel(UnitigArgs A, uint64_t n_pos) {
    const uint32_t k = A.k;
    const u128_t mask = (k >= 64) ? ~(u128_t)0 : (((u128_t)1 << (2 * k)) - 1);
    for (uint64_t p = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; p < n_pos; p += (uint64_t)gridDim.x * blockDim.x) {
        if (!unitig_valid(A, p)) {
            A.nb[2 * p] = A.nb[2 * p + 1] = NO_LINK;
            A.ext[p] = 0;
            continue;
        }
        const u128_t x = key_at(A, p);
        const u128_t xr = kmer_rc128(x, k);
        const uint32_t sub = A.out_sub[p];
        uint32_t degs = 0;
        for (uint32_t s = 0; s < 2; s++) {
            uint32_t cnt = 0, only = NO_LINK;
            for (uint32_t b = 0; b < 4; b++) {
                // neighbour y and its reverse complement, both from x and rc(x)
                const u128_t y = s ? ((x >> 2) | ((u128_t)b << (2 * (k - 1)))) : (((x << 2) | b) & mask);
                const u128_t yr = s ? (((xr << 2) | (b ^ 2u)) & mask) : ((xr >> 2) | ((u128_t)(b ^ 2u) << (2 * (k - 1))));
                const u128_t c = y < yr ? y : yr;
                const uint32_t q = lookup_in_sub(A, c, sub);
                if (q == NO_LINK) continue;
                cnt++;
                const uint32_t t = (y == c) ? (s ? 0u : 1u) : (s ? 1u : 0u);
                only = (q << 1) | t;
            }
            A.nb[2 * p + s] = cnt == 1 ? only : NO_LINK;
            degs |= (cnt > 2 ? 2u : cnt) << (2 * s);
        }
        A.ext[p] = (uint8_t)degs;
    }
}

__global__ void link_kernel(UnitigArgs A, uint64_t n_pos) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < 2 * n_pos;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t v = A.nb[e];
        uint32_t out = NO_LINK;
        if (v != NO_LINK) {
            const uint64_t p = e >> 1, q = v >> 1;
            const uint32_t s = (uint32_t)(e & 1), t = v & 1;
            if (q != p && !side_flagged(A, p, s) && !side_flagged(A, q, t) && A.nb[v] == (uint32_t)e) out = v;
        }
        A.link[e] = out;
    }
}

// Last base of k-mer q read in orientation t (0: canonical, 1: reverse
// complement): base k-1 of the key, or the complement of its base 0.
GG_DEV uint32_t entry_base(const UnitigArgs& A, uint64_t q, uint32_t t) {
    if (t) return ((uint32_t)A.out_key[q * A.key_words] & 3u) ^ 2u;
    const uint32_t j = A.k - 1;
    const uint64_t w = A.out_key[q * A.key_words + (j >> 5)];
    return (uint32_t)(w >> (2 * (j & 31))) & 3u;
}

// Writes the bytes of one unitig at seq[off...] through 4-byte stores; the
// first and last words, which it may share with its neighbours, byte by byte.
struct SeqWriter {
    uint8_t* seq;
    uint64_t off, a;
    uint32_t acc = 0;
    GG_DEV SeqWriter(uint8_t* s, uint64_t o) : seq(s), off(o), a(o) {}
    GG_DEV void put(uint32_t ch) {
        acc |= ch << (8 * (uint32_t)(a & 3));
        if ((a & 3) == 3) {
            if (a - 3 >= off) {
                *reinterpret_cast<uint32_t*>(seq + a - 3) = acc;
            } else {
                for (uint64_t b = off; b <= a; b++) seq[b] = (uint8_t)(acc >> (8 * (uint32_t)(b & 3)));
            }
            acc = 0;
        }
        a++;
    }
    GG_DEV void finish() {
        if ((a & 3) == 0) return;  // the last word was stored
        const uint64_t w0 = a & ~(uint64_t)3;
        for (uint64_t b = w0 > off ? w0 : off; b < a; b++) seq[b] = (uint8_t)(acc >> (8 * (uint32_t)(b & 3)));
    }
};

// Pass 1 (write == false): every end of a path walks to its other end; the
// end with the smaller (key, side) records the unitig's bytes.  Pass 2
// (write == true): those ends write the sequence, its sub-bucket and flags.
template <bool WRITE>
__global__ void walk_kernel(UnitigArgs A, uint64_t n_pos) {
    const uint32_t k = A.k;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < 2 * n_pos;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t p = e >> 1;
        const uint32_t s = (uint32_t)(e & 1);
        bool head = unitig_valid(A, p) && A.link[e] == NO_LINK;
        if (!WRITE) {
            uint32_t bytes = 0, cnt = 0;
            if (head) {
                uint64_t cur = e;
                uint64_t steps = 1;
                while (A.link[cur ^ 1] != NO_LINK && steps <= n_pos) {
                    cur = A.link[cur ^ 1];
                    steps++;
                }
                const uint64_t q = cur >> 1;
                const uint32_t qs = (uint32_t)((cur ^ 1) & 1);
                const u128_t kp = key_at(A, p), kq = key_at(A, q);
                const bool canon = kp < kq || (kp == kq && s < qs);
                if (canon) {
                    bytes = (uint32_t)(steps + k);  // k-mer, one base per further k-mer, '\n'
                    cnt = 1;
                }
            }
            A.head_bytes[e] = bytes;
            A.head_count[e] = cnt;
            continue;
        }
        if (!head || A.head_off[e + 1] == A.head_off[e]) continue;
        SeqWriter out(A.seq, A.head_off[e]);
        const uint64_t ui = A.head_idx[e];
        constexpr uint32_t ASCII = 'A' | ('C' << 8) | ('T' << 16) | ('G' << 24);
        const u128_t x = s == 0 ? key_at(A, p) : kmer_rc128(key_at(A, p), k);
        for (uint32_t i = 0; i < k; i++) out.put((ASCII >> (8 * ((uint32_t)(x >> (2 * i)) & 3u))) & 0xFFu);
        if (A.mark_visited) A.visited[p] = 1;
        uint32_t* pp = A.path_pos ? A.path_pos + (A.head_off[e] - ui * k) : nullptr;
        if (pp) *pp++ = (uint32_t)p;
        uint64_t cur = e;
        // the count pass found this path's end within n_pos steps
        for (uint64_t steps = 1; A.link[cur ^ 1] != NO_LINK && steps <= n_pos; steps++) {
            cur = A.link[cur ^ 1];
            const uint64_t q = cur >> 1;
            if (A.mark_visited) A.visited[q] = 1;
            if (pp) *pp++ = (uint32_t)q;
            out.put((ASCII >> (8 * entry_base(A, q, (uint32_t)(cur & 1)))) & 0xFFu);
        }
        out.put('\n');
        out.finish();
        const uint64_t qend = cur >> 1;
        const uint32_t exit_side = (uint32_t)((cur ^ 1) & 1);
        A.uni_sub[ui] = A.global ? 0u : A.out_sub[p];
        A.uni_flags[ui] = (uint8_t)((side_flagged(A, p, s) ? 1u : 0u) | (side_flagged(A, qend, exit_side) ? 2u : 0u));
    }
}

}  // namespace gg
