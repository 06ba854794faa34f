// bucketing.hip -- minimizer bucketing on gfx950.
//
// Restates, for the device, the per-run work of
//   crates/assembler_minimizer_bucketing/src/lib.rs:171-270  (process_sequence)
//   crates/hashes/src/cn_nthash.rs:19-109                     (canonical ntHash m-mers)
//   crates/hashes/src/rolling/batch_minqueue.rs:37-189        (window minima + splits)
//   crates/minimizer_bucketing/src/sequences_splitter.rs:15-40 (runs of >= k ACGT)
//
// Input layout: one byte array in HBM holding every record followed by a
// separator byte.  Any byte that is not A/C/G/T (either case) ends an ACGT run,
// exactly like 'N' after the reference's normalisation, so separators need no
// index.
//
// Geometry.  A "window" j is the (k-1)-mer starting at base j; its minimum is
// taken over the w = k - m m-mers j .. j+w-1.  A tile owns windows [t0, t0+T)
// and is processed by one wavefront.  The wave also computes the w windows
// before t0 (halo), enough to find the start of every group whose minimizer is
// unique (such groups span at most w windows).  Groups with a non-unique
// minimizer may be longer; their start is found in the global group-start
// bitmap once pass 1 has written all of it.
//
//   pass 1 (bucket_stage_kernel)  : group-start bitmap; every record built,
//                                   counted in its unit (the count is its rank)
//                                   and written to its unit's region at that
//                                   rank, or to the spill list past the region
//          (bucket_defer_kernel)  : records whose group starts before the halo
//   scan   (scan_*_kernel)        : unit offsets of the dense arena
//   pass 2 (unit_compact_kernel)  : each unit's region copied to its offsets
//          (spill_place_kernel)   : spilled records to unit offset + rank
//
// A group = maximal run of consecutive windows sharing the minimizer (value,
// and position when unique).  Its super-k-mer is the set of k-mers touching
// the group: bases [a-1, b+k) for a group of windows [a, b], clipped to the
// ACGT run (include_first / include_last when clipped).
#include "common.h"

namespace gg {

// Records are scattered by physical unit: the sub-bucket (the semantic unit,
// one kmers_transform job) refined by `xb` further bits of the minimizer hash
// for unique minimizers.  Units of one sub-bucket are contiguous, so the merge
// can process any run of them as one work item.
struct BucketArgs {
    const uint8_t* seq;
    uint64_t n;  // bytes in seq
    uint32_t k, m, w;
    uint32_t b, sb, xb;
    uint32_t canonical;
    uint32_t T;  // windows per tile (multiple of 64)
    uint64_t n_tiles;
    uint64_t tile0;  // index of the launch's first tile (launches over part of the tiles)
    uint32_t* seg_bits;  // [ceil(n/32)+2] group-start bitmap (written by pass 1)
    uint32_t* sub_count;  // [n_units] records per unit
    unsigned long long* unit_ck;  // [n_units] records << 32 | k-mer occurrences, counted in pass 1
    uint64_t n_slots;  // slots in the arena
    // Pass 1 writes record `rank` of unit u to its region slot
    // u * unit_cap + rank; records of rank >= unit_cap are appended to the
    // spill list with their unit and rank.
    uint32_t unit_cap;
    uint32_t* region;  // [n_units * unit_cap * slot_words]
    uint32_t* region_color;  // [n_units * unit_cap] (colored builds)
    uint32_t* spill_slots;  // [spill_cap * slot_words]
    uint32_t* spill_unit;  // [spill_cap]
    uint32_t* spill_rank;  // [spill_cap]
    uint32_t* spill_color;  // [spill_cap] (colored builds)
    unsigned long long* spill_cursor;  // spilled records (may exceed spill_cap: the host retries)
    uint64_t spill_cap;
    uint64_t* defer;  // [n_tiles * 3] records whose group starts before the tile: end window, minimizer, last flag
    unsigned int* defer_count;
    unsigned long long* kmer_total;  // k-mer occurrences of all records
    const uint64_t* sub_off;  // [n_units+1] slot offsets
    uint32_t* slots;
    uint32_t slot_words;
    uint32_t* pool;
    unsigned long long* pool_cursor;
    uint64_t pool_cap;  // words
    uint32_t* sub_kmers;  // [n_units] k-mer occurrences
    // after deduplication: records stored per unit (distinct region records
    // plus spilled ones), of which distinct region records
    uint32_t* unit_dcount;
    uint32_t* unit_dregion;
    // k-mer positions per unit once identical records are stored once (the
    // inserts the merge makes; sizes its work items and tables)
    uint32_t* unit_skmers;
    const uint64_t* color_starts;
    const uint32_t* colors;
    uint32_t n_colors;
    uint32_t* slot_color;
    uint32_t* error_flag;  // set when a consistency check fails
    uint64_t gtab[40];  // interleaved (fw, rc) rolling tables, index out*4+in
};

// LDS layout of one tile (host and device agree through this struct).
struct TileGeom {
    uint32_t NW, NWB, NM, NB;
    uint32_t off_mm, off_wv, off_wp, off_pk, off_inv, off_fwd, off_seg, off_ok, off_last, off_g;
    uint32_t bytes;
};

// With `regs` the window minima are formed in registers and the per-m-mer
// hash array is not needed.
__host__ __device__ inline TileGeom tile_geom(uint32_t T, uint32_t w, uint32_t k, uint32_t m, bool regs) {
    TileGeom g;
    g.NW = T + w + 1;
    g.NWB = (g.NW + w - 1) / w;
    g.NM = (g.NWB + 1) * w;
    // bases relative to borg: d0 in [1,16]; the furthest read is the word
    // after the one holding relative base d0 + max(NM+m-2, NW+k-2)
    uint32_t need = 48 + (g.NM + m > g.NW + k ? g.NM + m : g.NW + k);
    g.NB = (need + 31) / 32 * 32;
    uint32_t o = 0;
    g.off_g = o; o += 40 * 8;
    g.off_mm = o; o += regs ? 0 : g.NM * 8;
    if (regs) {
        // register path (w == 16): per window the low 32 bits of its minimum
        // (all the record needs: unique bit, bucket, second bucket and unit
        // bits) and the minimizer's offset in the window as a byte, per m-mer
        // its strand as a bit; each lane's 16 windows padded by one word so
        // that lanes' stores hit distinct LDS banks
        g.off_wv = o; o += (g.NWB * 17 + 1) * 4;
        o = (o + 7) / 8 * 8;
    } else {
        // wv / wp / fwd are padded after every 16 entries (Tile::WV, WP, FWD)
        // so that lanes writing 16-window blocks hit distinct LDS banks
        g.off_wv = o; o += (g.NWB * w + g.NWB * w / 16 + 1) * 8;
    }
    g.off_pk = o; o += g.NB / 16 * 4;
    g.off_inv = o; o += g.NB / 32 * 4;
    g.off_seg = o; o += (g.NW / 64 + 2) * 8;
    g.off_ok = o; o += (g.NW / 64 + 2) * 8;
    g.off_last = o; o += (g.NW / 64 + 2) * 8;
    if (regs) {
        g.off_wp = o; o += g.NWB * 20 + 4;
        g.off_fwd = o; o += (g.NWB + 2) * 2;
    } else {
        g.off_wp = o; o += (g.NWB * w + 2 * (g.NWB * w / 16) + 2) * 2;
        g.off_fwd = o; o += g.NM + 4 * (g.NM / 16) + 4;
    }
    g.bytes = (o + 15) / 16 * 16;
    return g;
}

GG_DEV bool is_acgt(uint8_t c) {
    uint8_t u = c & 0xDF;
    return u == 'A' || u == 'C' || u == 'G' || u == 'T';
}

struct Tile {
    const BucketArgs* A;
    TileGeom G;
    int64_t t0, wlo, borg;
    uint32_t d0;  // wlo - borg
    uint8_t* lds;
    uint64_t* mm;
    uint64_t* wv;
    uint16_t* wp;
    uint32_t* pk;
    uint32_t* inv;
    uint8_t* fwd;
    uint32_t* seg;
    uint32_t* ok;
    uint32_t* last;
    uint64_t* gt;
    bool regs;  // the w == 16 layout: wv32 / wp8 / fwd16
    uint32_t* wv32;
    uint8_t* wp8;
    uint16_t* fwd16;

    GG_DEV uint32_t code_rel(int64_t r) const { return (pk[r >> 4] >> ((r & 15) * 2)) & 3; }
    GG_DEV bool inv_rel(int64_t r) const { return (inv[r >> 5] >> (r & 31)) & 1; }
    // true when any base in [r, r+len) (relative) is not ACGT
    GG_DEV bool any_inv(int64_t r, uint32_t len) const {
        int64_t e = r + len;  // exclusive
        int64_t wi = r >> 5;
        int64_t we = (e - 1) >> 5;
        uint32_t lo_mask = ~0u << (r & 31);
        if (wi == we) {
            uint32_t hi_mask = ((e & 31) == 0) ? ~0u : ((1u << (e & 31)) - 1);
            return (inv[wi] & lo_mask & hi_mask) != 0;
        }
        if (inv[wi] & lo_mask) return true;
        for (int64_t q = wi + 1; q < we; q++)
            if (inv[q]) return true;
        uint32_t hi_mask = ((e & 31) == 0) ? ~0u : ((1u << (e & 31)) - 1);
        return (inv[we] & hi_mask) != 0;
    }
    GG_DEV bool bit(const uint32_t* a, uint32_t x) const { return (a[x >> 5] >> (x & 31)) & 1; }
    // window minimum, its m-mer position, and forward-strand flag per m-mer
    GG_DEV uint64_t& WV(uint32_t x) const { return wv[x + (x >> 4)]; }
    GG_DEV uint16_t& WP(uint32_t x) const { return wp[x + 2 * (x >> 4)]; }
    GG_DEV uint8_t& FWD(uint32_t x) const { return fwd[x + 4 * (x >> 4)]; }
    // either layout: window minimum (the w == 16 path keeps its low 32 bits),
    // its m-mer position, and the strand flag of m-mer p
    GG_DEV uint64_t wval(uint32_t x) const { return regs ? (uint64_t)wv32[x + (x >> 4)] : WV(x); }
    GG_DEV uint32_t wpos(uint32_t x) const { return regs ? x + wp8[x + 4 * (x >> 4)] : WP(x); }
    GG_DEV bool fwd_at(uint32_t p) const { return regs ? ((fwd16[p >> 4] >> (p & 15)) & 1u) != 0 : FWD(p) != 0; }
};

// 16 codes of relative positions [r, r+16) packed (position r at bits 0..1);
// positions below 0 read as zero.
GG_DEV uint32_t extract16(const Tile& t, int64_t r) {
    if (r >= 0) {
        uint32_t wi = (uint32_t)(r >> 4), s = (uint32_t)(r & 15) * 2;
        uint32_t lo = t.pk[wi];
        if (s == 0) return lo;
        return (lo >> s) | (t.pk[wi + 1] << (32 - s));
    }
    if (r <= -16) return 0;
    uint32_t sh = (uint32_t)(-r) * 2;
    return t.pk[0] << sh;
}

GG_DEV uint32_t rev_comp16(uint32_t x) {
    uint32_t y = __brev(x);
    y = ((y >> 1) & 0x55555555u) | ((y & 0x55555555u) << 1);
    return y ^ 0xAAAAAAAAu;
}

GG_DEV void tile_minima_lds(Tile& t);
GG_DEV void tile_minima_w16(Tile& t);
GG_DEV void tile_flags(Tile& t);

// Phases 0-3: load, hash, window minima, per-window flags.  WC == 16 selects
// the register path for w == 16 (k = 31, m = 15 and other k - m == 16).
template <int WC>
GG_DEV void tile_compute(Tile& t) {
    const BucketArgs& A = *t.A;
    const TileGeom& G = t.G;
    const uint32_t lane = threadIdx.x;

    for (uint32_t i = lane; i < 40; i += 64) t.gt[i] = A.gtab[i];

    // ---- phase 0: bases -> 2-bit codes + invalid bits --------------------
    const uint32_t nchunk = G.NB / 16;
    for (uint32_t c = lane; c < nchunk; c += 64) {
        int64_t a = t.borg + (int64_t)c * 16;
        uint32_t packed = 0, invm = 0;
        if (a >= 0 && a + 16 <= (int64_t)A.n) {
            uint4 v = *reinterpret_cast<const uint4*>(A.seq + a);
            uint32_t wds[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int q = 0; q < 4; q++) {
#pragma unroll
                for (int bb = 0; bb < 4; bb++) {
                    uint8_t ch = (wds[q] >> (8 * bb)) & 0xFF;
                    uint32_t idx = q * 4 + bb;
                    packed |= (uint32_t)((ch >> 1) & 3) << (2 * idx);
                    invm |= (is_acgt(ch) ? 0u : 1u) << idx;
                }
            }
        } else {
            for (uint32_t idx = 0; idx < 16; idx++) {
                int64_t p = a + idx;
                uint8_t ch = (p >= 0 && p < (int64_t)A.n) ? A.seq[p] : (uint8_t)'\n';
                packed |= (uint32_t)((ch >> 1) & 3) << (2 * idx);
                invm |= (is_acgt(ch) ? 0u : 1u) << idx;
            }
        }
        t.pk[c] = packed;
        reinterpret_cast<uint16_t*>(t.inv)[c] = (uint16_t)invm;
    }
    __syncthreads();

    if (WC == 16) {
        tile_minima_w16(t);  // also writes the flag bitmaps
    } else {
        tile_minima_lds(t);
        __syncthreads();
        tile_flags(t);
    }
}

// Phase 1 + 2, any w: every m-mer hash goes to LDS, then van Herk blocks.
GG_DEV void tile_minima_lds(Tile& t) {
    const BucketArgs& A = *t.A;
    const TileGeom& G = t.G;
    const uint32_t lane = threadIdx.x;
    const uint32_t m = A.m, w = A.w;

    // ---- phase 1: canonical ntHash of every m-mer (rolling strips) -------
    // fw_i = rotl1(fw_{i-1}) ^ rotl(h(s_{i-1}), m) ^ h(s_{i+m-1})
    // rc_i = rotr1(rc_{i-1}) ^ rotr1(rc(s_{i-1})) ^ rotl(rc(s_{i+m-1}), m-1)
    // with the "out" term dropped while the first m bases are fed.
    {
        const uint32_t S = (G.NM + 63) / 64;
        uint32_t p0 = lane * S, p1 = p0 + S;
        if (p1 > G.NM) p1 = G.NM;
        if (p0 < p1) {
            uint64_t fw = 0, rc = 0;
            const int64_t rb = (int64_t)t.d0;  // relative base of m-mer 0
            for (uint32_t j = 0; j + 1 < m; j++) {
                uint32_t e = 16 + t.code_rel(rb + p0 + j);
                fw = rotl1_64(fw) ^ t.gt[2 * e];
                rc = rotr1_64(rc) ^ t.gt[2 * e + 1];
            }
            for (uint32_t p = p0; p < p1; p++) {
                uint32_t in = t.code_rel(rb + p + m - 1);
                uint32_t out = (p == p0) ? 4u : t.code_rel(rb + p - 1);
                uint32_t e = out * 4 + in;
                fw = rotl1_64(fw) ^ t.gt[2 * e];
                rc = rotr1_64(rc) ^ t.gt[2 * e + 1];
                uint64_t mn = fw < rc ? fw : rc;
                t.mm[p] = (mn << 1) | (fw != rc ? 1ULL : 0ULL);
                t.FWD(p) = fw < rc ? 1 : 0;
            }
        }
    }
    __syncthreads();

    // ---- phase 2: window minima, van Herk / Gil-Werman blocks of w -------
    for (uint32_t q = lane; q < G.NWB; q += 64) {
        const uint32_t base = q * w;
        uint64_t sv = t.mm[base + w - 1];
        uint32_t sp = base + w - 1;
        t.WV(base + w - 1) = sv;
        t.WP(base + w - 1) = (uint16_t)sp;
        for (int i = (int)w - 2; i >= 0; i--) {
            mm_combine(t.mm[base + i], base + i, sv, sp, sv, sp);
            t.WV(base + i) = sv;
            t.WP(base + i) = (uint16_t)sp;
        }
        uint64_t pv = 0;
        uint32_t pp = 0;
        const uint32_t nb = base + w;
        for (uint32_t i = 1; i < w; i++) {
            if (i == 1) {
                pv = t.mm[nb];
                pp = nb;
            } else {
                mm_combine(pv, pp, t.mm[nb + i - 1], nb + i - 1, pv, pp);
            }
            uint64_t rv;
            uint32_t rp;
            mm_combine(t.WV(base + i), t.WP(base + i), pv, pp, rv, rp);
            t.WV(base + i) = rv;
            t.WP(base + i) = (uint16_t)rp;
        }
    }
}

// 64 bits of a bit array starting at bit `r` (r >= 0).
GG_DEV uint64_t bits64(const uint32_t* a, int64_t r) {
    const uint32_t wi = (uint32_t)(r >> 5), sh = (uint32_t)(r & 31);
    const uint64_t lo = (uint64_t)a[wi] | ((uint64_t)a[wi + 1] << 32);
    if (sh == 0) return lo;
    return (lo >> sh) | ((uint64_t)a[wi + 2] << (64 - sh));
}

// Phases 1-3 for w == 16 (k <= 47), one window block of 16 per lane (the
// tile geometry makes NWB <= 64).  The lane owning window block q hashes
// m-mer blocks q and q+1 (rolling update, tables in LDS), takes suffix minima
// of block q and running prefix minima of block q+1; window q*16+i is the
// combination of the two (van Herk / Gil-Werman in registers).  The window
// flags of phase 3 are formed in the same pass from a 64-bit slice of the
// invalid-base bitmap and written as 16-bit pieces.
GG_DEV void tile_minima_w16(Tile& t) {
    const TileGeom& G = t.G;
    const uint32_t lane = threadIdx.x;
    const uint32_t m = t.A->m, k = t.A->k;
    const int64_t rb = (int64_t)t.d0;  // relative base of m-mer 0 (and of window 0)
    const ulonglong2* gt2 = reinterpret_cast<const ulonglong2*>(t.gt);
    const uint32_t q = lane;
    const bool have = q < G.NWB;
    const uint32_t p0 = q * 16;
    uint32_t segm = 0, okm = 0, lastm = 0;
    bool need_prev = false;
    // the lane's first and last window minima (compared across lanes below)
    uint64_t first_v = 0, last_v = 0;
    uint32_t first_p = 0, last_p = 0;
    // every lane q <= NWB hashes m-mer block q once; lane q takes block q+1
    // from lane q+1 by a lane shift instead of hashing it again
    uint64_t sv[16], nv[16];
    uint32_t sp[16];
    uint32_t fbits = 0;
    if (q <= G.NWB) {
        uint64_t fw = 0, rc = 0;
        for (uint32_t j = 0; j + 1 < m; j++) {
            const ulonglong2 g = gt2[16 + t.code_rel(rb + p0 + j)];
            fw = rotl1_64(fw) ^ g.x;
            rc = rotr1_64(rc) ^ g.y;
        }
        const uint32_t inw = extract16(t, rb + p0 + m - 1);
        const uint32_t outw = extract16(t, rb + p0 - 1);
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const uint32_t in = (inw >> (2 * i)) & 3u, out = (outw >> (2 * i)) & 3u;
            const ulonglong2 g = gt2[(i == 0 ? 16u : out * 4u) + in];
            fw = rotl1_64(fw) ^ g.x;
            rc = rotr1_64(rc) ^ g.y;
            const uint64_t mn = fw < rc ? fw : rc;
            sv[i] = (mn << 1) | (fw != rc ? 1ULL : 0ULL);
            sp[i] = p0 + i;
            fbits |= (fw < rc ? 1u : 0u) << i;
        }
        t.fwd16[q] = (uint16_t)fbits;  // strands of m-mers p0 .. p0 + 15
    } else {
#pragma unroll
        for (int i = 0; i < 16; i++) {
            sv[i] = 0;
            sp[i] = 0;
        }
    }
#pragma unroll
    for (int i = 0; i < 16; i++) nv[i] = lane_next64(sv[i]);  // every lane takes part
    if (have) {
#pragma unroll
        for (int i = 14; i >= 0; i--) mm_combine(sv[i], sp[i], sv[i + 1], sp[i + 1], sv[i], sp[i]);
        // invalid bits of relative bases rb + p0 - 1 .. rb + p0 + 62
        const uint64_t im = bits64(t.inv, rb + p0 - 1);
        const uint64_t kmask = (1ULL << (k - 1)) - 1;
        {
            const uint32_t p1 = p0 + 16;
            uint64_t pv = 0, prv = 0;
            uint32_t pp = 0, prp = 0;
            uint32_t offs[4] = {0u, 0u, 0u, 0u};  // minimizer offsets in the window, one byte each
            uint32_t* wv_q = t.wv32 + q * 17;
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const uint64_t v = nv[i];  // m-mer p1 + i
                uint64_t rv = sv[i];
                uint32_t rp = sp[i];
                if (i > 0) mm_combine(sv[i], sp[i], pv, pp, rv, rp);
                wv_q[i] = (uint32_t)rv;
                offs[i >> 2] |= (rp - (p0 + i)) << (8 * (i & 3));
                if (i == 0) {
                    first_v = rv;
                    first_p = rp;
                }
                if (i == 15) {
                    last_v = rv;
                    last_p = rp;
                }
                // flags of window x = p0 + i (bases rb + x .. rb + x + k - 2)
                const uint32_t x = p0 + i;
                if (x >= 1 && x < G.NW) {
                    const bool valid = ((im >> (i + 1)) & kmask) == 0;
                    const bool first = (im >> i) & 1;
                    const bool lst = (im >> (i + k)) & 1;
                    const bool okf = valid && !(first && lst);
                    if (okf) {
                        okm |= 1u << i;
                        if (lst) lastm |= 1u << i;
                        if (first) {
                            segm |= 1u << i;
                        } else if (i == 0) {
                            need_prev = true;
                        } else if (rv != prv || ((rv & 1) && rp != prp)) {
                            segm |= 1u << i;
                        }
                    }
                }
                prv = rv;
                prp = rp;
                if (i == 0) {
                    pv = v;
                    pp = p1;
                } else {
                    mm_combine(pv, pp, v, p1 + i, pv, pp);
                }
            }
            uint32_t* o4 = reinterpret_cast<uint32_t*>(t.wp8 + 20 * q);
#pragma unroll
            for (int j = 0; j < 4; j++) o4[j] = offs[j];
        }
    }
    // window p0 starts a group when its minimum differs from window p0 - 1,
    // the previous lane's last window (every lane takes part in the shift)
    const uint64_t prev_v = lane_prev64(last_v);
    const uint32_t prev_p = lane_prev32(last_p);
    if (need_prev && (prev_v != first_v || ((first_v & 1) && first_p != prev_p))) segm |= 1u;
    // bitmap words past the last block stay zero
    const uint32_t npieces = (G.NW / 64 + 2) * 4;
    for (uint32_t pc = q; pc < npieces; pc += 64) {
        const bool mine = pc == q && have;
        reinterpret_cast<uint16_t*>(t.seg)[pc] = mine ? (uint16_t)segm : 0;
        reinterpret_cast<uint16_t*>(t.ok)[pc] = mine ? (uint16_t)okm : 0;
        reinterpret_cast<uint16_t*>(t.last)[pc] = mine ? (uint16_t)lastm : 0;
    }
    __syncthreads();
}

// Phase 3: per-window flags.
GG_DEV void tile_flags(Tile& t) {
    const BucketArgs& A = *t.A;
    const TileGeom& G = t.G;
    const uint32_t lane = threadIdx.x;
    const uint32_t k = A.k;
    const uint32_t nwords64 = (G.NW + 63) / 64;
    for (uint32_t c = 0; c < nwords64; c++) {
        uint32_t x = c * 64 + lane;
        bool okf = false, lastf = false, segf = false;
        if (x >= 1 && x < G.NW) {
            int64_t r = (int64_t)t.d0 + x;  // relative base of window start
            bool valid = !t.any_inv(r, k - 1);
            bool first = t.inv_rel(r - 1);
            bool lst = t.inv_rel(r + k - 1);
            okf = valid && !(first && lst);
            lastf = okf && lst;
            if (okf) {
                uint64_t v0 = t.WV(x - 1), v1 = t.WV(x);
                bool differs = (v0 != v1) || ((v1 & 1) && t.WP(x) != t.WP(x - 1));
                segf = first || differs;
            }
        }
        uint64_t bs = __ballot(segf), bo = __ballot(okf), bl = __ballot(lastf);
        if (lane == 0) {
            reinterpret_cast<uint64_t*>(t.seg)[c] = bs;
            reinterpret_cast<uint64_t*>(t.ok)[c] = bo;
            reinterpret_cast<uint64_t*>(t.last)[c] = bl;
        }
    }
    __syncthreads();
}

GG_DEV bool tile_is_end(const Tile& t, uint32_t x) {
    if (!t.bit(t.ok, x)) return false;
    return t.bit(t.last, x) || t.bit(t.seg, x + 1);
}

// Sub-bucket of a window minimum: bucket = (h >> 1) mod 2^b for a unique
// minimizer, the duplicates bucket 2^b otherwise; second bucket = next sb bits
// (crates/assembler_minimizer_bucketing/src/lib.rs:218-255,
//  crates/hashes/src/cn_nthash.rs:135-142).
GG_DEV uint32_t sub_of(uint64_t v, uint32_t b, uint32_t sb) {
    uint32_t bucket = (v & 1) ? (uint32_t)((v >> 1) & ((1ULL << b) - 1)) : (1u << b);
    uint32_t second = (uint32_t)((v >> (b + 1)) & ((1ULL << sb) - 1));
    return (bucket << sb) | second;
}

// Physical unit: the sub-bucket refined by xb more minimizer bits (unique
// minimizers only; duplicates-bucket records keep one unit per sub-bucket).
GG_DEV uint32_t unit_of(uint64_t v, uint32_t b, uint32_t sb, uint32_t xb) {
    uint32_t extra = (v & 1) ? (uint32_t)((v >> (b + 1 + sb)) & ((1ULL << xb) - 1)) : 0u;
    return (sub_of(v, b, sb) << xb) | extra;
}

GG_DEV void tile_init(Tile& t, const BucketArgs* A, uint64_t tile, uint8_t* lds, bool regs) {
    t.A = A;
    t.G = tile_geom(A->T, A->w, A->k, A->m, regs);
    t.t0 = (int64_t)tile * A->T;
    t.wlo = t.t0 - A->w;
    int64_t bo = t.wlo - 1;
    t.borg = (bo >= 0) ? (bo / 16) * 16 : -(((-bo) + 15) / 16) * 16;
    t.d0 = (uint32_t)(t.wlo - t.borg);
    t.lds = lds;
    t.gt = reinterpret_cast<uint64_t*>(lds + t.G.off_g);
    t.mm = reinterpret_cast<uint64_t*>(lds + t.G.off_mm);
    t.wv = reinterpret_cast<uint64_t*>(lds + t.G.off_wv);
    t.wp = reinterpret_cast<uint16_t*>(lds + t.G.off_wp);
    t.pk = reinterpret_cast<uint32_t*>(lds + t.G.off_pk);
    t.inv = reinterpret_cast<uint32_t*>(lds + t.G.off_inv);
    t.fwd = lds + t.G.off_fwd;
    t.regs = regs;
    t.wv32 = reinterpret_cast<uint32_t*>(lds + t.G.off_wv);
    t.wp8 = lds + t.G.off_wp;
    t.fwd16 = reinterpret_cast<uint16_t*>(lds + t.G.off_fwd);
    t.seg = reinterpret_cast<uint32_t*>(lds + t.G.off_seg);
    t.ok = reinterpret_cast<uint32_t*>(lds + t.G.off_ok);
    t.last = reinterpret_cast<uint32_t*>(lds + t.G.off_last);
}

// w == 16 path: group ends among the 16 windows of the lane's block that the
// tile owns (windows w .. w+T-1, absolute index below n).  Returns the mask
// and sets *x0 / *j0 (first window, relative and absolute) and *segm.
GG_DEV uint32_t lane_end_mask16(const Tile& t, uint32_t* x0, int64_t* j0, uint32_t* segm) {
    const uint32_t q = threadIdx.x;
    *x0 = q * 16;
    *j0 = t.t0 + (int64_t)q * 16 - 16;
    *segm = 0;
    if (q < 1 || q > t.A->T / 16) return 0;
    const uint32_t sg = reinterpret_cast<const uint16_t*>(t.seg)[q];
    const uint32_t okm = reinterpret_cast<const uint16_t*>(t.ok)[q];
    const uint32_t lastm = reinterpret_cast<const uint16_t*>(t.last)[q];
    const uint32_t next = t.bit(t.seg, *x0 + 16) ? 1u : 0u;
    const int64_t left = (int64_t)t.A->n - *j0;
    const uint32_t vm = left >= 16 ? 0xFFFFu : (left <= 0 ? 0u : (1u << left) - 1u);
    *segm = sg & vm;
    return okm & (lastm | (sg >> 1) | (next << 15)) & vm;
}

// Code of the base at absolute position p (LDS when inside the tile, else HBM).
GG_DEV uint32_t code_abs(const Tile& t, int64_t p) {
    int64_t r = p - t.borg;
    if (r >= 0 && r < (int64_t)t.G.NB) return t.code_rel(r);
    return (t.A->seq[p] >> 1) & 3;
}
GG_DEV bool invalid_abs(const Tile& t, int64_t p) {
    if (p < 0 || p >= (int64_t)t.A->n) return true;
    int64_t r = p - t.borg;
    if (r >= 0 && r < (int64_t)t.G.NB) return t.inv_rel(r);
    return !is_acgt(t.A->seq[p]);
}

constexpr uint32_t NO_UNIT = 0xFFFFFFFFu;

// Color of the record starting at byte `start`: the last color range that
// starts at or before it.
GG_DEV uint32_t color_of(const BucketArgs& A, int64_t start) {
    uint32_t lo = 0, hi = A.n_colors;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (A.color_starts[mid] <= (uint64_t)start) lo = mid; else hi = mid;
    }
    return A.colors[lo];
}

// Packs the bases [start, end) (reverse complement when rc) into dst words,
// reading the tile's LDS codes where they cover the bases and the input
// otherwise.
GG_DEV void pack_bases(const Tile& t, int64_t start, int64_t end, bool rc, uint32_t* dst) {
    const uint32_t len = (uint32_t)(end - start);
    uint32_t acc = 0;
    for (uint32_t i = 0; i < len; i++) {
        const uint32_t cd = rc ? (code_abs(t, end - 1 - i) ^ 2u) : code_abs(t, start + i);
        acc |= cd << (2 * (i & 15));
        if ((i & 15) == 15 || i + 1 == len) {
            dst[i >> 4] = acc;
            acc = 0;
        }
    }
}

// Long record: bases to the pool, the slot holds length and pool offset.
// Returns false when the pool is full (the host retries with a larger one).
GG_DEV bool write_long(const BucketArgs& A, const Tile& t, uint32_t* out, int64_t start, int64_t end, bool rc,
                       uint32_t flags, uint32_t minpos) {
    const uint32_t len = (uint32_t)(end - start);
    const uint64_t words = (len + 15) / 16;
    const unsigned long long off = atomicAdd(A.pool_cursor, (unsigned long long)words);
    out[0] = (flags << 8) | (minpos << 10) | META_LONG | (rc ? META_RC : 0u);
    out[1] = len;
    out[2] = (uint32_t)off;
    out[3] = (uint32_t)(off >> 32);
    for (uint32_t q = 4; q < A.slot_words; q++) out[q] = 0;
    if (off + words > A.pool_cap) return false;
    pack_bases(t, start, end, rc, A.pool + off);
    return true;
}

// Counts a record in its unit and returns its rank there (records << 32 |
// k-mer occurrences in one 64-bit atomic); flags a unit whose k-mer count
// would carry into the record count.
GG_DEV void count_record(const BucketArgs& A, uint32_t unit, uint32_t nk, unsigned long long* old) {
    *old = atomicAdd(&A.unit_ck[unit], (1ULL << 32) | nk);
}
GG_DEV uint32_t rank_of(const BucketArgs& A, unsigned long long old, uint32_t nk) {
    if ((uint32_t)old + nk < (uint32_t)old) atomicOr(A.error_flag, 8u);
    return (uint32_t)(old >> 32);
}

// One record of a tile: everything the slot needs.
struct RecInfo {
    int64_t start, end;
    uint32_t unit, meta, nk;  // meta: flags << 8 | minpos << 10 | rc | cross marks (length added on write)
    bool rc, ok;
};

GG_DEV RecInfo rec_info(const Tile& t, uint32_t x) {
    const BucketArgs& A = *t.A;
    const uint32_t w = A.w, k = A.k, m = A.m;
    RecInfo r;
    r.ok = false;
    const int64_t bwin = t.t0 + (int64_t)x - w;
    // group start: latest group-start bit at or before x inside the tile
    int64_t a = -1;
    {
        int32_t wi = (int32_t)(x >> 5);
        uint32_t msk = t.seg[wi] & ((2u << (x & 31)) - 1u);
        while (true) {
            if (msk) {
                a = t.wlo + (int64_t)wi * 32 + (31 - __clz(msk));
                break;
            }
            if (--wi < 0) break;
            msk = t.seg[wi];
        }
    }
    if (a < 0) return r;  // group starts before the halo: deferred
    const uint64_t v = t.wval(x);
    const bool unique = v & 1;
    const bool include_first = invalid_abs(t, a - 1);
    const bool include_last = t.bit(t.last, x);
    r.start = include_first ? a : a - 1;
    r.end = include_last ? bwin + k - 1 : bwin + k;
    const uint32_t prel = t.wpos(x);
    const int64_t mmabs = t.wlo + prel;
    r.rc = A.canonical && unique && !t.fwd_at(prel);
    uint32_t minpos = 0;
    if (unique) minpos = r.rc ? (uint32_t)(r.end - mmabs - m) : (uint32_t)(mmabs - r.start);
    const uint32_t flags = ((include_first ? 1u : 0u) << (r.rc ? 1 : 0)) | ((include_last ? 1u : 0u) << (r.rc ? 0 : 1));
    const uint32_t sub = sub_of(v, A.b, A.sb);
    r.unit = unit_of(v, A.b, A.sb, A.xb);
    // Boundary k-mers shared with a neighbouring group of the same
    // sub-bucket that was scattered to another unit: the merge must combine
    // the two halves (cross marks, in stored orientation).  A unique group
    // spans at most w windows, so window a-1 is inside the halo.
    uint32_t cross = 0;
    if (unique && A.xb) {
        if (!include_first) {
            const uint64_t vl = t.wval((uint32_t)(a - t.wlo - 1));
            if (sub_of(vl, A.b, A.sb) == sub && unit_of(vl, A.b, A.sb, A.xb) != r.unit) cross |= r.rc ? META_XEND : META_XBEGIN;
        }
        if (!include_last) {
            const uint64_t vr = t.wval(x + 1);
            if (sub_of(vr, A.b, A.sb) == sub && unit_of(vr, A.b, A.sb, A.xb) != r.unit) cross |= r.rc ? META_XBEGIN : META_XEND;
        }
    }
    r.meta = (flags << 8) | (minpos << 10) | (r.rc ? META_RC : 0u) | cross;
    r.nk = (uint32_t)(r.end - r.start) - k + 1;
    r.ok = true;
    return r;
}

// Writes the slot of record r to out (16-byte aligned, slot_words words).
// Returns false when a long record did not fit the pool.
GG_DEV bool write_record(const Tile& t, const RecInfo& r, uint32_t* out) {
    const BucketArgs& A = *t.A;
    const uint32_t SW = A.slot_words;
    const uint32_t inline_cap = (SW - 1) * 16;
    const uint32_t len = (uint32_t)(r.end - r.start);
    const int64_t rs = r.start - t.borg;
    if (len <= inline_cap && rs >= 0) {
        // 16-byte stores (slots are a multiple of 4 words)
        const uint32_t nw = (len + 15) / 16;
        uint4* o4 = reinterpret_cast<uint4*>(out);
        for (uint32_t g = 0; g < SW / 4; g++) {
            uint32_t w4[4];
#pragma unroll
            for (uint32_t j = 0; j < 4; j++) {
                const uint32_t q = 4 * g + j;
                uint32_t wq = 0;
                if (q == 0) {
                    wq = len | r.meta;
                } else if (q - 1 < nw) {
                    const uint32_t qi = q - 1;
                    if (!r.rc) {
                        wq = extract16(t, rs + 16 * qi);
                    } else {
                        wq = rev_comp16(extract16(t, (r.end - t.borg) - 16 * (int64_t)qi - 16));
                    }
                    const uint32_t rem = len - 16 * qi;
                    if (rem < 16) wq &= (1u << (2 * rem)) - 1u;
                }
                w4[j] = wq;
            }
            o4[g] = make_uint4(w4[0], w4[1], w4[2], w4[3]);
        }
        return true;
    }
    if (len > inline_cap) return write_long(A, t, out, r.start, r.end, r.rc, (r.meta >> 8) & 3u, (r.meta >> 10) & 0xFFu);
    // bases start before the tile's loaded range
    out[0] = len | r.meta;
    for (uint32_t q = 1; q < SW; q++) out[q] = 0;
    pack_bases(t, r.start, r.end, r.rc, out + 1);
    return true;
}

// Slot words (up to 8) of an inline record whose bases are in the tile's
// LDS, built in registers; false when the record takes the generic path
// (write_record: long records, bases before the tile, slots above 8 words).
GG_DEV bool record_words(const Tile& t, const RecInfo& r, uint4* w4) {
    const BucketArgs& A = *t.A;
    const uint32_t SW = A.slot_words;
    const uint32_t len = (uint32_t)(r.end - r.start);
    const int64_t rs = r.start - t.borg;
    if (SW > 8 || len > (SW - 1) * 16 || rs < 0) return false;
    const uint32_t nw = (len + 15) / 16;
#pragma unroll
    for (uint32_t g = 0; g < 2; g++) {
        if (4 * g >= SW) break;
        uint32_t wv[4];
#pragma unroll
        for (uint32_t j = 0; j < 4; j++) {
            const uint32_t q = 4 * g + j;
            uint32_t wq = 0;
            if (q == 0) {
                wq = len | r.meta;
            } else if (q - 1 < nw) {
                const uint32_t qi = q - 1;
                wq = !r.rc ? extract16(t, rs + 16 * qi) : rev_comp16(extract16(t, (r.end - t.borg) - 16 * (int64_t)qi - 16));
                const uint32_t rem = len - 16 * qi;
                if (rem < 16) wq &= (1u << (2 * rem)) - 1u;
            }
            wv[j] = wq;
        }
        w4[g] = make_uint4(wv[0], wv[1], wv[2], wv[3]);
    }
    return true;
}

// Where record `rank` of `unit` goes: its region slot, or the next spill
// entry (nullptr when the spill list is full: the host retries with the
// counted size).  *color_at receives the matching color entry.
GG_DEV uint32_t* record_dest(const BucketArgs& A, uint32_t unit, uint32_t rank, uint32_t** color_at) {
    const uint32_t SW = A.slot_words;
    if (rank < A.unit_cap) {
        const uint64_t slot = (uint64_t)unit * A.unit_cap + rank;
        *color_at = A.region_color ? A.region_color + slot : nullptr;
        return A.region + slot * SW;
    }
    const unsigned long long pos = atomicAdd(A.spill_cursor, 1ULL);
    *color_at = nullptr;
    if (pos >= A.spill_cap) return nullptr;
    A.spill_unit[pos] = unit;
    A.spill_rank[pos] = rank;
    if (A.spill_color) *color_at = A.spill_color + pos;
    return A.spill_slots + pos * SW;
}

// Records handled per lane and pass of the record loop; their unit atomics
// are all in flight while their slot words are built.
constexpr uint32_t STAGE_R = 2;

// Pass 1 (one wavefront per tile): window minima and group ends; every record
// whose group starts inside the tile or its halo is counted in its unit
// (which gives its rank there), built and staged at the tile's positions; the
// first record of a tile whose group starts further back is deferred to
// bucket_defer_kernel, which runs once the group-start bitmap is complete.
template <int WC>
__global__ __launch_bounds__(64) void bucket_stage_kernel(BucketArgs A) {
    extern __shared__ __align__(16) uint8_t lds[];
    Tile t;
    tile_init(t, &A, A.tile0 + blockIdx.x, lds, WC != 0);
    tile_compute<WC>(t);
    const uint32_t lane = threadIdx.x;
    const uint32_t w = A.w;
    const uint32_t SW = A.slot_words;
    __shared__ uint16_t ends[1024];
    __shared__ uint32_t n_ends;
    if (WC == 16) {
        uint32_t x0, segm;
        int64_t j0;
        uint32_t endm = lane_end_mask16(t, &x0, &j0, &segm);
        if (lane >= 1 && lane <= A.T / 16) reinterpret_cast<uint16_t*>(A.seg_bits)[j0 >> 4] = (uint16_t)segm;
        const uint32_t cnt = __popc(endm);
        uint32_t inc = cnt;
        for (uint32_t d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(inc, d);
            if (lane >= d) inc += y;
        }
        uint32_t o = inc - cnt;
        while (endm) {
            const uint32_t i = __ffs(endm) - 1;
            endm &= endm - 1;
            ends[o++] = (uint16_t)(x0 + i);
        }
        if (lane == 63) n_ends = inc;
        __syncthreads();
    } else {
        if (lane == 0) n_ends = 0;
        __syncthreads();
        for (uint32_t c = 0; c < A.T; c += 64) {
            const uint32_t x = w + c + lane;
            const int64_t j = t.t0 + c + lane;
            const bool sg = t.bit(t.seg, x) && j < (int64_t)A.n;
            const uint64_t bs = __ballot(sg);
            if (lane == 0) {
                const uint64_t wi = (uint64_t)(t.t0 + c) >> 5;
                A.seg_bits[wi] = (uint32_t)bs;
                A.seg_bits[wi + 1] = (uint32_t)(bs >> 32);
            }
            const bool e = j < (int64_t)A.n && tile_is_end(t, x);
            const uint64_t be = __ballot(e);
            if (e) ends[n_ends + __popcll(be & ((1ULL << lane) - 1))] = (uint16_t)x;
            __syncthreads();
            if (lane == 0) n_ends += (uint32_t)__popcll(be);
            __syncthreads();
        }
    }
    const uint32_t n_end = n_ends;
    for (uint32_t e0 = 0; e0 < n_end; e0 += 64 * STAGE_R) {
        RecInfo ri[STAGE_R];
        unsigned long long old[STAGE_R];
#pragma unroll
        for (uint32_t r = 0; r < STAGE_R; r++) {
            const uint32_t ei = e0 + 64 * r + lane;
            ri[r].ok = false;
            old[r] = 0;
            if (ei >= n_end) continue;
            const uint32_t x = ends[ei];
            ri[r] = rec_info(t, x);
            if (ri[r].ok) {
                count_record(A, ri[r].unit, ri[r].nk, &old[r]);
            } else {
                // group starts before the halo (non-unique minimizer): defer
                const uint32_t d = atomicAdd(A.defer_count, 1u);
                if (d < A.n_tiles) {
                    A.defer[3 * (uint64_t)d] = (uint64_t)(t.t0 + (int64_t)x - w);
                    A.defer[3 * (uint64_t)d + 1] = t.wval(x);
                    A.defer[3 * (uint64_t)d + 2] = t.bit(t.last, x) ? 1 : 0;
                } else {
                    atomicOr(A.error_flag, 4u);
                }
            }
        }
        uint4 wv[STAGE_R][2];
        bool in_regs[STAGE_R];
#pragma unroll
        for (uint32_t r = 0; r < STAGE_R; r++) in_regs[r] = ri[r].ok && record_words(t, ri[r], wv[r]);
#pragma unroll
        for (uint32_t r = 0; r < STAGE_R; r++) {
            if (!ri[r].ok) continue;
            uint32_t* color_at;
            uint32_t* out = record_dest(A, ri[r].unit, rank_of(A, old[r], ri[r].nk), &color_at);
            if (!out) continue;  // spill list full: retried
            if (in_regs[r]) {
                uint4* o4 = reinterpret_cast<uint4*>(out);
                o4[0] = wv[r][0];
                if (SW == 8) o4[1] = wv[r][1];
            } else if (!write_record(t, ri[r], out)) {
                continue;  // pool full: retried
            }
            if (color_at) *color_at = color_of(A, ri[r].start);
        }
    }
}

// Deferred records (group start before the tile's halo; only groups with a
// non-unique minimizer are that long): the group-start bitmap is complete
// now, so the start is the last group start at or before the end window.
// One thread per record; bases are read from the input; staged in the
// overflow region.
__global__ void bucket_defer_kernel(BucketArgs A) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint64_t n_defer = *A.defer_count;
    if (i >= n_defer || i >= A.n_tiles) return;
    const int64_t bwin = (int64_t)A.defer[3 * i];
    const uint64_t v = A.defer[3 * i + 1];
    const bool include_last = A.defer[3 * i + 2] != 0;
    const uint32_t k = A.k, SW = A.slot_words, inline_cap = (SW - 1) * 16;
    int64_t gw = bwin >> 5;
    uint32_t gm = A.seg_bits[gw] & ((2u << (bwin & 31)) - 1u);
    while (gm == 0 && gw > 0) gm = A.seg_bits[--gw];
    if (!gm) {
        atomicOr(A.error_flag, 4u);
        return;
    }
    const int64_t a = gw * 32 + (31 - __clz(gm));
    const bool include_first = (a == 0) || !is_acgt(A.seq[a - 1]);
    const int64_t start = include_first ? a : a - 1;
    const int64_t end = include_last ? bwin + k - 1 : bwin + k;
    const uint32_t len = (uint32_t)(end - start);
    // a non-unique minimizer: forward orientation, no minimizer position
    const uint32_t flags = (include_first ? 1u : 0u) | ((include_last ? 1u : 0u) << 1);
    const uint32_t unit = unit_of(v, A.b, A.sb, A.xb);
    unsigned long long old;
    count_record(A, unit, len - k + 1, &old);
    uint32_t* color_at;
    uint32_t* out = record_dest(A, unit, rank_of(A, old, len - k + 1), &color_at);
    if (!out) return;  // spill list full: the host retries with the counted size
    Tile none;  // no LDS codes: every base comes from the input
    none.A = &A;
    none.borg = 0;
    none.G.NB = 0;
    bool ok = true;
    if (len > inline_cap) {
        ok = write_long(A, none, out, start, end, false, flags, 0);
    } else {
        out[0] = len | (flags << 8);
        for (uint32_t q = 1; q < SW; q++) out[q] = 0;
        pack_bases(none, start, end, false, out + 1);
    }
    if (ok && color_at) *color_at = color_of(A, start);
}

// Splits the packed per-unit counts into records and k-mer occurrences.
// Also sums the k-mer occurrences into total[0] and the records into
// total[1] (cleared by the reset kernel; blocks of 256 threads).
__global__ __launch_bounds__(256) void unit_split_kernel(const unsigned long long* ck, uint32_t* count, uint32_t* kmers,
                                                        uint32_t n, unsigned long long* total) {
    __shared__ unsigned long long part[4], rpart[4];
    unsigned long long sum = 0, rsum = 0;  // k-mer occurrences, records
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
        const unsigned long long c = ck[u];
        count[u] = (uint32_t)(c >> 32);
        kmers[u] = (uint32_t)c;
        sum += (uint32_t)c;
        rsum += c >> 32;
    }
    for (uint32_t d = 32; d; d >>= 1) {
        sum += __shfl_down(sum, d);
        rsum += __shfl_down(rsum, d);
    }
    if ((threadIdx.x & 63) == 0) {
        part[threadIdx.x >> 6] = sum;
        rpart[threadIdx.x >> 6] = rsum;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long t = part[0] + part[1] + part[2] + part[3];
        const unsigned long long r = rpart[0] + rpart[1] + rpart[2] + rpart[3];
        if (t) atomicAdd(total, t);
        if (r) atomicAdd(total + 1, r);
    }
}

// Records and bases per sub-bucket (sums over its 2^xb units).
__global__ void sub_totals_kernel(const uint32_t* unit_count, const uint32_t* unit_kmers, uint32_t n_sub, uint32_t xb,
                                  uint32_t k, uint64_t* records, uint64_t* bases) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_sub) return;
    uint64_t r = 0, b = 0;
    for (uint32_t u = s << xb; u < (s + 1) << xb; u++) {
        r += unit_count[u];
        b += unit_kmers[u] + (uint64_t)unit_count[u] * (k - 1);
    }
    records[s] = r;
    bases[s] = b;
}

// Clears the cursors of pass 1 (a kernel, so the pass can be captured in a graph).
__global__ void bucket_reset_kernel(BucketArgs A) {
    if (threadIdx.x == 0) {
        *A.spill_cursor = 0;
        *A.pool_cursor = 0;
        A.kmer_total[0] = 0;
        A.kmer_total[1] = 0;
        *A.defer_count = 0;
        *A.error_flag = 0;
    }
}

// Pass 2a: identical records of a unit counted once, as the reference's
// compactor merges identical super-k-mers of a bucket
// (crates/minimizer_bucketing/src/compactor.rs:128-196: same bases, length
// and flags; multiplicities summed).  Identical super-k-mers share their
// minimizer, so all copies of a record live in one unit.  One workgroup per
// unit whose region holds 2..DEDUP_WIDE records: 64 threads for units of up to
// DEDUP_SMALL records, 256 above (two launches; the small one also passes the
// other units through).  The records are loaded into LDS with 16-byte loads,
// each is inserted into an LDS table keyed by its slot words and color, a
// record equal to an earlier inserted one adds to that one's multiplicity (up
// to MAX_MULT per stored record), and the stored records are written back, in
// region order, to the front of the region with multiplicity - 1 in their
// meta word.  Spilled records (rank >= unit_cap) are stored as they are.
constexpr uint32_t DEDUP_SMALL = 256;
constexpr uint32_t DEDUP_WIDE = 1024;

GG_DEV uint32_t slot_hash(const uint32_t* w, uint32_t sw, uint32_t color) {
    uint32_t h = 0x9E3779B9u ^ color;
    for (uint32_t i = 0; i < sw; i++) {
        h ^= w[i];
        h *= 0x85EBCA6Bu;
        h ^= h >> 13;
    }
    h *= 0xC2B2AE35u;
    return h ^ (h >> 16);
}

// Dynamic LDS of a dedup workgroup taking up to max_rec records: their slot
// words, then (colored builds) their colors.
__host__ __device__ inline uint32_t dedup_lds_bytes(uint32_t max_rec, uint32_t slot_words, bool colored) {
    return max_rec * slot_words * 4 + (colored ? max_rec * 4 : 0);
}

// NT threads per workgroup, units of LO < records <= HI (and, with PASS, the
// other units kept as they are, except those of (DEDUP_SMALL, DEDUP_WIDE]
// when the wide launch takes them).  At most 8 16-byte pieces per thread:
// HI * slot_words / 4 <= 8 * NT.
template <uint32_t NT, uint32_t LO, uint32_t HI, bool PASS>
GG_DEV void unit_dedup_body(const BucketArgs& A, uint32_t n_units, bool wide) {
    extern __shared__ __align__(16) uint8_t lds[];
    __shared__ uint32_t mult[HI];  // 0: a copy of another record
    __shared__ uint32_t table[2 * HI];  // record index + 1, 0 = free
    __shared__ uint32_t wave_n[NT / 64];
    __shared__ uint32_t folded_k;  // k-mer positions of the copies folded into earlier records
    constexpr uint32_t TABLE = 2 * HI;
    const uint32_t SW = A.slot_words, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint32_t* rec = reinterpret_cast<uint32_t*>(lds);  // slot words
    uint32_t* col = rec + HI * SW;  // colored builds only
    auto mine = [&](uint32_t n) { return n > LO && n <= HI; };
    // a unit's region is loaded into registers while the previous unit is
    // processed
    uint4 p0 = {}, p1 = {}, p2 = {}, p3 = {}, p4 = {}, p5 = {}, p6 = {}, p7 = {};
#define GG_DEDUP_PREFETCH(UU, CNT)                                                                 \
    do {                                                                                           \
        const uint32_t n_ = (CNT) < A.unit_cap ? (CNT) : A.unit_cap;                               \
        if (mine(n_)) {                                                                            \
            const uint4* s_ = reinterpret_cast<const uint4*>(A.region + (uint64_t)(UU) * A.unit_cap * SW); \
            const uint32_t m_ = n_ * (SW / 4);                                                     \
            if (tid < m_) p0 = s_[tid];                                                            \
            if (tid + NT < m_) p1 = s_[tid + NT];                                                  \
            if (tid + 2 * NT < m_) p2 = s_[tid + 2 * NT];                                          \
            if (tid + 3 * NT < m_) p3 = s_[tid + 3 * NT];                                          \
            if (tid + 4 * NT < m_) p4 = s_[tid + 4 * NT];                                          \
            if (tid + 5 * NT < m_) p5 = s_[tid + 5 * NT];                                          \
            if (tid + 6 * NT < m_) p6 = s_[tid + 6 * NT];                                          \
            if (tid + 7 * NT < m_) p7 = s_[tid + 7 * NT];                                          \
        }                                                                                          \
    } while (0)
    uint32_t u = blockIdx.x;
    uint32_t cnt = u < n_units ? A.sub_count[u] : 0u;
    if (u < n_units) GG_DEDUP_PREFETCH(u, cnt);
    for (; u < n_units; u += gridDim.x) {
        const uint32_t n = cnt < A.unit_cap ? cnt : A.unit_cap;
        const uint32_t spilled = cnt - n;
        uint32_t* reg = A.region + (uint64_t)u * A.unit_cap * SW;
        uint32_t* regc = A.region_color ? A.region_color + (uint64_t)u * A.unit_cap : nullptr;
        const uint32_t un = u + gridDim.x;
        const uint32_t cnt_next = un < n_units ? A.sub_count[un] : 0u;
        if (!mine(n)) {
            // kept as it is (by the small launch for every unit the wide
            // launch does not take)
            if (PASS && tid == 0 && !(wide && n > DEDUP_SMALL && n <= DEDUP_WIDE)) {
                A.unit_dregion[u] = n;
                A.unit_dcount[u] = cnt;
                A.unit_skmers[u] = A.sub_kmers[u];
            }
            cnt = cnt_next;
            if (un < n_units) GG_DEDUP_PREFETCH(un, cnt);
            continue;
        }
        const uint32_t n4 = n * (SW / 4);
        uint4* r4 = reinterpret_cast<uint4*>(rec);
        if (tid < n4) r4[tid] = p0;
        if (tid + NT < n4) r4[tid + NT] = p1;
        if (tid + 2 * NT < n4) r4[tid + 2 * NT] = p2;
        if (tid + 3 * NT < n4) r4[tid + 3 * NT] = p3;
        if (tid + 4 * NT < n4) r4[tid + 4 * NT] = p4;
        if (tid + 5 * NT < n4) r4[tid + 5 * NT] = p5;
        if (tid + 6 * NT < n4) r4[tid + 6 * NT] = p6;
        if (tid + 7 * NT < n4) r4[tid + 7 * NT] = p7;
        for (uint32_t i = tid; i < n; i += NT) {
            if (regc) col[i] = regc[i];
            mult[i] = 1;
        }
        for (uint32_t i = tid; i < TABLE; i += NT) table[i] = 0;
        if (tid == 0) folded_k = 0;
        __syncthreads();
        // the next unit's records are in flight meanwhile
        cnt = cnt_next;
        if (un < n_units) GG_DEDUP_PREFETCH(un, cnt);
        uint32_t my_folded = 0;
        for (uint32_t i = tid; i < n; i += NT) {
            const uint32_t* w = rec + i * SW;
            if (w[0] & META_LONG) continue;  // pool records: bases elsewhere, kept as they are
            const uint32_t ci = regc ? col[i] : 0u;
            uint32_t slot = slot_hash(w, SW, ci) & (TABLE - 1);
            for (uint32_t probe = 0; probe < TABLE; probe++) {
                uint32_t cur = table[slot];
                if (cur == 0) {
                    cur = atomicCAS(&table[slot], 0u, i + 1);
                    if (cur == 0) break;  // first of its kind: stored
                }
                const uint32_t j = cur - 1;
                const uint32_t* o = rec + j * SW;
                bool eq = !regc || col[j] == ci;
                for (uint32_t q = 0; q < SW && eq; q++) eq = o[q] == w[q];
                if (eq) {
                    // a copy of record j: counted there while j has room
                    uint32_t m = mult[j];
                    while (m < MAX_MULT) {
                        const uint32_t seen = atomicCAS(&mult[j], m, m + 1);
                        if (seen == m) break;
                        m = seen;
                    }
                    if (m < MAX_MULT) {
                        mult[i] = 0;
                        my_folded += meta_len(w[0]) - A.k + 1;  // inserts this copy no longer makes
                    }
                    break;
                }
                slot = (slot + 1) & (TABLE - 1);
            }
        }
        // one LDS atomic per wavefront for the folded copies' k-mer positions
        for (uint32_t d = 32; d >= 1; d >>= 1) my_folded += __shfl_down(my_folded, d);
        if (lane == 0 && my_folded) atomicAdd(&folded_k, my_folded);
        __syncthreads();
        // stored records to the front of the region, in region order
        uint32_t placed = 0;
        for (uint32_t i0 = 0; i0 < n; i0 += NT) {
            const uint32_t i = i0 + tid;
            const uint32_t m = i < n ? mult[i] : 0u;
            const uint64_t bal = __ballot(m != 0);
            uint32_t before = 0, total = (uint32_t)__popcll(bal);
            if (NT > 64) {
                if (lane == 0) wave_n[wave] = total;
                __syncthreads();
                total = 0;
                for (uint32_t q = 0; q < NT / 64; q++) {
                    const uint32_t c = wave_n[q];
                    if (q < wave) before += c;
                    total += c;
                }
            }
            if (m) {
                const uint32_t dst = placed + before + (uint32_t)__popcll(bal & ((1ULL << lane) - 1));
                const uint32_t* w = rec + i * SW;
                uint32_t* d = reg + (uint64_t)dst * SW;
                d[0] = w[0] | ((m - 1) << META_MULT_SHIFT);
                for (uint32_t q = 1; q < SW; q++) d[q] = w[q];
                if (regc) regc[dst] = col[i];
            }
            placed += total;
            if (NT > 64) __syncthreads();  // wave_n is reused by the next step
        }
        if (tid == 0) {
            A.unit_dregion[u] = placed;
            A.unit_dcount[u] = placed + spilled;
            A.unit_skmers[u] = A.sub_kmers[u] - folded_k;
        }
        __syncthreads();  // LDS reused by the next unit
    }
#undef GG_DEDUP_PREFETCH
}

// wide: unit_dedup_wide_kernel runs too and takes the larger units
__global__ __launch_bounds__(64) void unit_dedup_kernel(BucketArgs A, uint32_t n_units, uint32_t wide) {
    unit_dedup_body<64, 1, DEDUP_SMALL, true>(A, n_units, wide != 0);
}

__global__ __launch_bounds__(256) void unit_dedup_wide_kernel(BucketArgs A, uint32_t n_units) {
    unit_dedup_body<256, DEDUP_SMALL, DEDUP_WIDE, false>(A, n_units, true);
}

// Pass 2b: each unit's stored region records (the region's first
// unit_dregion[u] slots) to their place in the dense unit-major arena; one
// wavefront per unit, 16-byte copies.
__global__ __launch_bounds__(256) void unit_compact_kernel(BucketArgs A, uint32_t n_units) {
    const uint32_t SW = A.slot_words;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    for (uint32_t u = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); u < n_units; u += waves) {
        const uint32_t n = A.unit_dregion[u];
        const uint64_t o0 = A.sub_off[u];
        const uint4* src = reinterpret_cast<const uint4*>(A.region + (uint64_t)u * A.unit_cap * SW);
        uint4* dst = reinterpret_cast<uint4*>(A.slots + o0 * SW);
        const uint32_t n4 = n * (SW / 4);
        for (uint32_t i = lane; i < n4; i += 64) dst[i] = src[i];
        if (A.region_color)
            for (uint32_t i = lane; i < n; i += 64) A.slot_color[o0 + i] = A.region_color[(uint64_t)u * A.unit_cap + i];
    }
}

// Pass 2c: spilled records (rank >= unit_cap) after their unit's stored
// region records.
__global__ void spill_place_kernel(BucketArgs A, uint64_t n_spill) {
    const uint32_t SW = A.slot_words;
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n_spill;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t unit = A.spill_unit[i];
        const uint64_t slot = A.sub_off[unit] + A.unit_dregion[unit] + (A.spill_rank[i] - A.unit_cap);
        if (slot >= A.sub_off[unit + 1]) {
            atomicOr(A.error_flag, 1u);
            continue;
        }
        const uint4* src = reinterpret_cast<const uint4*>(A.spill_slots + i * SW);
        uint4* dst = reinterpret_cast<uint4*>(A.slots + slot * SW);
        for (uint32_t g = 0; g < SW / 4; g++) dst[g] = src[g];
        if (A.spill_color) A.slot_color[slot] = A.spill_color[i];
    }
}

// Device-wide exclusive scan of uint32 counts into uint64 offsets (n+1
// entries): per-chunk sums, one-block scan of the sums, per-chunk scan.
constexpr uint32_t SCAN_CHUNK = 4096;  // 1024 threads x 4 elements

GG_DEV uint64_t block_exclusive_scan(uint64_t v, uint64_t* tmp, uint64_t& total) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    uint64_t x = v;
    for (uint32_t d = 1; d < 64; d <<= 1) {
        uint64_t y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    if (lane == 63) tmp[wave] = x;
    __syncthreads();
    if (wave == 0) {
        uint64_t s = lane < nw ? tmp[lane] : 0;
        for (uint32_t d = 1; d < 64; d <<= 1) {
            uint64_t y = __shfl_up(s, d);
            if (lane >= d) s += y;
        }
        if (lane < nw) tmp[lane] = s;
    }
    __syncthreads();
    uint64_t base = wave ? tmp[wave - 1] : 0;
    total = tmp[nw - 1];
    __syncthreads();
    return base + x - v;
}

__global__ __launch_bounds__(1024) void scan_reduce_kernel(const uint32_t* counts, uint64_t* partial, uint32_t n) {
    __shared__ uint64_t tmp[16];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_CHUNK;
    uint64_t s = 0;
    for (uint32_t q = 0; q < 4; q++) {
        uint64_t i = base + (uint64_t)q * 1024 + threadIdx.x;
        if (i < n) s += counts[i];
    }
    uint64_t total;
    block_exclusive_scan(s, tmp, total);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void scan_partials_kernel(uint64_t* partial, uint32_t nb) {
    __shared__ uint64_t tmp[16];
    uint64_t carry = 0;
    for (uint32_t c = 0; c < nb; c += 1024) {
        uint32_t i = c + threadIdx.x;
        uint64_t v = i < nb ? partial[i] : 0;
        uint64_t total;
        uint64_t ex = block_exclusive_scan(v, tmp, total);
        if (i < nb) partial[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) partial[nb] = carry;
}

__global__ __launch_bounds__(1024) void scan_apply_kernel(const uint32_t* counts, const uint64_t* partial,
                                                          uint64_t* off, uint32_t n, uint32_t nb) {
    __shared__ uint64_t tmp[16];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_CHUNK;
    // thread t owns elements base + 4t .. base + 4t + 3
    uint32_t c[4];
    uint64_t s = 0;
    for (uint32_t q = 0; q < 4; q++) {
        uint64_t i = base + 4ull * threadIdx.x + q;
        c[q] = i < n ? counts[i] : 0;
        s += c[q];
    }
    uint64_t total;
    uint64_t run = partial[blockIdx.x] + block_exclusive_scan(s, tmp, total);
    for (uint32_t q = 0; q < 4; q++) {
        uint64_t i = base + 4ull * threadIdx.x + q;
        if (i < n) off[i] = run;
        run += c[q];
    }
    if (blockIdx.x == nb - 1 && threadIdx.x == 0) off[n] = partial[nb];
}

// Exclusive prefix sum of counts -> offsets (n+1 entries), one workgroup.
__global__ __launch_bounds__(1024) void scan_kernel(const uint32_t* counts, uint64_t* off, uint32_t n) {
    __shared__ uint64_t part[1024];
    const uint32_t tid = threadIdx.x;
    const uint32_t per = (n + 1023) / 1024;
    uint32_t s0 = tid * per, s1 = s0 + per;
    if (s1 > n) s1 = n;
    uint64_t sum = 0;
    for (uint32_t i = s0; i < s1; i++) sum += counts[i];
    part[tid] = sum;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {
        uint64_t v = (tid >= d) ? part[tid - d] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    uint64_t run = (tid == 0) ? 0 : part[tid - 1];
    for (uint32_t i = s0; i < s1; i++) {
        off[i] = run;
        run += counts[i];
    }
    if (tid == 1023) off[n] = part[1023];
}

}  // namespace gg
