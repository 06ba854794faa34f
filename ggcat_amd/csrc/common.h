// common.h -- shared device/host definitions for the gfx950 GGCAT front-stage
// kernels (minimizer bucketing + per-sub-bucket k-mer merge).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GG_DEV __device__ __forceinline__

typedef unsigned __int128 u128_t;

// Slot meta word (word 0 of every arena slot).
//   bits  0..7   length in bases for inline records (0 for long records)
//   bits  8..9   READ_FLAG_INCL_BEGIN / READ_FLAG_INCL_END as stored
//   bits 10..17  minimizer position inside the stored record
//   bit  18      long record (bases live in the pool: word1 = length,
//                words 2..3 = 64-bit pool word offset)
//   bit  19      record was reverse-complemented when stored
//   bit  20      first k-mer (stored orientation) is shared with a group of
//                the same sub-bucket that lives in another physical unit
//   bit  21      same for the last k-mer
//   bits 22..31  multiplicity - 1: identical records of a unit are stored
//                once (crates/minimizer_bucketing/src/compactor.rs:128-196)
constexpr uint32_t META_LONG = 1u << 18;
constexpr uint32_t META_RC = 1u << 19;
constexpr uint32_t META_XBEGIN = 1u << 20;
constexpr uint32_t META_XEND = 1u << 21;

GG_DEV uint32_t meta_len(uint32_t m) { return m & 0xFF; }
GG_DEV uint32_t meta_flags(uint32_t m) { return (m >> 8) & 3; }
GG_DEV uint32_t meta_minpos(uint32_t m) { return (m >> 10) & 0xFF; }
constexpr uint32_t META_MULT_SHIFT = 22;
constexpr uint32_t MAX_MULT = 1024;  // per stored record
__host__ __device__ inline uint32_t meta_mult(uint32_t m) { return ((m >> META_MULT_SHIFT) & 0x3FF) + 1; }

// ntHash constants (crates/hashes/src/nthash_base.rs:49-61)
constexpr uint64_t NT_MULT = 0x397f178c6ae330f9ULL;

GG_DEV uint64_t rotl1_64(uint64_t x) { return (x << 1) | (x >> 63); }
GG_DEV uint64_t rotr1_64(uint64_t x) { return (x >> 1) | (x << 63); }

// Windowed-minimum combine.  Items carry v = (canonical hash << 1) | unique,
// the unique bit being cleared whenever the minimum value occurs twice.  The
// operation is associative and reproduces BatchMinQueue::get_minimizers
// (crates/hashes/src/rolling/batch_minqueue.rs:37-124) for any reduction
// order: when both sides hold the same value the result keeps the value with
// the unique bit cleared; otherwise the strictly smaller side wins.
GG_DEV void mm_combine(uint64_t av, uint32_t ap, uint64_t bv, uint32_t bp, uint64_t& rv, uint32_t& rp) {
    if (av == bv) {
        rv = av & ~1ULL;
        rp = ap;
    } else if (bv < av) {
        rv = bv;
        rp = bp;
    } else {
        rv = av;
        rp = ap;
    }
}

// Multiplicative hash for LDS/global open-addressing tables.
// Value of the next lane (lane + 1) of the wavefront: a DPP wavefront shift
// (wave_shl:1); the last lane gets 0.
GG_DEV uint32_t lane_next32(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x130, 0xF, 0xF, false);
}
GG_DEV uint64_t lane_next64(uint64_t x) {
    return (uint64_t)lane_next32((uint32_t)x) | ((uint64_t)lane_next32((uint32_t)(x >> 32)) << 32);
}
// Value of the previous lane (lane - 1): DPP wave_shr:1; lane 0 gets 0.
GG_DEV uint32_t lane_prev32(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x138, 0xF, 0xF, false);
}
GG_DEV uint64_t lane_prev64(uint64_t x) {
    return (uint64_t)lane_prev32((uint32_t)x) | ((uint64_t)lane_prev32((uint32_t)(x >> 32)) << 32);
}

GG_DEV uint32_t tab_hash64(uint64_t x) {
    x ^= x >> 29;
    x *= 0xbf58476d1ce4e5b9ULL;
    x ^= x >> 32;
    return (uint32_t)x;
}
