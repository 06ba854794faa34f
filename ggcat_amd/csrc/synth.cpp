// synth.cpp -- seeded synthetic genomes / reads / colored genome sets used by
// the benchmark and the tests.  Counter-based RNG (splitmix64 over
// (seed, stream, counter)) so that every read is reproducible independently
// of thread count.
//
// Output format for read sets: the "joined" layout consumed by the bucketing
// entry points -- every record followed by one '\n' byte.  A separator byte is
// not an A/C/G/T and therefore ends an ACGT run exactly like an 'N' does in
// the reference splitter (crates/minimizer_bucketing/src/sequences_splitter.rs:26-35),
// so record boundaries need no extra index.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

inline uint64_t splitmix(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

struct Rng {
    uint64_t s;
    Rng(uint64_t seed, uint64_t stream) : s(splitmix(seed * 0x100000001B3ULL ^ splitmix(stream))) {}
    uint64_t next() {
        s += 0x9E3779B97F4A7C15ULL;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    uint64_t below(uint64_t n) { return n ? (uint64_t)(((unsigned __int128)next() * n) >> 64) : 0; }
    double normal() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

const uint8_t BASES[4] = {'A', 'C', 'G', 'T'};

inline uint8_t comp(uint8_t c) {
    switch (c) {
        case 'A': return 'T';
        case 'C': return 'G';
        case 'G': return 'C';
        case 'T': return 'A';
    }
    return 'N';
}

inline uint8_t substitute(uint8_t c, Rng& r) {
    // one of the three other bases, uniformly
    int idx = (c == 'A') ? 0 : (c == 'C') ? 1 : (c == 'G') ? 2 : 3;
    int o = (int)r.below(3);
    return BASES[(idx + 1 + o) & 3];
}

template <class F>
void parallel_for(uint64_t n, int threads, F&& f) {
    if (threads < 1) threads = 1;
    std::vector<std::thread> th;
    for (int t = 0; t < threads; t++) {
        uint64_t a = n * t / threads, b = n * (t + 1) / threads;
        th.emplace_back([=, &f] { f(a, b); });
    }
    for (auto& x : th) x.join();
}

}  // namespace

extern "C" {

// i.i.d. uniform ACGT genome; when repeat_frac > 0, segments of uniform
// length in [rep_min, rep_max] are copied to random destinations with a
// per-copy divergence uniform in [0, max_div] until repeat_frac * G bases
// were overwritten by copies.
void synth_genome(uint64_t seed, uint64_t G, double repeat_frac, uint32_t rep_min, uint32_t rep_max,
                  double max_div, uint8_t* out, int threads) {
    parallel_for(G, threads, [&](uint64_t a, uint64_t b) {
        const uint64_t CH = 1 << 16;
        for (uint64_t c0 = a / CH * CH; c0 < b; c0 += CH) {
            Rng r(seed, c0 / CH);
            for (uint64_t i = c0; i < c0 + CH && i < G; i++) {
                uint8_t v = BASES[r.next() >> 62];
                if (i >= a && i < b) out[i] = v;
            }
        }
    });
    if (repeat_frac <= 0.0 || G < rep_max + 1) return;
    Rng r(seed, 0xC0FFEEULL);
    uint64_t copied = 0, target = (uint64_t)(repeat_frac * (double)G);
    std::vector<uint8_t> seg;
    while (copied < target) {
        uint32_t L = rep_min + (uint32_t)r.below(rep_max - rep_min + 1);
        uint64_t s = r.below(G - L), d = r.below(G - L);
        double div = r.uniform() * max_div;
        seg.assign(out + s, out + s + L);
        for (uint32_t i = 0; i < L; i++) {
            uint8_t c = seg[i];
            if (r.uniform() < div) c = substitute(c, r);
            out[d + i] = c;
        }
        copied += L;
    }
}

// n_reads reads of read_len bases sampled uniformly from genome[0..G),
// reverse-complemented with probability 1/2, each base substituted with
// probability err.  Written in the joined layout: read_len bases + '\n'.
void synth_reads(const uint8_t* genome, uint64_t G, uint64_t seed, uint64_t n_reads, uint32_t read_len,
                 double err, uint8_t* out, int threads) {
    const uint64_t stride = (uint64_t)read_len + 1;
    parallel_for(n_reads, threads, [&](uint64_t a, uint64_t b) {
        for (uint64_t i = a; i < b; i++) {
            Rng r(seed, i);
            uint64_t st = r.below(G - read_len + 1);
            bool rc = (r.next() >> 63) != 0;
            uint8_t* o = out + i * stride;
            if (!rc) {
                memcpy(o, genome + st, read_len);
            } else {
                for (uint32_t j = 0; j < read_len; j++) o[j] = comp(genome[st + read_len - 1 - j]);
            }
            if (err > 0) {
                for (uint32_t j = 0; j < read_len; j++)
                    if (r.uniform() < err) o[j] = substitute(o[j], r);
            }
            o[read_len] = '\n';
        }
    });
}

// One mutated descendant of `anc` (length A): SNP rate snp, indel rate indel
// (length uniform 1..50, insertion or deletion with equal probability).  The
// descendant is then cut or extended to target_len: shorter targets keep a
// random contiguous window, longer ones append mutated ancestor segments.
// Returns the produced length (<= cap).
uint64_t synth_descendant(const uint8_t* anc, uint64_t A, uint64_t seed, uint64_t stream, double snp, double indel,
                          uint64_t target_len, uint8_t* out, uint64_t cap) {
    Rng r(seed, stream);
    std::vector<uint8_t> g;
    g.reserve(A + A / 10);
    auto mutate_into = [&](const uint8_t* src, uint64_t n) {
        for (uint64_t i = 0; i < n; i++) {
            double u = r.uniform();
            if (u < indel) {
                uint32_t L = 1 + (uint32_t)r.below(50);
                if (r.next() & 1) {
                    for (uint32_t j = 0; j < L; j++) g.push_back(BASES[r.next() >> 62]);
                    g.push_back(src[i]);
                } else {
                    i += L - 1;  // deletion of L bases (src[i] included)
                }
            } else if (u < indel + snp) {
                g.push_back(substitute(src[i], r));
            } else {
                g.push_back(src[i]);
            }
        }
    };
    mutate_into(anc, A);
    while (g.size() < target_len) {
        uint64_t L = std::min<uint64_t>(A / 4 + 1, target_len - g.size() + 1000);
        uint64_t s = r.below(A - L + 1);
        mutate_into(anc + s, L);
    }
    uint64_t n = std::min<uint64_t>(g.size(), target_len);
    uint64_t st = (g.size() > n) ? r.below(g.size() - n + 1) : 0;
    n = std::min(n, cap);
    memcpy(out, g.data() + st, n);
    return n;
}

// log-normal length in [lo, hi] (clipped), median `median`, shape sigma.
uint64_t synth_lognormal_len(uint64_t seed, uint64_t stream, double median, double sigma, double lo, double hi) {
    Rng r(seed, stream ^ 0xABCDEF);
    double v = median * std::exp(sigma * r.normal());
    v = std::max(lo, std::min(hi, v));
    return (uint64_t)v;
}

// Query set: n queries, lengths log-uniform in [lmin, lmax]; half are
// substrings of the given genomes (chosen uniformly) with substitution rate
// mut, half i.i.d. ACGT.  Joined layout ('\n' after each query).  Returns the
// number of bytes written; lengths are written to out_len.
uint64_t synth_queries(const uint8_t* genomes, const uint64_t* g_off, uint64_t n_genomes, uint64_t seed,
                       uint64_t stream, uint64_t n, uint32_t lmin, uint32_t lmax, double mut, uint8_t* out,
                       uint32_t* out_len) {
    Rng r(seed, stream);
    uint64_t pos = 0;
    double la = std::log((double)lmin), lb = std::log((double)lmax);
    for (uint64_t q = 0; q < n; q++) {
        uint32_t L = (uint32_t)std::exp(la + (lb - la) * r.uniform());
        L = std::max(lmin, std::min(lmax, L));
        bool from_genome = (r.next() >> 63) != 0;
        if (from_genome && n_genomes > 0) {
            uint64_t gi = r.below(n_genomes);
            uint64_t glen = g_off[gi + 1] - g_off[gi];
            if (glen < L) L = (uint32_t)glen;
            uint64_t st = g_off[gi] + r.below(glen - L + 1);
            for (uint32_t j = 0; j < L; j++) {
                uint8_t c = genomes[st + j];
                if (r.uniform() < mut) c = substitute(c, r);
                out[pos + j] = c;
            }
        } else {
            for (uint32_t j = 0; j < L; j++) out[pos + j] = BASES[r.next() >> 62];
        }
        out[pos + L] = '\n';
        out_len[q] = L;
        pos += (uint64_t)L + 1;
    }
    return pos;
}

}  // extern "C"
