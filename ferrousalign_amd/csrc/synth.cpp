// Seeded synthetic workload generator for the seed-extension path.
//
// PRNG: xoshiro256** (Blackman & Vigna) seeded through SplitMix64.  Every
// pair k draws from its own stream seeded with splitmix64(seed + k * phi), so
// the output is identical for any thread count.
//
// Pair model (right extension from a seed, the convention of
// src/pipelines/linear/batch_extension/collect_soa.rs:187-209 with h0 taken as
// seed_len * A as on the left side, :181):
//   reference : i.i.d. uniform ACGT
//   read      : copied from a uniform reference position with per-base
//               substitutions (another base, uniform) and per-base indel
//               events (length U{1,2,3}, insertion/deletion 50/50)
//   seed      : length L ~ U{19..60} at read offset s ~ U{0..R-L}
//   query     : read[s+L .. R)   (qlen == 0 is redrawn)
//   target    : reference from the coordinate of read[s+L], length qlen+100
//   h0        : L * A
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct Xoshiro {
    uint64_t s[4];
    explicit Xoshiro(uint64_t seed) {
        uint64_t x = seed;
        for (auto &v : s) v = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    inline uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    // uniform double in [0,1)
    inline double uniform() { return (next() >> 11) * 0x1.0p-53; }
    // uniform integer in [0, n)
    inline uint32_t below(uint32_t n) { return (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32); }
    inline double normal() {  // Box-Muller
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

inline uint64_t pair_seed(uint64_t seed, uint64_t k) {
    uint64_t x = seed + k * 0x9E3779B97F4A7C15ULL;
    return splitmix64(x);
}

// Simulate a read of length R starting at reference position p.  map[k] is
// the reference coordinate aligned to read base k (for inserted bases, the
// next reference base).
struct ReadModel {
    double sub_lo, sub_hi;  // substitution rate rises linearly lo -> hi along the read
    double indel;           // per-base indel event rate
};

void simulate_read(Xoshiro &rng, const uint8_t *ref, int64_t p, int R, const ReadModel &mdl,
                   uint8_t *read, int64_t *map) {
    int64_t r = p;
    int k = 0;
    while (k < R) {
        if (mdl.indel > 0 && rng.uniform() < mdl.indel) {
            int len = 1 + (int)rng.below(3);
            if (rng.below(2) == 0) {  // insertion into the read
                for (int x = 0; x < len && k < R; ++x) {
                    read[k] = (uint8_t)rng.below(4);
                    map[k] = r;
                    ++k;
                }
                continue;
            }
            r += len;  // deletion from the read
        }
        double rate = R > 1 ? mdl.sub_lo + (mdl.sub_hi - mdl.sub_lo) * (double)k / (double)(R - 1)
                            : mdl.sub_lo;
        uint8_t b = ref[r];
        if (rng.uniform() < rate) b = (uint8_t)((b + 1 + rng.below(3)) & 3);
        read[k] = b;
        map[k] = r;
        ++r;
        ++k;
    }
}

template <class F>
void parallel_for(int64_t n, int nthreads, F &&f) {
    if (nthreads <= 0) nthreads = (int)std::thread::hardware_concurrency();
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    if (n < 4096 || nthreads == 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    int64_t per = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        int64_t lo = per * t, hi = lo + per < n ? lo + per : n;
        if (lo >= hi) break;
        th.emplace_back([=, &f] { f(lo, hi); });
    }
    for (auto &x : th) x.join();
}

}  // namespace

extern "C" {

// Uniform i.i.d. ACGT reference of length len.
void fa_synth_reference(uint64_t seed, int64_t len, uint8_t *out) {
    Xoshiro rng(seed);
    for (int64_t i = 0; i < len; ++i) out[i] = (uint8_t)(rng.next() >> 62);
}

// Extension-pair workloads.
//   kind 0: 151 bp reads, 1% substitutions, 0.1% indels
//   kind 1: 150 bp reads, substitutions rising 0.2% -> 2%, 0.1% indels
//   kind 2: read length in {100,151,250} (weights .3/.5/.2), substitutions
//           rising 0.2% -> 2%, 0.1% indels; 5% of targets drawn from a
//           uniform reference position (non-homologous); 10% use w2.
// Each pair k writes qlen[k] codes at qseq + k*qstride and tlen[k] codes at
// tseq + k*tstride (qstride >= 250, tstride >= 350); the target is the reference
// window [tstart, tstart + tlen), tstart[k] written when tstart_out is not NULL.
// Returns 0, or -1 on bad arguments.
int fa_synth_extension_pairs(int kind, uint64_t seed, const uint8_t *ref, int64_t ref_len,
                             int64_t n, uint8_t *qseq, int64_t qstride, int32_t *qlen,
                             uint8_t *tseq, int64_t tstride, int32_t *tlen, int32_t *h0,
                             int32_t *w, int32_t match_score, int32_t w1, int32_t w2,
                             int nthreads, int64_t *tstart_out) {
    if (kind < 0 || kind > 2 || n < 0 || ref_len < 4096) return -1;
    if (qstride < 250 || tstride < 350) return -1;
    parallel_for(n, nthreads, [&](int64_t lo, int64_t hi) {
        uint8_t read[256];
        int64_t map[256];
        for (int64_t k = lo; k < hi; ++k) {
            Xoshiro rng(pair_seed(seed, (uint64_t)k));
            int R;
            ReadModel mdl;
            if (kind == 0) {
                R = 151;
                mdl = {0.01, 0.01, 0.001};
            } else if (kind == 1) {
                R = 150;
                mdl = {0.002, 0.02, 0.001};
            } else {
                double u = rng.uniform();
                R = u < 0.3 ? 100 : (u < 0.8 ? 151 : 250);
                mdl = {0.002, 0.02, 0.001};
            }
            int64_t p = (int64_t)(rng.uniform() * (double)(ref_len - R - 800));
            simulate_read(rng, ref, p, R, mdl, read, map);
            int L, s, ql;
            do {
                L = 19 + (int)rng.below(42);
                s = (int)rng.below((uint32_t)(R - L + 1));
                ql = R - s - L;
            } while (ql == 0);
            int tl = ql + 100;
            int64_t tstart = map[s + L];
            bool random_target = kind == 2 && rng.uniform() < 0.05;
            bool wide = kind == 2 && rng.uniform() < 0.10;
            if (random_target) tstart = (int64_t)(rng.uniform() * (double)(ref_len - tl - 1));
            std::memcpy(qseq + k * qstride, read + s + L, (size_t)ql);
            std::memcpy(tseq + k * tstride, ref + tstart, (size_t)tl);
            if (tstart_out) tstart_out[k] = tstart;
            qlen[k] = ql;
            tlen[k] = tl;
            h0[k] = L * match_score;
            w[k] = wide ? w2 : w1;
        }
    });
    return 0;
}

// Mate-rescue workload: a 151 bp query (1% substitutions, 0.1% indels)
// against a 1,500 bp reference window placed for an insert size drawn from
// N(400, 50) around an anchor; 20% of windows are drawn from a uniform
// position (no mate inside).  Query k at qseq + k*151, window at
// tseq + k*1500.  has_mate[k] = 1 when the window holds the mate.
int fa_synth_mate_pairs(uint64_t seed, const uint8_t *ref, int64_t ref_len, int64_t n,
                        uint8_t *qseq, uint8_t *tseq, uint8_t *has_mate, int nthreads) {
    const int R = 151, T = 1500;
    if (n < 0 || ref_len < 8192) return -1;
    parallel_for(n, nthreads, [&](int64_t lo, int64_t hi) {
        uint8_t read[256];
        int64_t map[256];
        ReadModel mdl = {0.01, 0.01, 0.001};
        for (int64_t k = lo; k < hi; ++k) {
            Xoshiro rng(pair_seed(seed, (uint64_t)k));
            int64_t anchor = 2000 + (int64_t)(rng.uniform() * (double)(ref_len - 6000));
            double ins = 400.0 + 50.0 * rng.normal();
            if (ins < 151.0) ins = 151.0;
            if (ins > 1100.0) ins = 1100.0;
            int64_t mate = anchor + (int64_t)ins - R;
            int64_t wstart = anchor + 400 - T / 2;
            bool nomate = rng.uniform() < 0.20;
            if (nomate) wstart = (int64_t)(rng.uniform() * (double)(ref_len - T - 1));
            simulate_read(rng, ref, mate, R, mdl, read, map);
            std::memcpy(qseq + k * R, read, R);
            std::memcpy(tseq + k * T, ref + wstart, T);
            has_mate[k] = nomate ? 0 : 1;
        }
    });
    return 0;
}

// CIGAR-regeneration workload (the pairs region.rs:1058 aligns): a 150 bp
// read (substitutions rising 0.2% -> 2%, 0.1% indels) against the reference
// span it maps to, [rb, re) of its alignment region (the aligner starts at the
// region's first base); band w = |tlen - qlen| + 10.  Query k at
// qseq + k*qstride (>= 150), window at tseq + k*tstride (>= 180).
int fa_synth_region_pairs(uint64_t seed, const uint8_t *ref, int64_t ref_len, int64_t n,
                          uint8_t *qseq, int64_t qstride, int32_t *qlen, uint8_t *tseq,
                          int64_t tstride, int32_t *tlen, int32_t *w, int nthreads) {
    const int R = 150;
    if (n < 0 || ref_len < 8192 || qstride < R || tstride < R + 30) return -1;
    parallel_for(n, nthreads, [&](int64_t lo, int64_t hi) {
        uint8_t read[256];
        int64_t map[256];
        ReadModel mdl = {0.002, 0.02, 0.001};
        for (int64_t k = lo; k < hi; ++k) {
            Xoshiro rng(pair_seed(seed, (uint64_t)k));
            const int64_t p = 16 + (int64_t)(rng.uniform() * (double)(ref_len - R - 64));
            simulate_read(rng, ref, p, R, mdl, read, map);
            const int64_t a = map[0];
            int64_t b = map[R - 1] + 1;
            if (b - a > tstride) b = a + tstride;
            std::memcpy(qseq + k * qstride, read, R);
            std::memcpy(tseq + k * tstride, ref + a, (size_t)(b - a));
            qlen[k] = R;
            tlen[k] = (int32_t)(b - a);
            w[k] = (tlen[k] > R ? tlen[k] - R : R - tlen[k]) + 10;
        }
    });
    return 0;
}

// Chain-extension workload (seeds and chains in, regions out): a read of R
// bases (substitutions rising 0.2% -> 2%, 0.1% indels) from a uniform
// position on a uniform strand of the bidirectional reference [0, 2 l_pac)
// (p >= l_pac: reverse complement of forward base 2 l_pac - 1 - p).  Its
// seeds are the maximal exact runs of at least kmin bases on the true
// alignment (at most max_seeds, in query order, which is also reference
// order), in bidirectional coordinates.  Half the reads carry a decoy
// seed of U{kmin..kmin+11} bases at a uniform position (decoy_len 0: none).
// Read k: query at qseq + k*R, seeds at [k*max_seeds, +nseeds[k]).
int fa_synth_chain_reads(uint64_t seed, const uint8_t *ref, int64_t l_pac, int64_t n, int32_t R,
                         int32_t kmin, int32_t max_seeds, uint8_t *qseq, int64_t *seed_rp,
                         int32_t *seed_qp, int32_t *seed_len, int32_t *nseeds, int64_t *decoy_rp,
                         int32_t *decoy_qp, int32_t *decoy_len, int64_t *true_pos, int nthreads) {
    if (n < 0 || R < 2 * kmin || R > 250 || kmin < 1 || max_seeds < 1 || l_pac < 8192) return -1;
    parallel_for(n, nthreads, [&](int64_t lo, int64_t hi) {
        uint8_t read[256];
        int64_t map[256];
        int32_t rs[256], re[256];
        ReadModel mdl = {0.002, 0.02, 0.001};
        for (int64_t k = lo; k < hi; ++k) {
            Xoshiro rng(pair_seed(seed, (uint64_t)k));
            const int64_t p = 64 + (int64_t)(rng.uniform() * (double)(l_pac - R - 256));
            const bool rev = rng.below(2) == 1;
            simulate_read(rng, ref, p, R, mdl, read, map);
            // maximal exact runs on the true alignment, forward read coordinates
            int nr = 0, s = 0;
            for (int j = 0; j <= R; ++j) {
                const bool ext = j < R && read[j] == ref[map[j]] && (j == s || map[j] == map[j - 1] + 1);
                if (ext) continue;
                if (j - s >= kmin) {
                    rs[nr] = s;
                    re[nr] = j;
                    ++nr;
                }
                s = (j < R && read[j] == ref[map[j]]) ? j : j + 1;
            }
            uint8_t *q = qseq + k * R;
            int32_t ns = 0;
            if (!rev) {
                std::memcpy(q, read, (size_t)R);
                for (int x = 0; x < nr && ns < max_seeds; ++x, ++ns) {
                    seed_rp[k * max_seeds + ns] = map[rs[x]];
                    seed_qp[k * max_seeds + ns] = rs[x];
                    seed_len[k * max_seeds + ns] = re[x] - rs[x];
                }
                true_pos[k] = map[0];
            } else {
                for (int j = 0; j < R; ++j) q[j] = (uint8_t)(3 - read[R - 1 - j]);
                for (int x = nr - 1; x >= 0 && ns < max_seeds; --x, ++ns) {
                    seed_rp[k * max_seeds + ns] = 2 * l_pac - 1 - map[re[x] - 1];
                    seed_qp[k * max_seeds + ns] = R - re[x];
                    seed_len[k * max_seeds + ns] = re[x] - rs[x];
                }
                true_pos[k] = 2 * l_pac - 1 - map[R - 1];
            }
            nseeds[k] = ns;
            if (rng.below(2) == 0) {
                const int32_t len = kmin + (int32_t)rng.below(12);
                decoy_len[k] = len;
                decoy_qp[k] = (int32_t)rng.below((uint32_t)(R - len + 1));
                decoy_rp[k] = (int64_t)(rng.uniform() * (double)(2 * l_pac - len));
            } else {
                decoy_len[k] = 0;
                decoy_qp[k] = 0;
                decoy_rp[k] = 0;
            }
        }
    });
    return 0;
}

}  // extern "C"
