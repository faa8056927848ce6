// synth.cpp -- seeded synthetic genomes and reads for the benchmark workloads.
//
// The reference has no data generator (its sample data sits on an external FTP
// site, README.md:36); these are the workloads the benchmarks are defined on:
// uniform or GC-biased genomes, repeat copies with divergence, multi-contig
// genomes joined by an N separator, and reads sampled uniformly on both strands
// with either a fixed number of errors per read or per-base error rates.
#include <array>
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "saruman_hip.h"

namespace {

// xoshiro256** (Blackman & Vigna), seeded through splitmix64
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed)
    {
        uint64_t z = seed;
        for (auto &x : s) {
            z += 0x9e3779b97f4a7c15ULL;
            uint64_t t = z;
            t = (t ^ (t >> 30)) * 0xbf58476d1ce4e5b9ULL;
            t = (t ^ (t >> 27)) * 0x94d049bb133111ebULL;
            x = t ^ (t >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next()
    {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    uint64_t below(uint64_t n) { return n ? next() % n : 0; }
};

const char kBases[4] = {'A', 'C', 'G', 'T'};

inline char base_gc(Rng &r, double gc)
{
    double u = r.uniform();
    if (u < gc) return (u < gc / 2) ? 'C' : 'G';
    return (u < gc + (1 - gc) / 2) ? 'A' : 'T';
}

inline char other_base(Rng &r, char c)
{
    char o;
    do {
        o = kBases[r.below(4)];
    } while (o == c);
    return o;
}

inline char comp(char c)
{
    // A<->T, C<->G, every other byte unchanged: one table load, no branch per base
    static constexpr auto t = [] {
        std::array<char, 256> a{};
        for (int i = 0; i < 256; ++i) a[(size_t)i] = (char)i;
        a['A'] = 'T';
        a['T'] = 'A';
        a['C'] = 'G';
        a['G'] = 'C';
        return a;
    }();
    return t[(uint8_t)c];
}

}  // namespace

extern "C" {

/* genome of len bases with GC fraction gc (0.5 = uniform ACGT) */
int srm_synth_genome(uint64_t seed, size_t len, double gc, char *out)
{
    const int T = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    size_t per = (len + T - 1) / T;
    for (int t = 0; t < T; ++t) {
        th.emplace_back([=] {
            Rng r(seed * 1000003ULL + (uint64_t)t);
            size_t b = (size_t)t * per, e = std::min(len, b + per);
            for (size_t i = b; i < e; ++i) out[i] = base_gc(r, gc);
        });
    }
    for (auto &x : th) x.join();
    return SRM_OK;
}

/*
 * Overwrites n_copies segments of the genome with diverged copies of n_families
 * repeat families (family length uniform in [fam_min, fam_max]).  Copy lengths
 * are drawn so the copies cover about cover_frac of the genome in total; each
 * copy is a random segment of its family with per-base substitution rate
 * divergence.  Returns the number of bases overwritten.
 */
long long srm_synth_repeats(uint64_t seed, char *genome, size_t len, int n_families, int fam_min,
                            int fam_max, int n_copies, double cover_frac, double divergence)
{
    Rng r(seed ^ 0x5eedULL);
    std::vector<std::string> fam;
    for (int f = 0; f < n_families; ++f) {
        int fl = fam_min + (int)r.below((uint64_t)(fam_max - fam_min + 1));
        std::string s((size_t)fl, 'A');
        for (auto &c : s) c = kBases[r.below(4)];
        fam.push_back(s);
    }
    double mean_copy = cover_frac * (double)len / std::max(1, n_copies);
    long long written = 0;
    for (int c = 0; c < n_copies; ++c) {
        const std::string &f = fam[r.below(fam.size())];
        int cl = (int)std::llround(mean_copy * (0.5 + r.uniform()));
        cl = std::max(50, std::min(cl, (int)f.size()));
        size_t fs = r.below(f.size() - (size_t)cl + 1);
        size_t gp = 1 + r.below(len - (size_t)cl - 1);
        for (int i = 0; i < cl; ++i) {
            char b = f[fs + (size_t)i];
            if (r.uniform() < divergence) b = other_base(r, b);
            genome[gp + (size_t)i] = b;
        }
        written += cl;
    }
    return written;
}

/*
 * Reads sampled from [lo, hi) ranges of the genome (contigs; a read never spans
 * two), read length uniform in [Lmin, Lmax], strand 50/50 (reverse reads are the
 * reverse complement of the sampled segment, errors applied before), start
 * uniform over the positions of the chosen contig that leave room for the read
 * and never 0 (the reference cannot report start 0, src/hash.c:448).
 *   mode 0: n_err uniform in [0, max_err] per read, each error a substitution
 *           (p_sub), 1-bp insertion (p_ins) or 1-bp deletion (rest);
 *   mode 1: per-base substitution rate p_sub, per-base indel rate p_ins (indel
 *           length uniform in 1..indel_max, insertions and deletions equally).
 * p_random: share of reads that are uniform random sequence.
 * Contigs are chosen with probability proportional to their length.
 * Outputs: reads concatenated into out (size sum of lengths), per read offset,
 * length, true start (-1 for random reads) and strand (0 fwd, 1 rev).
 */
int srm_synth_reads(uint64_t seed, const char *genome, const uint64_t *contig_lo,
                    const uint64_t *contig_hi, int n_contigs, size_t n_reads, int Lmin, int Lmax,
                    int mode, int max_err, double p_sub, double p_ins, int indel_max,
                    double p_random, char *out, uint64_t *off, uint16_t *len, int64_t *true_pos,
                    uint8_t *strand)
{
    if (Lmin < 1 || Lmax < Lmin || n_contigs < 1) return SRM_EINVAL;
    // lengths and offsets first (sequential, from their own stream)
    {
        Rng r(seed ^ 0x1e9ULL);
        uint64_t acc = 0;
        for (size_t i = 0; i < n_reads; ++i) {
            int L = Lmin + (int)r.below((uint64_t)(Lmax - Lmin + 1));
            len[i] = (uint16_t)L;
            off[i] = acc;
            acc += (uint64_t)L;
        }
    }
    std::vector<double> cum((size_t)n_contigs);
    double tot = 0;
    for (int c = 0; c < n_contigs; ++c) {
        tot += (double)(contig_hi[c] - contig_lo[c]);
        cum[(size_t)c] = tot;
    }
    const int T = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    const size_t per = (n_reads + T - 1) / T;
    const int slack = 3 * (indel_max + 2) + 8 + (mode == 0 ? max_err : 0) + 64;
    for (int t = 0; t < T; ++t) {
        th.emplace_back([=, &cum] {
            size_t b = (size_t)t * per, e = std::min(n_reads, b + per);
            std::string buf;
            for (size_t i = b; i < e; ++i) {
                Rng r(seed * 0x100000001b3ULL + i);
                const int L = len[i];
                char *o = out + off[i];
                if (r.uniform() < p_random) {
                    for (int j = 0; j < L; ++j) o[j] = kBases[r.below(4)];
                    true_pos[i] = -1;
                    strand[i] = (uint8_t)(r.below(2));
                    continue;
                }
                double u = r.uniform() * tot;
                int c = (int)(std::lower_bound(cum.begin(), cum.end(), u) - cum.begin());
                c = std::min(c, n_contigs - 1);
                uint64_t lo = std::max<uint64_t>(contig_lo[c], 1);
                uint64_t hi = contig_hi[c];
                uint64_t span = (uint64_t)L + (uint64_t)slack;
                if (hi < lo + span + 1) {         // contig too short: random read
                    for (int j = 0; j < L; ++j) o[j] = kBases[r.below(4)];
                    true_pos[i] = -1;
                    strand[i] = 0;
                    continue;
                }
                uint64_t start = lo + r.below(hi - span - lo);
                buf.clear();
                const char *g = genome + start;
                size_t gi = 0;
                if (mode == 0) {
                    int nerr = (int)r.below((uint64_t)max_err + 1);
                    buf.assign(g, (size_t)L + (size_t)max_err + 2);
                    for (int x = 0; x < nerr; ++x) {
                        size_t p = r.below((uint64_t)L);
                        double v = r.uniform();
                        if (v < p_sub)
                            buf[p] = other_base(r, buf[p]);
                        else if (v < p_sub + p_ins)
                            buf.insert(buf.begin() + (long)p, kBases[r.below(4)]);
                        else
                            buf.erase(buf.begin() + (long)p);
                    }
                } else {
                    const size_t gmax = (size_t)(hi - start);   // stay inside the contig
                    while ((int)buf.size() < L) {
                        if (gi >= gmax) {
                            buf.push_back(kBases[r.below(4)]);
                            continue;
                        }
                        double v = r.uniform();
                        if (v < p_sub) {
                            buf.push_back(other_base(r, g[gi++]));
                        } else if (v < p_sub + p_ins) {
                            int il = 1 + (int)r.below((uint64_t)indel_max);
                            if (r.below(2)) {
                                for (int x = 0; x < il; ++x) buf.push_back(kBases[r.below(4)]);
                            } else {
                                gi += (size_t)il;
                            }
                        } else {
                            buf.push_back(g[gi++]);
                        }
                    }
                }
                buf.resize((size_t)L);
                const bool rev = r.below(2) != 0;
                if (rev) {
                    for (int j = 0; j < L; ++j) o[j] = comp(buf[(size_t)(L - 1 - j)]);
                } else {
                    std::memcpy(o, buf.data(), (size_t)L);
                }
                true_pos[i] = (int64_t)start;
                strand[i] = rev ? 1 : 0;
            }
        });
    }
    for (auto &x : th) x.join();
    return SRM_OK;
}

}  // extern "C"
