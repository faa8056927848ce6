// CPU build of the device filter's per-read code (saruman_amd/csrc/filter_pair.h:
// q-gram lookup, run_seashore pairing, add_start_position), driven the way the
// kernels of filter_gpu.hip drive it -- count, exclusive scan, fill, pairing with
// 16/4-entry slots, then 1024 and 65534-entry slots for the reads that overflow --
// and compared with the host filter (srm_filter_run) on seeded workloads.  Built
// with AddressSanitizer by tests/test_filter_pair_cpu.py, so every read and write
// of the kernels' indexing is checked against its buffer on the host.
//
//   filter_pair_cpu <seed> <genome_len> <n_reads> <Lmin> <Lmax> <chunk> <repeats 0/1> <odd 0/1>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "filter_internal.h"
#include "filter_pair.h"

using namespace srm_filter_impl;
using namespace srm_pair;

namespace {

struct Emu {
    const srm_index *ix;
    const std::string &genome;
    std::vector<DQ> dq;
    std::vector<std::vector<uint64_t>> oh;
    std::vector<std::vector<uint32_t>> op;
    uint64_t overflow = 0;

    Emu(const srm_index *ix_, const std::string &g) : ix(ix_), genome(g), dq(32)
    {
        std::memset(dq.data(), 0, dq.size() * sizeof(DQ));
        oh.resize(ix->idx.size());
        op.resize(ix->idx.size());
        for (size_t i = 0; i < ix->idx.size(); ++i) {
            const QIndex &X = ix->idx[i];
            odd_table(X, oh[i], op[i]);
            DQ &d = dq[(size_t)X.q];
            d.keys = X.keys.data();
            d.pre = X.pre.data();
            d.odd_hash = oh[i].data();
            d.odd_pos = op[i].data();
            d.n_odd = (uint32_t)oh[i].size();
            d.q = X.q;
            d.pos_bits = X.pos_bits;
            d.pre_bases = X.pre_bases;
        }
    }

    // one pairing pass over `todo`, compacted per read as the kernel's meta words
    void pass(FArgs a, const std::vector<uint32_t> &todo, uint32_t cap, uint32_t pcap,
              std::vector<uint32_t> &meta, std::vector<int32_t> &pos, std::vector<int8_t> &flag,
              std::vector<int32_t> &perf)
    {
        meta.assign(todo.size() * 3, 0);
        pos.clear();
        flag.clear();
        perf.clear();
        for (size_t j = 0; j < todo.size(); ++j) {
            // exactly-sized slots: an access past cap is a heap overflow
            std::vector<int32_t> sp(cap), sq(pcap);
            std::vector<int8_t> sf(cap);
            Store s;
            s.pos = sp.data();
            s.flag = sf.data();
            s.perf = sq.data();
            s.stride = 1;
            s.cap = cap;
            s.pcap = pcap;
            seashore(a, todo[j], s);
            meta[3 * j + 0] = (uint32_t)pos.size();
            meta[3 * j + 1] = (uint32_t)perf.size();
            if (s.ovf) {
                meta[3 * j + 2] = OVERFLOW_MARK;
                continue;
            }
            meta[3 * j + 2] = s.n | (s.np << 16);
            for (uint32_t e = 0; e < s.n; ++e) {
                pos.push_back(s.P(e));
                flag.push_back(s.F(e));
            }
            for (uint32_t e = 0; e < s.np; ++e) perf.push_back(s.Q(e));
        }
    }

    srm_filter *run(const uint8_t *blocked, const char *reads, const uint64_t *roff,
                    const uint16_t *rlen, const uint8_t *rk, size_t n, int dr, int cv)
    {
        ReadPlan plan;
        std::string err;
        if (plan_reads(ix, roff, rlen, rk, n, plan, err)) {
            fprintf(stderr, "plan: %s\n", err.c_str());
            exit(2);
        }
        std::vector<DParam> dpar;
        std::vector<uint16_t> gram;
        device_params(plan.params, dpar, gram);
        const uint64_t nl = plan.list_base[n];
        std::vector<uint32_t> list_cnt(nl + 1), list_start(nl + 1), list_lo(nl + 1);
        FArgs a{};
        a.qix = dq.data();
        a.par = dpar.data();
        a.gram = gram.data();
        a.genome = genome.data();
        a.glen = genome.size();
        a.reads = reads;
        a.roff = roff;
        a.rlen = rlen;
        a.pid = plan.pid.data();
        a.n_reads = (uint32_t)n;
        a.dr = dr;
        a.cv = (uint32_t)cv;
        a.list_base = plan.list_base.data();
        a.list_cnt = list_cnt.data();
        a.list_lo = list_lo.data();
        a.list_start = list_start.data();
        // count (fl_count): the list lengths and key ranges, checked against lookup
        for (uint32_t r = 0; r < n; ++r) {
            const DParam P = a.par[a.pid[r]];
            const DQ X = a.qix[P.q];
            for (int x = 0; x < P.n; ++x) {
                uint32_t lo;
                const uint32_t c = lookup<false>(X, a.genome, reads + roff[r], P.L, dr,
                                                 a.gram[P.gram_off + x], nullptr, &lo);
                if (c != lookup<false>(X, a.genome, reads + roff[r], P.L, dr,
                                       a.gram[P.gram_off + x], nullptr)) {
                    fprintf(stderr, "lookup count with the key range differs (read %u list %d)\n", r, x);
                    exit(1);
                }
                list_cnt[a.list_base[r] + x] = c + 1;
                list_lo[a.list_base[r] + x] = lo;
            }
        }
        list_cnt[nl] = 0;
        std::exclusive_scan(list_cnt.begin(), list_cnt.end(), list_start.begin(), 0u);
        std::vector<uint32_t> flat(list_start[nl]);   // exactly sized
        a.flat = flat.data();
        // fill (fl_fill)
        for (uint32_t r = 0; r < n; ++r) {
            const DParam P = a.par[a.pid[r]];
            const DQ X = a.qix[P.q];
            const uint64_t mask = ((uint64_t)1 << X.pos_bits) - 1;
            for (int x = 0; x < P.n; ++x) {
                uint32_t *dst = a.flat + list_start[a.list_base[r] + x];
                const uint32_t c = lookup<true>(X, a.genome, reads + roff[r], P.L, dr,
                                                a.gram[P.gram_off + x], dst);
                // fl_fill copies the keys of the range fl_count found: the same positions
                const uint32_t lo = list_lo[a.list_base[r] + x];
                if (lo != NO_RANGE)
                    for (uint32_t i = 0; i < c; ++i)
                        if ((uint32_t)(X.keys[lo + i] & mask) != dst[i]) {
                            fprintf(stderr, "key range differs (read %u list %d)\n", r, x);
                            exit(1);
                        }
                dst[c] = 0;
            }
        }
        // pairing (fl_pair), then the overflow passes of run_device
        std::vector<uint32_t> all(n);
        std::iota(all.begin(), all.end(), 0u);
        std::vector<uint32_t> meta;
        std::vector<int32_t> pos, perf;
        std::vector<int8_t> flag;
        pass(a, all, 16, 4, meta, pos, flag, perf);
        std::vector<uint32_t> over;
        for (size_t r = 0; r < n; ++r)
            if (meta[3 * r + 2] == OVERFLOW_MARK) over.push_back((uint32_t)r);
        overflow += over.size();
        for (uint32_t cap : {1024u, 65534u}) {
            if (over.empty()) break;
            std::vector<uint32_t> m2, still;
            std::vector<int32_t> p2, q2;
            std::vector<int8_t> f2;
            pass(a, over, cap, cap, m2, p2, f2, q2);
            const uint32_t es = (uint32_t)pos.size(), ps = (uint32_t)perf.size();
            pos.insert(pos.end(), p2.begin(), p2.end());
            flag.insert(flag.end(), f2.begin(), f2.end());
            perf.insert(perf.end(), q2.begin(), q2.end());
            for (size_t t = 0; t < over.size(); ++t) {
                const uint32_t r = over[t];
                if (m2[3 * t + 2] == OVERFLOW_MARK) {
                    still.push_back(r);
                    continue;
                }
                meta[3 * r + 0] = es + m2[3 * t + 0];
                meta[3 * r + 1] = ps + m2[3 * t + 1];
                meta[3 * r + 2] = m2[3 * t + 2];
            }
            over.swap(still);
        }
        if (!over.empty()) {
            fprintf(stderr, "read %u overflows every slot size\n", over[0]);
            exit(2);
        }
        return batch_pass(MetaAcc{meta, pos, flag, perf}, n, rlen, rk, genome.size(), blocked,
                          cv);
    }
};

template <class T>
bool same(const std::vector<T> &a, const std::vector<T> &b, const char *what, int dr)
{
    if (a == b) return true;
    size_t i = 0;
    while (i < a.size() && i < b.size() && a[i] == b[i]) ++i;
    fprintf(stderr, "strand %d: %s differ (sizes %zu / %zu, first difference at %zu)\n", dr, what,
            a.size(), b.size(), i);
    return false;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc != 9) {
        fprintf(stderr, "usage: %s seed glen nreads Lmin Lmax chunk repeats odd\n", argv[0]);
        return 2;
    }
    const uint64_t seed = strtoull(argv[1], nullptr, 0);
    const size_t glen = strtoull(argv[2], nullptr, 0);
    const size_t nreads = strtoull(argv[3], nullptr, 0);
    const int Lmin = atoi(argv[4]), Lmax = atoi(argv[5]), chunk = atoi(argv[6]);
    const bool repeats = atoi(argv[7]) != 0, odd = atoi(argv[8]) != 0;

    std::string g(glen, 'A');
    srm_synth_genome(seed, glen, 0.5, &g[0]);
    if (repeats) srm_synth_repeats(seed + 1, &g[0], glen, 3, 300, 600, 40, 0.5, 0.001);
    uint64_t x = seed * 0x9E3779B97F4A7C15ull + 1;
    auto next = [&] {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        return x;
    };
    if (odd) {
        for (size_t i = 0; i < glen / 400; ++i) g[next() % glen] = 'N';
        for (int t = 0; t < 5; ++t) {
            const size_t s = next() % (glen - 30);
            for (size_t i = 0; i < 25; ++i) g[s + i] = 'N';
        }
        for (size_t i = 0; i < glen / 1500; ++i) g[next() % glen] = 'R';
    }
    const uint64_t lo = 0, hi = glen;
    std::vector<char> rb(nreads * (size_t)Lmax);
    std::vector<uint64_t> off(nreads);
    std::vector<uint16_t> len(nreads);
    std::vector<int64_t> tp(nreads);
    std::vector<uint8_t> st(nreads), k(nreads);
    srm_synth_reads(seed + 2, g.data(), &lo, &hi, 1, nreads, Lmin, Lmax, Lmin == Lmax ? 0 : 1,
                    Lmin == Lmax ? 3 : 0, Lmin == Lmax ? 0.8 : 0.01, Lmin == Lmax ? 0.1 : 0.001, 2,
                    0.02, rb.data(), off.data(), len.data(), tp.data(), st.data());
    for (size_t i = 0; i < nreads; ++i) {
        k[i] = Lmin == Lmax ? 3 : (uint8_t)((len[i] + 19) / 20);   // ceil(5% L)
        if (odd && next() % 10 == 0) rb[off[i] + next() % len[i]] = 'N';
    }
    std::vector<int> qs;
    for (size_t i = 0; i < nreads; ++i) {
        int q, rest, n, et;
        srm_filter_params(len[i], k[i], &q, &rest, &n, &et);
        if (std::find(qs.begin(), qs.end(), q) == qs.end()) qs.push_back(q);
    }
    srm_index *ix = nullptr;
    if (srm_index_build(g.data(), glen, qs.data(), (int)qs.size(), 0, &ix) != SRM_OK) return 2;
    Emu emu(ix, g);
    bool ok = true;
    size_t total = 0;
    for (int dr : {0, 4}) {
        srm_filter *H = nullptr;
        if (srm_filter_run(ix, g.data(), glen, nullptr, rb.data(), off.data(), len.data(), k.data(),
                           nreads, dr, chunk, 0, &H) != SRM_OK)
            return 2;
        srm_filter *E = emu.run(nullptr, rb.data(), off.data(), len.data(), k.data(), nreads, dr,
                                chunk);
        ok &= same(H->c_read, E->c_read, "candidate reads", dr);
        ok &= same(H->c_pos, E->c_pos, "candidate positions", dr);
        ok &= same(H->c_flag, E->c_flag, "candidate flags", dr);
        ok &= same(H->c_w, E->c_w, "window lengths", dr);
        ok &= same(H->batch_end, E->batch_end, "batch ends", dr);
        ok &= same(H->p_read, E->p_read, "perfect reads", dr);
        ok &= same(H->p_pos, E->p_pos, "perfect positions", dr);
        total += H->c_read.size();
        srm_filter_destroy(H);
        delete E;
    }
    srm_index_destroy(ix);
    printf("%s candidates=%zu overflow_reads=%llu\n", ok ? "OK" : "MISMATCH", total,
           (unsigned long long)emu.overflow);
    return ok ? 0 : 1;
}
