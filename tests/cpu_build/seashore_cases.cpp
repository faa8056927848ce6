// Runs the product's two implementations of SARUMAN's per-read filter on inputs
// given literally on stdin, for the hand-worked cases of tests/test_filter_cases.py:
//   - the host filter's add_start and seashore_read (saruman_amd/csrc/filter.cpp,
//     compiled into this program so that its internal functions are reachable);
//   - the device filter's add_start and seashore (saruman_amd/csrc/filter_pair.h,
//     the code the kernels of filter_gpu.hip run, compiled for the host).
//
// Input (whitespace-separated tokens):
//   dedup <k> <exact_tresh> <n_calls>  then n_calls x  <pos> <flag> <possible>
//   seashore <L> <k> <do_reverse> <n_reads> <chunk_value>
//       <genome_len> <genome text>
//       n_reads x <read text, as in the read file (forward)>
//       n_reads x n lists: <count> <position>...   (the data_array of the chunk:
//       each list is followed by its 0 terminator in the array this builds)
// Output, one line per implementation (and read):
//   <impl> rets=<r,...> ent=<pos:flag,...> queued=<n>          (dedup)
//   <impl> read=<r> ent=<pos:flag,...> perf=<S,...> queued=<n>  (seashore)
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

#include "filter.cpp"
#include "filter_pair.h"

namespace {

// the data_array of the reads' chunks, as filter.cpp's seashore_read reads it: a
// read's lists run on into the next reads' lists up to the end of its chunk of
// chunk_value reads, and read as 0 beyond (hash.c:377-461 builds one zero-padded
// array per chunk)
struct FlatLists {
    const char *reads;
    const uint64_t *roff;
    const std::vector<uint32_t> *flat;
    const std::vector<uint32_t> *list_start;   // per list, then the end
    const std::vector<uint32_t> *list_base;    // per read, then the total
    size_t n_reads, chunk_value;
    uint32_t D(size_t id, int x, long idx) const
    {
        const size_t hi = std::min(n_reads, (id / chunk_value + 1) * chunk_value);
        const uint64_t end = (*list_start)[(*list_base)[hi]];
        const uint64_t p = (uint64_t)(*list_start)[(*list_base)[id] + (size_t)x] + (uint64_t)idx;
        return p < end ? (*flat)[p] : 0u;
    }
};

std::string ents(const std::vector<std::pair<int32_t, int>> &e)
{
    std::string s;
    for (size_t i = 0; i < e.size(); ++i)
        s += (i ? "," : "") + std::to_string(e[i].first) + ":" + std::to_string(e[i].second);
    return s;
}

std::string ints(const std::vector<int> &v)
{
    std::string s;
    for (size_t i = 0; i < v.size(); ++i) s += (i ? "," : "") + std::to_string(v[i]);
    return s;
}

int dedup()
{
    int k, exact, n;
    std::cin >> k >> exact >> n;
    std::vector<int> pos((size_t)n), flag((size_t)n), possible((size_t)n);
    for (int i = 0; i < n; ++i) std::cin >> pos[(size_t)i] >> flag[(size_t)i] >> possible[(size_t)i];
    // host filter
    {
        ReadOut ro;
        std::vector<int> rets;
        for (int i = 0; i < n; ++i)
            rets.push_back(add_start(ro, pos[(size_t)i], flag[(size_t)i], possible[(size_t)i], exact, k));
        std::vector<std::pair<int32_t, int>> e;
        for (auto &x : ro.ent) e.emplace_back(x.pos, (int)x.flag);
        printf("host rets=%s ent=%s queued=%d\n", ints(rets).c_str(), ents(e).c_str(), ro.queued);
    }
    // device filter code
    {
        std::vector<int32_t> sp(1024), sq(16);
        std::vector<int8_t> sf(1024);
        srm_pair::Store s;
        s.pos = sp.data();
        s.flag = sf.data();
        s.perf = sq.data();
        s.stride = 1;
        s.cap = 1024;
        s.pcap = 16;
        std::vector<int> rets;
        for (int i = 0; i < n; ++i)
            rets.push_back(srm_pair::add_start(s, pos[(size_t)i], flag[(size_t)i], possible[(size_t)i],
                                               exact, k));
        std::vector<std::pair<int32_t, int>> e;
        int queued = 0;
        for (uint32_t i = 0; i < s.n; ++i) {
            e.emplace_back(s.P(i), (int)s.F(i));
            queued += s.P(i) > 0;
        }
        printf("device rets=%s ent=%s queued=%d\n", ints(rets).c_str(), ents(e).c_str(), queued);
    }
    return 0;
}

int seashore_cases()
{
    int L, k, dr, nr, cv;
    size_t glen;
    std::string genome;
    std::cin >> L >> k >> dr >> nr >> cv >> glen >> genome;
    if (cv < 1) {
        fprintf(stderr, "chunk_value must be at least 1\n");
        return 2;
    }
    if (genome.size() != glen) {
        fprintf(stderr, "genome length %zu, expected %zu\n", genome.size(), glen);
        return 2;
    }
    std::string reads;
    std::vector<uint64_t> roff;
    std::vector<uint16_t> rlen;
    for (int r = 0; r < nr; ++r) {
        std::string rd;
        std::cin >> rd;
        if ((int)rd.size() != L) {
            fprintf(stderr, "read %d has length %zu, expected %d\n", r, rd.size(), L);
            return 2;
        }
        roff.push_back(reads.size());
        rlen.push_back((uint16_t)L);
        reads += rd;
    }
    const Params P = derive(L, k);
    std::vector<uint32_t> flat, list_start, list_base;
    for (int r = 0; r < nr; ++r) {
        list_base.push_back((uint32_t)list_start.size());
        for (int x = 0; x < P.n; ++x) {
            list_start.push_back((uint32_t)flat.size());
            int c;
            std::cin >> c;
            for (int i = 0; i < c; ++i) {
                long v;
                std::cin >> v;
                flat.push_back((uint32_t)v);
            }
            flat.push_back(0);
        }
    }
    list_base.push_back((uint32_t)list_start.size());
    list_start.push_back((uint32_t)flat.size());
    if (!std::cin) {
        fprintf(stderr, "malformed input\n");
        return 2;
    }
    // host filter
    FlatLists ls{reads.data(), roff.data(), &flat, &list_start, &list_base, (size_t)nr,
                 (size_t)cv};
    for (int r = 0; r < nr; ++r) {
        ReadOut ro;
        seashore_read(ls, (size_t)r, P, genome.data(), genome.size(), 0, dr, ro);
        std::vector<std::pair<int32_t, int>> e;
        for (auto &x : ro.ent) e.emplace_back(x.pos, (int)x.flag);
        std::vector<int> pf(ro.perfect.begin(), ro.perfect.end());
        printf("host read=%d ent=%s perf=%s queued=%d\n", r, ents(e).c_str(), ints(pf).c_str(),
               ro.queued);
    }
    // device filter code
    std::vector<srm_pair::DParam> dpar;
    std::vector<uint16_t> gram;
    srm_pair::device_params({P}, dpar, gram);
    std::vector<uint16_t> pid((size_t)nr, 0);
    srm_pair::FArgs a{};
    a.par = dpar.data();
    a.gram = gram.data();
    a.genome = genome.data();
    a.glen = genome.size();
    a.reads = reads.data();
    a.roff = roff.data();
    a.rlen = rlen.data();
    a.pid = pid.data();
    a.n_reads = (uint32_t)nr;
    a.dr = dr;
    a.cv = (uint32_t)cv;
    a.list_base = list_base.data();
    a.list_start = list_start.data();
    a.flat = flat.data();
    for (int r = 0; r < nr; ++r) {
        std::vector<int32_t> sp(1024), sq(64);
        std::vector<int8_t> sf(1024);
        srm_pair::Store s;
        s.pos = sp.data();
        s.flag = sf.data();
        s.perf = sq.data();
        s.stride = 1;
        s.cap = 1024;
        s.pcap = 64;
        srm_pair::seashore(a, (uint32_t)r, s);
        if (s.ovf) {
            fprintf(stderr, "slot overflow\n");
            return 2;
        }
        std::vector<std::pair<int32_t, int>> e;
        int queued = 0;
        for (uint32_t i = 0; i < s.n; ++i) {
            e.emplace_back(s.P(i), (int)s.F(i));
            queued += s.P(i) > 0;
        }
        std::vector<int> pf;
        for (uint32_t i = 0; i < s.np; ++i) pf.push_back(s.Q(i));
        printf("device read=%d ent=%s perf=%s queued=%d\n", r, ents(e).c_str(), ints(pf).c_str(),
               queued);
    }
    return 0;
}

}  // namespace

int main()
{
    std::string mode;
    std::cin >> mode;
    if (mode == "dedup") return dedup();
    if (mode == "seashore") return seashore_cases();
    fprintf(stderr, "unknown mode %s\n", mode.c_str());
    return 2;
}
