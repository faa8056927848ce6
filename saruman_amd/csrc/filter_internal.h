// filter_internal.h -- pieces of the q-gram filter shared by the host filter
// (filter.cpp) and the device filter (filter_gpu.hip): the q-gram index, the
// per-read parameters, the result object and the sequential batch pass.
#pragma once

#include <array>
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "saruman_hip.h"

namespace srm_filter_impl {

inline int code_of(char c)
{
    switch (c) {
    case 'A': return 0;
    case 'C': return 1;
    case 'G': return 2;
    case 'T': return 3;
    default: return -1;
    }
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

// Index of every q-gram of the genome at offsets 0..len-q (util.c:456-499), in
// ascending offset order per q-gram.  ACGT-only q-grams are 2-bit codes in one
// sorted array of (code << pos_bits) | offset keys with a prefix table over the
// first pre_bases bases; q-grams holding any other byte are kept by their text.
struct QIndex {
    int q = 0;
    int pos_bits = 0;
    std::vector<uint64_t> keys;            // (code << pos_bits) | pos, ascending
    int pre_bases = 0;
    std::vector<uint32_t> pre;             // offsets by the first pre_bases bases
    std::unordered_map<std::string, std::vector<uint32_t>> odd;   // q-grams with non-ACGT

    void lookup(const char *g, std::vector<uint32_t> &out) const
    {
        uint64_t code = 0;
        for (int i = 0; i < q; ++i) {
            int c = code_of(g[i]);
            if (c < 0) {
                auto it = odd.find(std::string(g, (size_t)q));
                if (it != odd.end()) out.insert(out.end(), it->second.begin(), it->second.end());
                return;
            }
            code = (code << 2) | (uint64_t)c;
        }
        uint64_t p = code >> (2 * (q - pre_bases));
        size_t lo = pre[p], hi = pre[p + 1];
        const uint64_t mask = ((uint64_t)1 << pos_bits) - 1;
        auto first = std::lower_bound(keys.begin() + (long)lo, keys.begin() + (long)hi,
                                      code << pos_bits);
        for (auto it = first; it != keys.begin() + (long)hi && ((*it) >> pos_bits) == code; ++it)
            out.push_back((uint32_t)(*it & mask));
    }
};

// per-read parameters, saruman.c:42-75 and :119
struct Params {
    int L = 0, k = 0, q = 0, rest = 0, n = 0, half = 0, exact_tresh = 0;
    std::vector<int> gram;   // q-gram offsets in the read
};

// float arithmetic as in the reference
inline Params derive(int L, int k)
{
    Params p;
    p.L = L;
    p.k = k;
    p.q = (int)std::ceil((float)L / (float)(k + 1)) - 1;
    if (p.q < 1) p.q = 1;
    p.rest = L % p.q;
    p.n = L / p.q;
    if (p.rest != 0) p.n *= 2;
    for (int i = 0; i < L / p.q; ++i) p.gram.push_back(i * p.q);
    for (int z = 1; z <= L / p.q; ++z) p.gram.push_back(L - z * p.q);
    p.gram.resize((size_t)p.n);
    p.half = p.n / 2;
    p.exact_tresh = (int)std::floor((float)L / (float)p.q) - k + 1;
    return p;
}

}  // namespace srm_filter_impl

struct srm_index {
    std::vector<srm_filter_impl::QIndex> idx;
    size_t glen = 0;
    uint32_t base = 0;   // genome position of offset 0 (a -G chunk's start)
    const srm_filter_impl::QIndex *get(int q) const
    {
        for (auto &x : idx)
            if (x.q == q) return &x;
        return nullptr;
    }
};

struct srm_filter {
    std::vector<int32_t> c_read;
    std::vector<int64_t> c_pos;
    std::vector<uint8_t> c_flag;
    std::vector<uint16_t> c_w;
    std::vector<uint64_t> batch_end;
    std::vector<int32_t> p_read;
    std::vector<int64_t> p_pos;
};

namespace srm_filter_impl {

// The batch builder (hash.c:841-851 cuda_queue_check, hash.c:898-1111
// prepare_cuda_alignments) over the per-read filter results.  Acc gives, for read
// r: n_ent(r), pos_at(r, i), flag_at(r, i) -- the result-hash entries of the read
// in insertion order --, n_perf(r), perf_at(r, i) -- the offsets of the perfect hits
// the filter printed -- and queued(r), the entries added with pos > 0
// (hash.c:87-89, 168-170).
//
// A batch is flushed after the read whose queued count takes the queue past
// chunk_value, and once at the end.  prepare_cuda_alignments walks every read of
// the result hash in insertion (= read) order.  Queued positions are negated,
// perfect ones are negative, so only entries of reads since the last flush and
// entries skipped as out of the genome (never negated, Q10) are positive; every
// other entry just resets the s/e state.  `live` marks positive entries;
// `sticky` lists older reads that still hold one.
template <class Acc>
srm_filter *batch_pass(const Acc &ro, size_t n_reads, const uint16_t *read_len,
                       const uint8_t *read_k, size_t glen, const uint8_t *blocked,
                       int chunk_value)
{
    srm_filter *F = new srm_filter();
    std::vector<size_t> next_with(n_reads + 1, n_reads);   // next read with entries
    for (size_t r = n_reads; r-- > 0;)
        next_with[r] = (r + 1 < n_reads && ro.n_ent(r + 1) != 0) ? r + 1 : next_with[r + 1];
    // live flag of entry i of read r at live[base[r] + i]
    std::vector<size_t> base(n_reads + 1, 0);
    size_t n_perf_all = 0;
    for (size_t r = 0; r < n_reads; ++r) {
        base[r + 1] = base[r] + ro.n_ent(r);
        n_perf_all += ro.n_perf(r);
    }
    std::vector<uint8_t> live(base[n_reads], 0);
    // an entry becomes a candidate at most once: room for all of them up front
    F->c_read.reserve(base[n_reads]);
    F->c_pos.reserve(base[n_reads]);
    F->c_flag.reserve(base[n_reads]);
    F->c_w.reserve(base[n_reads]);
    F->p_read.reserve(n_perf_all);
    F->p_pos.reserve(n_perf_all);
    std::vector<size_t> sticky;
    uint64_t queue = 0;
    size_t pending_lo = 0;
    auto flush = [&](size_t lo, size_t hi) {
        std::vector<size_t> order(sticky);
        for (size_t r = lo; r < hi; ++r) {
            const size_t ne = ro.n_ent(r);
            if (ne == 0) continue;
            for (size_t i = 0; i < ne; ++i) live[base[r] + i] = ro.pos_at(r, i) > 0;
            order.push_back(r);
        }
        int start_offset = 0, end_offset = 0;
        size_t before = F->c_read.size();
        size_t prev = (size_t)-1;
        std::vector<size_t> still;
        for (size_t r : order) {
            if (prev != (size_t)-1 && next_with[prev] != r) {
                start_offset = 0;   // an intervening read's non-positive entry resets
                end_offset = 0;
            }
            prev = r;
            const int L = read_len[r], k = read_k[r];
            bool keep = false;
            const size_t ne = ro.n_ent(r);
            for (size_t i = 0; i < ne; ++i) {
                if (live[base[r] + i]) {
                    const int32_t pos = ro.pos_at(r, i);
                    const int flag = ro.flag_at(r, i);
                    switch (flag) {   // hash.c:938-979
                    case 1: case 5: end_offset = k; break;
                    case 2: case 6: start_offset = k; break;
                    case 3: case 7: start_offset = k; end_offset = k; break;
                    default: break;
                    }
                    const int64_t w = (int64_t)L + start_offset + end_offset;
                    bool skip = (uint64_t)pos + (uint64_t)w > (uint64_t)glen;   // :982
                    if (!skip && blocked)
                        for (int64_t j = 0; j < w && !skip; ++j) skip = blocked[pos + j] != 0;
                    if (skip) {            // `continue`: no reset, stays positive (Q10)
                        keep = true;
                        continue;
                    }
                    F->c_read.push_back((int32_t)r);
                    F->c_pos.push_back(pos);
                    F->c_flag.push_back((uint8_t)flag);
                    F->c_w.push_back((uint16_t)w);
                    live[base[r] + i] = 0;   // :1077 negated
                }
                end_offset = 0;
                start_offset = 0;
            }
            if (keep) still.push_back(r);
        }
        sticky.swap(still);
        if (F->c_read.size() > before) F->batch_end.push_back(F->c_read.size());
    };
    for (size_t r = 0; r < n_reads; ++r) {
        const size_t np = ro.n_perf(r);
        for (size_t i = 0; i < np; ++i) {
            F->p_read.push_back((int32_t)r);
            F->p_pos.push_back(ro.perf_at(r, i));
        }
        queue += (uint64_t)ro.queued(r);
        if (queue > (uint64_t)chunk_value) {
            flush(pending_lo, r + 1);
            pending_lo = r + 1;
            queue = 0;
        }
    }
    if (queue > 0) flush(pending_lo, n_reads);
    return F;
}

// Per-call read plan of the device filter: the parameters of every distinct
// (length, k), each read's parameter set and the index of its first q-gram list.
struct ReadPlan {
    std::vector<Params> params;
    std::vector<uint16_t> pid;          // per read
    std::vector<uint32_t> list_base;    // per read + 1: first list, then the total
    uint64_t read_bytes = 0;            // extent of the reads buffer used
};

// returns SRM_OK or an SRM_E* code with a message in err
inline int plan_reads(const srm_index *ix, const uint64_t *read_off, const uint16_t *read_len,
                      const uint8_t *read_k, size_t n_reads, ReadPlan &P, std::string &err)
{
    // parameter set of each (length, k): a table over the lengths present and k
    // 0..255, filled on first use
    int maxlen = 0;
    for (size_t i = 0; i < n_reads; ++i) maxlen = std::max(maxlen, (int)read_len[i]);
    std::vector<int32_t> slot(((size_t)maxlen + 1) * 256, -1);
    P.pid.resize(n_reads);
    P.list_base.resize(n_reads + 1);
    uint64_t nl = 0;
    for (size_t i = 0; i < n_reads; ++i) {
        const int L = read_len[i], k = read_k[i];
        int32_t &id = slot[(size_t)L * 256 + (size_t)k];
        if (id < 0) {
            if (L < 1) {
                err = "a read has length 0";
                return SRM_EINVAL;
            }
            P.params.push_back(derive(L, k));
            if (!ix->get(P.params.back().q)) {
                err = "the index lacks q = " + std::to_string(P.params.back().q);
                return SRM_EINVAL;
            }
            if (P.params.size() > 65535) {
                err = "too many (length, k) classes";
                return SRM_EINVAL;
            }
            id = (int32_t)P.params.size() - 1;
        }
        P.pid[i] = (uint16_t)id;
        P.list_base[i] = (uint32_t)nl;
        nl += (uint64_t)P.params[(size_t)id].n;
        P.read_bytes = std::max<uint64_t>(P.read_bytes, read_off[i] + read_len[i]);
        if (nl >= ((uint64_t)1 << 31)) {
            err = "too many q-gram lists in one call";
            return SRM_EINVAL;
        }
    }
    P.list_base[n_reads] = (uint32_t)nl;
    return SRM_OK;
}

// Per-read filter results held as meta words (3 per read: entry offset, perfect
// offset, n_entries | n_perfect << 16) over compact entry and perfect arrays;
// the accessor batch_pass reads.
struct MetaAcc {
    const std::vector<uint32_t> &meta;
    const std::vector<int32_t> &pos;
    const std::vector<int8_t> &flag;
    const std::vector<int32_t> &perf;
    size_t n_ent(size_t r) const { return meta[3 * r + 2] & 0xFFFFu; }
    int32_t pos_at(size_t r, size_t i) const { return pos[meta[3 * r] + i]; }
    int flag_at(size_t r, size_t i) const { return flag[meta[3 * r] + i]; }
    size_t n_perf(size_t r) const { return meta[3 * r + 2] >> 16; }
    int32_t perf_at(size_t r, size_t i) const { return perf[meta[3 * r + 1] + i]; }
    // entries are only ever added or have their flag lowered, so the count of
    // additions with pos > 0 is the count of such entries
    int queued(size_t r) const
    {
        int c = 0;
        const size_t n = n_ent(r);
        for (size_t i = 0; i < n; ++i) c += pos_at(r, i) > 0;
        return c;
    }
};

}  // namespace srm_filter_impl
