// filter_pair.h -- the per-read part of the q-gram filter as host/device code:
// q-gram lookup in the device copy of the index, run_seashore's pairing
// (src/hash.c:502-834) and add_start_position (src/hash.c:40-178) over the flat
// position-list array.  Compiled into the kernels of filter_gpu.hip; the same
// functions also compile for the host, which lets a CPU build check the
// kernels' indexing.
#pragma once

#include <stdint.h>

#include <vector>

#include "filter_internal.h"

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define SRM_HD __host__ __device__
#else
#define SRM_HD
#endif

namespace srm_pair {

constexpr uint32_t OVERFLOW_MARK = 0xFFFFFFFFu;

struct DQ {                  // one q's index on the device
    const uint64_t *keys;    // (code << pos_bits) | offset, ascending
    const uint32_t *pre;     // [4^pre_bases + 1] key ranges by the first bases
    const uint64_t *odd_hash;   // FNV-1a of the q-gram text, ascending, then offset
    const uint32_t *odd_pos;
    uint32_t n_odd;
    int q, pos_bits, pre_bases;
};

struct DParam {
    int L, k, q, rest, n, half, exact;
    uint32_t gram_off;       // into FArgs::gram
};

struct FArgs {
    const DQ *qix;           // [32], indexed by q
    const DParam *par;
    const uint16_t *gram;
    const char *genome;
    uint64_t glen;
    const char *reads;
    const uint64_t *roff;
    const uint16_t *rlen;
    const uint16_t *pid;
    uint32_t n_reads;
    int dr;                  // 0 forward, 4 reverse
    uint32_t cv;             // chunk_value
    const uint32_t *list_base;   // [n_reads + 1] first list of every read
    uint32_t *list_cnt;          // [n_lists + 1] list length + 1 (terminator)
    uint32_t *list_lo;           // [n_lists] first key of the list (NO_RANGE: search again)
    const uint32_t *list_start;  // [n_lists + 1] exclusive sum of list_cnt
    uint32_t *flat;
    // pairing
    const uint32_t *todo;    // reads to pair; null = all reads
    uint32_t n_todo;
    int32_t *slot_pos;       // HBM slots (second and later passes)
    int8_t *slot_flag;
    int32_t *slot_perf;
    uint32_t cap, pcap;
    uint32_t *counters;      // [2] compact entries / perfect hits written so far
    int32_t *out_pos;
    int8_t *out_flag;
    int32_t *out_perf;
    uint32_t *meta;          // [3 * n_todo]: ent_off, perf_off, n_ent | n_perf << 16
};

SRM_HD inline char dcomp(char c)
{
    switch (c) {
    case 'A': return 'T';
    case 'C': return 'G';
    case 'G': return 'C';
    case 'T': return 'A';
    default: return c;
    }
}

SRM_HD inline int dcode(char c)
{
    switch (c) {
    case 'A': return 0;
    case 'C': return 1;
    case 'G': return 2;
    case 'T': return 3;
    default: return -1;
    }
}

// byte j of the read as the filter sees it (reverse complement for strand 4,
// util.c:35-57)
SRM_HD inline char obyte(const char *rd, int L, int dr, int j)
{
    return dr ? dcomp(rd[L - 1 - j]) : rd[j];
}

// FNV-1a, 64 bit
SRM_HD inline uint64_t fnv_step(uint64_t h, unsigned char c)
{
    return (h ^ c) * 1099511628211ull;
}
constexpr uint64_t FNV_BASIS = 14695981039346656037ull;

// list_lo of a q-gram whose positions are not one run of keys (it holds a byte
// other than ACGT: its positions come from the hash table and a byte compare)
constexpr uint32_t NO_RANGE = 0xFFFFFFFFu;

// positions of the q-gram at read offset g, ascending; writes them to out when
// FILL, returns their number; lo_out (when given) receives the index of the first
// key (the positions are keys lo .. lo + count - 1), or NO_RANGE for a q-gram
// holding another byte
template <bool FILL>
SRM_HD uint32_t lookup(const DQ &X, const char *genome, const char *rd, int L, int dr, int g,
                       uint32_t *out, uint32_t *lo_out = nullptr)
{
    const int q = X.q;
    uint64_t code = 0;
    bool odd = false;
    // all q bytes, with no exit from the loop (on gfx950 a loop that left at the
    // first other byte gave fl_count wrong list lengths for some lanes once the key
    // range was stored as well; the same source on the host gave the right ones)
    for (int i = 0; i < q; ++i) {
        const int c = dcode(obyte(rd, L, dr, g + i));
        odd |= c < 0;
        code = (code << 2) | (uint64_t)(c & 3);
    }
    uint32_t cnt = 0;
    if (odd) {
        if (lo_out) *lo_out = NO_RANGE;
        uint64_t h = FNV_BASIS;
        for (int i = 0; i < q; ++i) h = fnv_step(h, (unsigned char)obyte(rd, L, dr, g + i));
        uint32_t lo = 0, hi = X.n_odd;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (X.odd_hash[mid] < h) lo = mid + 1;
            else hi = mid;
        }
        for (uint32_t i = lo; i < X.n_odd && X.odd_hash[i] == h; ++i) {
            const uint32_t p = X.odd_pos[i];
            bool same = true;
            for (int t = 0; t < q && same; ++t) same = genome[p + t] == obyte(rd, L, dr, g + t);
            if (same) {
                if (FILL) out[cnt] = p;
                ++cnt;
            }
        }
        return cnt;
    }
    const uint64_t p = code >> (2 * (q - X.pre_bases));
    uint32_t lo = X.pre[p], hi = X.pre[p + 1];
    const uint64_t key = code << X.pos_bits;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (X.keys[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    if (lo_out) *lo_out = lo;
    const uint64_t mask = ((uint64_t)1 << X.pos_bits) - 1;
    const uint32_t end = X.pre[p + 1];
    for (uint32_t i = lo; i < end && (X.keys[i] >> X.pos_bits) == code; ++i) {
        if (FILL) out[cnt] = (uint32_t)(X.keys[i] & mask);
        ++cnt;
    }
    return cnt;
}

// a read's result-hash entries (pos, flag) and printed perfect hits
struct Store {
    int32_t *pos;
    int8_t *flag;
    int32_t *perf;
    uint32_t stride, cap, pcap;
    uint32_t n = 0, np = 0;
    bool ovf = false;
    SRM_HD int32_t &P(uint32_t i) { return pos[i * stride]; }
    SRM_HD int8_t &F(uint32_t i) { return flag[i * stride]; }
    SRM_HD int32_t &Q(uint32_t i) { return perf[i * stride]; }
    SRM_HD void push(int32_t p, int f)
    {
        if (n == cap) {
            ovf = true;
            return;
        }
        P(n) = p;
        F(n) = (int8_t)f;
        ++n;
    }
};

// first index of the ascending run flat[lo, hi) whose value is >= t
SRM_HD inline uint64_t run_lower(const uint32_t *flat, uint64_t lo, uint64_t hi, long t)
{
    if (t <= 0) return lo;
    if (t > 0xFFFFFFFFl) return hi;
    const uint32_t u = (uint32_t)t;
    while (lo < hi) {
        const uint64_t m = (lo + hi) >> 1;
        if (flat[m] < u) lo = m + 1;
        else hi = m;
    }
    return lo;
}

// The pair passes' matches of one list in the reference's scan order.  The scan
// compares every target tg(err), err = -lim .. lim, with the ascending run
// flat[lo, min(hi, end)) -- cut at its first 0, as the reference's 0-terminated
// scans are -- and, in the first pass, first with one more value p0 (p0 == 0:
// none).  The targets are consecutive (ascending from t_lo when up, descending
// from t_hi otherwise), so the matches are the run's values in [t_lo, t_hi] found
// by one search and a walk, p0 merged in before a run value of the same target.
// Counts them into *hits and returns true when the count reaches need.
SRM_HD inline bool scan_run(const uint32_t *flat, uint64_t lo, uint64_t hi, uint64_t end,
                            long t_lo, long t_hi, bool up, uint32_t p0, int need, int &hits)
{
    if (hi > end) hi = end;
    if (lo >= hi || flat[lo] == 0) hi = lo;    // an empty run
    bool p0_left = p0 != 0 && (long)p0 >= t_lo && (long)p0 <= t_hi;
    if (up) {
        uint64_t i = run_lower(flat, lo, hi, t_lo);
        for (;;) {
            const bool run_next = i < hi && (long)flat[i] <= t_hi;
            if (p0_left && (!run_next || (long)p0 <= (long)flat[i])) {
                p0_left = false;
                if (++hits == need) return true;
                continue;
            }
            if (!run_next) return false;
            if (++hits == need) return true;
            ++i;
        }
    }
    // descending (second pass only: no p0)
    uint64_t i = run_lower(flat, lo, hi, t_hi + 1);
    while (i > lo && (long)flat[i - 1] >= t_lo) {
        if (++hits == need) return true;
        --i;
    }
    return false;
}

// add_start_position, hash.c:40-178: 0 first entry, 1 rejected, 2 added or
// flag lowered.  |pos - e| <= k is the reference's scan of pos + i, i in [-k, k].
SRM_HD inline int add_start(Store &s, int32_t pos, int flag, int possible, int exact, int k)
{
    if (possible != exact) return 1;
    const int dir = flag > 3 ? 1 : 0;
    if (s.n == 0) {
        s.push(pos, flag);
        return 0;
    }
    for (uint32_t e = 0; e < s.n; ++e) {
        const int32_t ep = s.P(e);
        const int ef = s.F(e);
        const int tdir = ef > 3 ? 1 : 0;
        const int64_t d1 = (int64_t)ep - pos, d2 = (int64_t)ep + pos;
        if (ef <= flag && ((d1 >= -k && d1 <= k) || (d2 >= -k && d2 <= k))) return 1;
        if (ep == pos && tdir == dir && ef > flag) {
            s.F(e) = (int8_t)flag;
            return 2;
        } else if ((ep == pos || ep == -pos) && tdir == dir) {
            return 1;
        }
    }
    s.push(pos, flag);
    return 2;
}

SRM_HD inline void seashore(const FArgs &a, uint32_t r, Store &s)
{
    const DParam P = a.par[a.pid[r]];
    const int q = P.q, half = P.half, k = P.k, L = P.L, rest = P.rest, dr = a.dr;
    const char *rd = a.reads + a.roff[r];
    const uint32_t lb = a.list_base[r];
    const uint32_t next_chunk = (r / a.cv + 1) * a.cv;
    const uint32_t hi_read = next_chunk < a.n_reads ? next_chunk : a.n_reads;
    const uint64_t flat_end = a.list_start[a.list_base[hi_read]];
    auto D = [&](int x, uint64_t idx) -> uint32_t {
        const uint64_t p = (uint64_t)a.list_start[lb + x] + idx;
        return p < flat_end ? a.flat[p] : 0u;
    };
    for (int aa = 0; aa < half; ++aa) {
        const bool fqm = (aa == 0);
        const long in_read = (long)aa * q + 1;
        uint64_t off = 0;
        uint32_t Pa = D(aa, off);
        while (Pa > 0) {
            const long gs = (long)Pa - in_read;
            if (gs + 1 > 0) {
                // perfect check (strncpy + strcmp, hash.c:550-583)
                const long S = gs + 1;
                if ((uint64_t)S + (uint64_t)L <= a.glen) {
                    bool perfect = true;
                    const char *g = a.genome + S;
                    for (int j = 0; j < L && perfect; ++j) perfect = g[j] == obyte(rd, L, dr, j);
                    if (perfect && add_start(s, (int32_t)(-S), 0, P.exact, P.exact, k) != 1) {
                        if (s.np == s.pcap) s.ovf = true;
                        else s.Q(s.np++) = (int32_t)S;
                    }
                    if (s.ovf) return;
                }
                // the pair passes, hash.c:606-794.  Every match raises `possible` and
                // calls add_start_position, which acts only when possible equals
                // exact_tresh (Q13): the need-th match of the two passes in scan order,
                // need = exact_tresh - 1, is the only one that can insert; the scan stops
                // there.
                const int need = P.exact - 1;
                int hits = 0;
                bool found = need < 1;
                int32_t hpos = 0;
                int hflag = 0;
                // first pass, hash.c:606-692.  For every err the reference scans
                // D(b, off), then D(b, 1), D(b, 2), ... up to the next 0 (Q12): the
                // element at a's index, then one ascending run of list b (or of the
                // list after b when b is empty).
                for (int b = aa + 1; b < half && !found; ++b) {
                    const uint32_t P0 = D(b, off);
                    if (P0 == 0) continue;                 // the scan never starts
                    const long rstart = (long)b * q + 1;
                    uint64_t lo = (uint64_t)a.list_start[lb + b] + 1, hi = lo;
                    const uint64_t nb = a.list_start[lb + b + 1];
                    if (lo < nb) {
                        hi = nb - 1;                       // list b from index 1
                    } else if ((uint64_t)lb + b + 2 <= a.list_base[a.n_reads]) {
                        hi = (uint64_t)a.list_start[lb + b + 2] - 1;   // the next list
                    }
                    if (scan_run(a.flat, lo, hi, flat_end, gs + rstart - k, gs + rstart + k, true,
                                 P0, need, hits)) {
                        found = true;
                        if (b == half - 1 && fqm) {
                            hflag = 1 + dr;
                            hpos = (int32_t)(gs + 1);
                        } else if (aa == half - 1) {
                            hflag = ((long)Pa + q > L ? 3 : 2) + dr;
                            hpos = (int32_t)(gs - k + 1);
                        } else if (fqm) {
                            hflag = 1 + dr;
                            hpos = (int32_t)(gs + 1);
                        } else {
                            hflag = 3 + dr;
                            hpos = (int32_t)(gs - k + 1);
                        }
                    }
                }
                // second pass, hash.c:695-794: each err scans list half + j from its
                // start up to its 0; the targets gs + nin + err run downwards when
                // j < a (the reference's err is negated there)
                for (int j = 0; j < half && !found; ++j) {
                    const int lim = half - 1 - j - aa < 0 ? -(half - 1 - j - aa) : half - 1 - j - aa;
                    const long nin = (long)(half - j - 1) * q + 1 + rest;
                    const uint64_t lo = a.list_start[lb + half + j];
                    const uint64_t hi = (uint64_t)a.list_start[lb + half + j + 1] - 1;
                    if (scan_run(a.flat, lo, hi, flat_end, gs + nin - lim, gs + nin + lim, j - aa >= 0,
                                 0u, need, hits)) {
                        found = true;
                        if (j == 0 && fqm) {
                            hflag = 1 + dr;
                            hpos = (int32_t)(gs + 1);
                        } else if (j == 0) {
                            hflag = (nin + q > L ? 3 : 2) + dr;
                            hpos = (int32_t)(gs - k + 1);
                        } else if (fqm) {
                            hflag = 1 + dr;
                            hpos = (int32_t)(gs + 1);
                        } else {
                            hflag = 3 + dr;
                            hpos = (int32_t)(gs - k + 1);
                        }
                    }
                }
                if (found && need >= 1) {
                    add_start(s, hpos, hflag, P.exact, P.exact, k);
                    if (s.ovf) return;
                }
            }
            ++off;
            Pa = D(aa, off);
        }
    }
}

// the device copies of the per-read parameter sets and their q-gram offsets
inline void device_params(const std::vector<srm_filter_impl::Params> &params,
                          std::vector<DParam> &dpar, std::vector<uint16_t> &gram)
{
    for (auto &p : params) {
        dpar.push_back({p.L, p.k, p.q, p.rest, p.n, p.half, p.exact_tresh, (uint32_t)gram.size()});
        for (int g : p.gram) gram.push_back((uint16_t)g);
    }
}

// the device copy of one q's index: the odd (non-ACGT) q-grams as (FNV-1a of
// the text, offset) pairs in ascending order
inline void odd_table(const srm_filter_impl::QIndex &X, std::vector<uint64_t> &hash,
                      std::vector<uint32_t> &pos)
{
    std::vector<std::pair<uint64_t, uint32_t>> odd;
    for (auto &kv : X.odd) {
        uint64_t h = FNV_BASIS;
        for (char c : kv.first) h = fnv_step(h, (unsigned char)c);
        for (uint32_t p : kv.second) odd.emplace_back(h, p);
    }
    std::sort(odd.begin(), odd.end());
    hash.resize(odd.size());
    pos.resize(odd.size());
    for (size_t t = 0; t < odd.size(); ++t) {
        hash[t] = odd[t].first;
        pos[t] = odd[t].second;
    }
}

}  // namespace srm_pair
