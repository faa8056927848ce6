// records.cpp -- host side of the output path: rebuilds the two alignment rows the
// reference's kernel leaves in a candidate's slot (get_alignment,
// src/cuda-functions.cu:108-315) from the compact transcript, and formats the TSV
// record of the reference's post-processing (src/cuda-functions.cu:828-994).
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "saruman_hip.h"

namespace {

inline unsigned op_at(const uint32_t *ops, int j) { return (ops[j >> 4] >> ((j & 15) * 2)) & 3u; }

// byte i of a slot of capacity cap; outside the slot reads as NUL (the zero
// padding the reference's pool memset leaves after every slot)
inline char at(const char *X, int cap, int i) { return (i >= 0 && i < cap) ? X[i] : '\0'; }

}  // namespace

extern "C" int srm_slot_rows(const srm_result *res, const uint32_t *ops, const char *read,
                             const char *win, char *R, char *G, int cap)
{
    if (cap < 0) return -1;
    std::memset(R, 0, (size_t)cap);
    std::memset(G, 0, (size_t)cap);
    const int n = res->n_ops;
    if (n > cap) return -1;
    // the transcript is in traceback order; replaying it backwards walks the
    // alignment from its first column, which is slot position 0 after the
    // reference's in-place reversal (cuda-functions.cu:306-307)
    int row = res->end_row, col = res->end_col;
    for (int p = 0; p < n; ++p) {
        unsigned op = op_at(ops, n - 1 - p);
        if (op == SRM_OP_DIAG) {
            R[p] = read[col];
            G[p] = win[row];
            ++row;
            ++col;
        } else if (op == SRM_OP_TOP) {
            R[p] = '_';
            G[p] = win[row];
            ++row;
        } else if (op == SRM_OP_LEFT) {
            R[p] = read[col];
            G[p] = '_';
            ++col;
        } else {
            return -1;
        }
    }
    return n;
}

extern "C" int srm_format_record(const char *name, unsigned int pos, int flag, int k, int L, int w,
                                 const srm_result *res, const uint32_t *ops, const char *read,
                                 const char *win, char *out, int outcap)
{
    const int cap = w + k;
    if (cap <= 0 || cap > SRM_MAX_W + 256) return -1;
    char R[SRM_MAX_W + 256], G[SRM_MAX_W + 256];
    if (srm_slot_rows(res, ops, read, win, R, G, cap) < 0) return -1;

    // opening gaps: leading '_' of the read row shift the start (:878-886)
    unsigned short lead = 0;
    while (at(R, cap, lead) == '_') ++lead;
    const unsigned int start = pos + lead;

    // closing gaps: '_' among the trailing NUL/'_' run of the slot (:897-907)
    unsigned short trail = 0;
    for (int i = cap - 1; i >= 0; --i) {
        char x = at(R, cap, i);
        if (x == '_')
            ++trail;
        else if (x != '\0')
            break;
    }

    // indel scan over the aligned part (:910-959)
    unsigned short del = 0, ins = 0;
    int i = 0, z = 0;
    short end_of_read = 0;
    while (i < w - trail - lead + z) {
        char x = at(R, cap, lead + i);
        if (x == '\0') {
            end_of_read = (short)i;
            break;
        }
        if (x == '_')
            ++del, ++z;
        else if (at(G, cap, lead + i) == '_')
            ++ins, ++z;
        ++i;
    }
    // characters printed (:972-980); the reference takes the global read length
    // here, the candidate's own L is the same for uniform batches (Q8)
    const short chars = end_of_read > 0 ? (short)(end_of_read - trail) : (short)(L + del);

    // %.*s stops at the first NUL (:987-994, Q9)
    // (a negative precision means no precision; it cannot occur for a walk that
    // emitted a read character, but is kept printf-exact)
    const int lim = chars < 0 ? cap : chars;
    int rn = 0, gn = 0;
    while (rn < lim && at(R, cap, lead + rn) != '\0') ++rn;
    while (gn < lim && at(G, cap, lead + gn) != '\0') ++gn;
    const unsigned int stop = (unsigned)chars - ins + start - 1;
    return std::snprintf(out, (size_t)(outcap > 0 ? outcap : 0), "%s\t%d\t%d\t%s\t%.*s\t%.*s\t%d\n",
                         name, (int)start, (int)stop, flag > 3 ? "<<" : ">>", rn, R + lead, gn,
                         G + lead, (int)res->score);
}

// ---------------------------------------------------------------------------
// bulk formatting
// ---------------------------------------------------------------------------

/*
 * Records of every hit (score <= k) of n candidates, in candidate order, into
 * out (capacity outcap).  Windows are taken from the ASCII genome at pos[t];
 * reads from reads_fwd (flags 0..3) or reads_rev (flags 4..7) at read_off[t];
 * names are name_prefix followed by read_id[t].  Returns the number of bytes the
 * records need (the text is complete only if that is <= outcap), or -1.
 * n_hits receives the number of records.
 */
extern "C" long long srm_format_records(int n, const srm_result *res, const uint32_t *ops,
                                        const srm_cand *cand, const char *genome,
                                        const int64_t *pos, const char *reads_fwd,
                                        const char *reads_rev, const uint64_t *read_off,
                                        const int32_t *read_id, const char *name_prefix,
                                        char *out, size_t outcap, long long *n_hits,
                                        int threads)
{
    if (n < 0) return -1;
    int T = threads > 0 ? threads : (int)std::thread::hardware_concurrency();
    T = std::max(1, std::min(T, 16));
    std::vector<std::string> part((size_t)T);
    std::vector<long long> hits((size_t)T, 0);
    std::vector<int> bad((size_t)T, 0);
    const size_t per = ((size_t)n + (size_t)T - 1) / (size_t)T;
    auto work = [&](int t) {
        size_t b = (size_t)t * per, e = std::min((size_t)n, b + per);
        char line[2 * (SRM_MAX_W + 256) + 256];
        char nm[96];
        std::string &dst = part[(size_t)t];
        for (size_t i = b; i < e; ++i) {
            const srm_cand &c = cand[i];
            if (res[i].score > (int)c.k) continue;
            snprintf(nm, sizeof nm, "%s%d", name_prefix, read_id[i]);
            const char *rd = (c.flag > 3 ? reads_rev : reads_fwd) + read_off[i];
            int len = srm_format_record(nm, (unsigned)pos[i], c.flag, c.k, c.L, c.w, &res[i],
                                        ops + c.ops_off, rd, genome + pos[i], line,
                                        (int)sizeof line);
            if (len < 0 || len >= (int)sizeof line) {
                bad[(size_t)t] = 1;
                continue;
            }
            dst.append(line, (size_t)len);
            hits[(size_t)t]++;
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < T; ++t) th.emplace_back(work, t);
    work(0);
    for (auto &x : th) x.join();
    long long total = 0, nh = 0;
    for (int t = 0; t < T; ++t) {
        if (bad[(size_t)t]) return -1;
        total += (long long)part[(size_t)t].size();
        nh += hits[(size_t)t];
    }
    if (n_hits) *n_hits = nh;
    if ((size_t)total <= outcap) {
        size_t o = 0;
        for (auto &p : part) {
            std::memcpy(out + o, p.data(), p.size());
            o += p.size();
        }
    }
    return total;
}
