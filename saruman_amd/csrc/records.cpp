// records.cpp -- host side of the output path: rebuilds the two alignment rows the
// reference's kernel leaves in a candidate's slot (get_alignment,
// src/cuda-functions.cu:108-315) from the compact transcript, and formats the TSV
// record of the reference's post-processing (src/cuda-functions.cu:828-994).
#include <cstdio>
#include <cstring>
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
