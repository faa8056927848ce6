// CPU check of the record writer's word-level transcript scans
// (saruman_amd/csrc/rec_ops.h) against the op-by-op scans of the reference's
// post-processing (src/cuda-functions.cu:866-980), restated here one op at a time:
// random transcripts of every length from 0 to 400 ops with runs of each op at
// the ends, and window lengths around the row length, including ones <= 0.
//
//   rec_measure_cpu <seed> <cases>      prints "OK <cases>" or the first mismatch
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "rec_ops.h"

using srm_rec::Measure;

namespace {

uint64_t state;
uint32_t rnd()
{
    state ^= state << 13;
    state ^= state >> 7;
    state ^= state << 17;
    return (uint32_t)(state >> 11);
}

// the reference's loops, one op at a time: position p holds op n - 1 - p
Measure by_ops(const std::vector<int> &op, int w)
{
    const int n = (int)op.size();
    auto at = [&](int p) { return op[(size_t)(n - 1 - p)]; };
    Measure m;
    while (m.lead < n && at(m.lead) == 1) ++m.lead;
    for (int p = n - 1; p >= 0 && at(p) == 1; --p) ++m.trail;
    int i = 0, z = 0;
    while (i < w - m.trail - m.lead + z) {
        const int p = m.lead + i;
        if (p >= n) {
            m.eor = i;
            break;
        }
        if (at(p) == 1) {
            ++m.del;
            ++z;
        } else if (at(p) == 2) {
            ++m.ins;
            ++z;
        }
        ++i;
    }
    return m;
}

}  // namespace

int main(int argc, char **argv)
{
    state = strtoull(argv[1], nullptr, 0) * 0x9E3779B97F4A7C15ull + 1;
    const int cases = atoi(argv[2]);
    for (int c = 0; c < cases; ++c) {
        const int n = (int)(rnd() % 401);
        std::vector<int> op((size_t)n);
        const int style = (int)(rnd() % 4);
        for (int j = 0; j < n; ++j) {
            const uint32_t r = rnd() % 100;
            if (style == 0) op[(size_t)j] = (int)(rnd() % 3);             // any op
            else op[(size_t)j] = r < 90 ? 0 : (r < 95 ? 1 : 2);          // mostly diagonal
        }
        // runs of TOP at either end (leading and closing gaps), or all TOP
        const int lead_run = style >= 2 ? (int)(rnd() % 40) : 0;
        const int trail_run = style == 3 ? (int)(rnd() % 40) : 0;
        for (int j = 0; j < lead_run && j < n; ++j) op[(size_t)(n - 1 - j)] = 1;
        for (int j = 0; j < trail_run && j < n; ++j) op[(size_t)j] = 1;
        if (rnd() % 50 == 0)
            for (int &x : op) x = 1;
        std::vector<uint32_t> words((size_t)(n + 15) / 16 + 1, 0u);
        for (int j = 0; j < n; ++j) words[(size_t)(j >> 4)] |= (uint32_t)op[(size_t)j] << (2 * (j & 15));
        // bits above op n - 1 in the last word are garbage the scans must ignore
        if (n & 15) words[(size_t)(n >> 4)] |= rnd() << (2 * (n & 15));
        const int w = n - 60 + (int)(rnd() % 121);
        const Measure a = srm_rec::measure_ops(words.data(), n, w), b = by_ops(op, w);
        if (a.lead != b.lead || a.trail != b.trail || a.del != b.del || a.ins != b.ins ||
            a.eor != b.eor) {
            printf("MISMATCH case %d n %d w %d: words lead %d trail %d del %d ins %d eor %d, "
                   "ops lead %d trail %d del %d ins %d eor %d\n",
                   c, n, w, a.lead, a.trail, a.del, a.ins, a.eor, b.lead, b.trail, b.del, b.ins,
                   b.eor);
            return 1;
        }
    }
    printf("OK %d\n", cases);
    return 0;
}
