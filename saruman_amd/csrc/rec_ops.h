// rec_ops.h -- the record post-processing's scans of a 2-bit transcript done a
// word (16 ops) at a time.  Host/device code: the record kernels use it, and a
// CPU build checks it against the op-by-op scans of src/cuda-functions.cu:866-980
// on random and edge-case transcripts.
//
// Transcript: op j in bits 2(j & 15) of word j >> 4, DIAG 0, TOP 1, LEFT 2;
// printed position p holds op n - 1 - p (the reference reverses its rows).
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define SRM_RHD __host__ __device__
#else
#define SRM_RHD
#endif

namespace srm_rec {

SRM_RHD inline int popc32(uint32_t x) { return __builtin_popcount(x); }

// highest set bit of x != 0
SRM_RHD inline int top_bit(uint32_t x) { return 31 - __builtin_clz(x); }

// even-bit mask of the fields f_lo .. f_hi of a word (0 <= f_lo <= f_hi <= 15)
SRM_RHD inline uint32_t field_mask(int f_lo, int f_hi)
{
    const uint32_t upto = f_hi >= 15 ? 0xFFFFFFFFu : ((1u << (2 * f_hi + 2)) - 1u);
    return upto & ~((1u << (2 * f_lo)) - 1u) & 0x55555555u;
}

// per field of x, one bit at the field's even position: op == TOP, op == LEFT,
// op == DIAG, op != TOP
SRM_RHD inline uint32_t tops_of(uint32_t x) { return x & ~(x >> 1) & 0x55555555u; }
SRM_RHD inline uint32_t lefts_of(uint32_t x) { return (x >> 1) & ~x & 0x55555555u; }
SRM_RHD inline uint32_t diags_of(uint32_t x) { return ~(x | (x >> 1)) & 0x55555555u; }
SRM_RHD inline uint32_t non_tops_of(uint32_t x) { return (~x | (x >> 1)) & 0x55555555u; }

// bit index of the m-th set bit of s counted from the top (1 <= m <= popc(s)): the
// largest b with popc(s >> b) >= m, by a binary search over b
SRM_RHD inline int nth_bit_from_top(uint32_t s, int m)
{
    int b = 0;
    for (int step = 16; step >= 1; step >>= 1)
        if (popc32(s >> (b + step)) >= m) b += step;
    return b;
}

struct Measure {
    int lead = 0;    // leading read-row gaps (cuda-functions.cu:878-886)
    int trail = 0;   // closing read-row gaps (:897-907)
    int del = 0;     // gaps of the indel scan (:910-959) in the read row
    int ins = 0;     // and in the genome row
    int eor = 0;     // positions scanned when the scan ran into the row's end, else 0
};

// The scans of one hit's transcript (n ops, window length w).  The indel scan
// visits positions lead, lead + 1, .. while i < w - trail - lead + z (i positions
// done, z of them gaps), that is while fewer than w - trail - lead DIAG ops have
// been seen, and stops with eor = i at the row's end.  ld(W) gives word W.  The
// top eight words and word 0 are asked for together: they hold the leading and the
// closing gap runs (each ends in its first word unless that word is all TOP ops)
// and the indel scan's first words, so that a device thread mostly waits for
// memory once for all three scans; the indel scan goes on eight words at a time.
template <class Load>
SRM_RHD inline Measure measure_ops_by(Load ld, int n, int w)
{
    constexpr int B = 8;
    Measure m;
    if (n <= 0) return m;   // the scan, if any, meets the NUL at once: eor 0
    const int wtop = (n - 1) >> 4, ftop = (n - 1) & 15;
    uint32_t x[B];   // x[k]: word base - k of the batch at hand
#pragma unroll
    for (int k = 0; k < B; ++k) x[k] = wtop - k >= 0 ? ld(wtop - k) : 0u;
    const uint32_t w0 = ld(0);
    // leading TOP ops from op n - 1 downwards: the top words first
    bool found = false;
#pragma unroll
    for (int k = 0; k < B; ++k) {
        const int W = wtop - k;
        if (found || W < 0) continue;
        const int f_hi = k == 0 ? ftop : 15;
        const uint32_t nt = non_tops_of(x[k]) & field_mask(0, f_hi);
        if (nt) {
            m.lead += f_hi - (top_bit(nt) >> 1);
            found = true;
        } else {
            m.lead += f_hi + 1;
        }
    }
    for (int W = wtop - B; !found && W >= 0; --W) {   // a run of more than 8 words
        const uint32_t nt = non_tops_of(ld(W)) & 0x55555555u;
        if (nt) {
            m.lead += 15 - (top_bit(nt) >> 1);
            found = true;
        } else {
            m.lead += 16;
        }
    }
    // closing TOP ops from op 0 upwards
    for (int W = 0; W <= wtop; ++W) {
        const int f_hi = W == wtop ? ftop : 15;
        const uint32_t nt = non_tops_of(W == 0 ? w0 : ld(W)) & field_mask(0, f_hi);
        if (nt) {
            m.trail += __builtin_ctz(nt) >> 1;
            break;
        }
        m.trail += f_hi + 1;
    }
    const int need = w - m.trail - m.lead;   // DIAG ops the scan waits for
    const int jtop = n - 1 - m.lead;         // op of position lead
    if (need <= 0 || jtop < 0) return m;     // no pass, or the NUL at once (eor 0)
    const int jw = jtop >> 4;
    int diag = 0;
    // batches of B words from the top word (the first one already loaded); words
    // above jtop's hold no field of the scan
    for (int base = wtop; base >= 0; base -= B) {
        if (base != wtop) {
#pragma unroll
            for (int k = 0; k < B; ++k) x[k] = base - k >= 0 ? ld(base - k) : 0u;
        }
#pragma unroll
        for (int k = 0; k < B; ++k) {
            const int W = base - k;
            if (W < 0 || W > jw) continue;
            const int f_hi = W == jw ? jtop & 15 : 15;
            const uint32_t valid = field_mask(0, f_hi);
            const uint32_t d = diags_of(x[k]) & valid;
            const int c = popc32(d);
            if (diag + c < need) {
                diag += c;
                m.del += popc32(tops_of(x[k]) & valid);
                m.ins += popc32(lefts_of(x[k]) & valid);
                continue;
            }
            // the (need - diag)-th DIAG field from the top ends the scan
            const uint32_t range = field_mask(nth_bit_from_top(d, need - diag) >> 1, f_hi);
            m.del += popc32(tops_of(x[k]) & range);
            m.ins += popc32(lefts_of(x[k]) & range);
            return m;
        }
    }
    m.eor = jtop + 1;                        // n - lead positions, then the NUL
    return m;
}

SRM_RHD inline Measure measure_ops(const uint32_t *ops, int n, int w)
{
    return measure_ops_by([ops](int W) { return ops[W]; }, n, w);
}

}  // namespace srm_rec
