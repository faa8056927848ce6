// nw_kernels.hip -- gfx950 kernels for SARUMAN's semiglobal Needleman-Wunsch with
// traceback (the work of compute_score/get_alignment, src/cuda-functions.cu:76-536).
//
// The kernels here compute the same results:
//   plain_kernel  one candidate per thread, full (w+1)x(L+1) int16 matrix in HBM,
//                 value-based traceback on that matrix.  Handles every candidate
//                 (any match/mismatch, raw-byte or 2-bit sequences).
//   band_kernel   persistent: each wave pulls 64-candidate tiles from a device queue;
//   trace_kernel  one candidate per lane; the exact k-error band of a DP column
//                 (d = row - col in [-k, max(2s+k, w-L)]) lives in VGPRs, saturated
//                 at k+1; each column is stored as nibbles in a per-wave scratch
//                 ([column][word][lane], coalesced 256-byte stores) and the traceback
//                 walks those nibbles.  Needs match >= 0, mismatch >= 1, k <= 14;
//                 serves the costs other than match 0 / mismatch 1 (unit costs run
//                 the bit-parallel kernels of band_bits.hip).
//
// All reproduce the reference recurrence and traceback decisions bit for bit
// (quirks Q4-Q7 kept, Q1 selectable per candidate, Q2/Q3 canonical).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nw_device.h"
#include "nw_seq.h"

namespace srm {

// ---------------------------------------------------------------------------
// plain kernel
// ---------------------------------------------------------------------------
template <bool CODES>
struct SeqRef {
    const uint32_t *words;
    const char *bytes;
    int64_t off;
    __device__ __forceinline__ int at(int i) const
    {
        if (CODES) return (int)code_at(words, off + i);
        return (int)(unsigned char)bytes[off + i];
    }
};

template <bool CODES>
__global__ __launch_bounds__(256) void plain_kernel(PlainArgs a)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= a.n) return;
    const uint32_t t = a.idx[j];
    const srm_cand cd = a.cand[t];
    const int L = cd.L, w = cd.w, k = cd.k;
    const int s = flag_s(cd.flag, k), e = flag_e(cd.flag, k);
    const int ma = a.ma, mm = a.mm;
    SeqRef<CODES> rd{a.read_words, a.read_bytes, (int64_t)cd.read_off};
    SeqRef<CODES> wn{a.win_words, a.win_bytes, (int64_t)cd.win_off};
    int16_t *__restrict__ M = a.mat + a.mat_off[j];
    const int S = L + 1;

    for (int c = 0; c <= L; ++c) M[c] = (int16_t)c;                     // :413-416
    {
        int16_t prev = 0;
        for (int h = 1; h <= w; ++h) {                                   // :419-430 (Q7)
            prev = (h < 2 * s + 1) ? prev : (int16_t)(prev + 1);
            M[h * S] = prev;
        }
    }
    for (int h = 0; h < w; ++h) {                                        // :444-494
        const int gh = wn.at(h);
        const int16_t *__restrict__ up = M + h * S;
        int16_t *__restrict__ cur = M + (h + 1) * S;
        int16_t diag = up[0];
        int16_t left = cur[0];
        const int corr_row = (h >= w - 2 * e) ? mm : 0;
        for (int c = 0; c < L; ++c) {
            const int16_t top = up[c + 1];
            const int corr = (c == L - 1) ? corr_row : 0;
            const int16_t D = (int16_t)(diag + (gh != rd.at(c) ? mm : ma));
            const int16_t Lf = (int16_t)(left + mm);
            const int16_t T = (int16_t)(top + mm - corr);
            int16_t m = D;                                               // :89-105
            if (m > Lf) m = Lf;
            if (m > T) m = T;
            cur[c + 1] = m;
            diag = top;
            left = m;
        }
    }
    const int score = M[w * S + L];
    srm_result r;
    r.score = (int16_t)score;
    r.n_ops = 0;
    r.end_row = (uint16_t)w;
    r.end_col = (uint16_t)L;
    if (score <= k) {                                                    // :504
        const bool stop_row0 = (cd.bits & SRM_CAND_STOP_ROW0) != 0;
        uint32_t *__restrict__ ops = a.ops + cd.ops_off;
        const int cap = w + k;
        int row = w, col = L, n = 0;
        uint32_t acc = 0;
        for (;;) {
            if (stop_row0 && row == 0) break;                            // Q1
            const int v = M[row * S + col];
            int Dg, Lf, Tp;
            if (col != 0) {
                Dg = row >= 1 ? M[(row - 1) * S + col - 1] : SRM_FOREIGN;
                Lf = M[row * S + col - 1];
            } else {
                Dg = SRM_COL0;
                Lf = SRM_COL0;
            }
            Tp = row >= 1 ? M[(row - 1) * S + col] : SRM_FOREIGN;
            const int wd = imax(col - 1, 0), ht = imax(row - 1, 0);
            const bool eq = wn.at(ht) == rd.at(wd);
            const int op = decide(ma, mm, v, Dg, Lf, Tp, wd, ht, L, w, s, e, eq);
            if (op == 3 || n >= cap) break;
            acc |= (uint32_t)op << ((n & 15) * 2);
            ++n;
            if ((n & 15) == 0) {
                ops[(n >> 4) - 1] = acc;
                acc = 0;
            }
            row -= (op != 2);
            col -= (op != 1);
        }
        if (n & 15) ops[n >> 4] = acc;
        r.n_ops = (uint16_t)n;
        r.end_row = (uint16_t)row;
        r.end_col = (uint16_t)col;
    }
    a.res[t] = r;
}

// ---------------------------------------------------------------------------
// banded persistent kernel
// ---------------------------------------------------------------------------
// A tile's band scratch holds the band values as nibbles in blocks of 8 columns:
// word (q, i) of lane l sits at ((q * BMAX + i) * 64 + l) and holds band cell
// (diagonal) i of columns 8q .. 8q+7 (column 8q + j in bits 4j..4j+3).  The
// forward pass fills the BMAX words of a block every 8 columns with coalesced
// 256-byte wave stores at immediate offsets; trace_kernel reads them back.
template <int BMAX>
__device__ __forceinline__ uint32_t band_word(const uint32_t *__restrict__ scr, int i, int q,
                                              int lane)
{
    return scr[((size_t)q * BMAX + i) * 64 + lane];
}

// occupancy each band class is compiled for (waves per SIMD): the register
// budget is 512 / waves VGPRs
template <int NW>
struct BandWaves {
    static constexpr int value =
        NW <= 1 ? 8 : NW <= 2 ? 7 : NW <= 3 ? 6 : NW <= 5 ? 4 : NW <= 7 ? 3 : 2;
};

template <int NW>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(BandWaves<NW>::value)))
void band_kernel(BandArgs a)
{
    constexpr int BMAX = NW * 8;            // band cells per column held in VGPRs
    constexpr int WW = (BMAX + 15) / 16;    // code words of the sliding window
    const int lane = threadIdx.x & 63;
    const int ma = a.ma, mm = a.mm;

    for (;;) {
        uint32_t tile = 0;
        if (lane == 0) tile = atomicAdd(a.queue, 1u);
        tile = __builtin_amdgcn_readfirstlane(tile);
        if (tile >= a.ntiles) break;
        // x: perm start, y: count, z: max L, w: max k -- one tile per wave: made
        // wave-uniform so the column loop and its every-8/16-column steps stay
        // scalar branches
        uint4 td = a.tiles[tile];
        td.x = __builtin_amdgcn_readfirstlane(td.x);
        td.y = __builtin_amdgcn_readfirstlane(td.y);
        td.z = __builtin_amdgcn_readfirstlane(td.z);
        td.w = __builtin_amdgcn_readfirstlane(td.w);
        uint32_t *__restrict__ scr = a.scratch + a.tile_scratch[tile];
        const bool active = (uint32_t)lane < td.y;
        const uint32_t t = active ? a.perm[td.x + lane] : 0u;
        srm_cand cd = a.cand[t];
        const int L = active ? (int)cd.L : 0;
        const int w = cd.w, k = cd.k;
        const int s = flag_s(cd.flag, k), e = flag_e(cd.flag, k);
        const int dlo = -k;
        const int K = k + 1;
        const int Lmax = (int)td.z;

        // ---- column 0: M[h][0] = max(0, h - 2s) (unit steps, Q7); rows < 0 -> K
        int v[BMAX];
        uint32_t acc[BMAX];
#pragma unroll
        for (int i = 0; i < BMAX; ++i) {
            const int h = dlo + i;
            const int m0 = h < 0 ? K : imin(imax(h - 2 * s, 0), K);
            v[i] = m0;
            acc[i] = (uint32_t)v[i];
        }
        // window codes for column 1: W[i] = g[dlo + i]; a 16*WW-code window plus a stream
        uint32_t gw[WW];
        const int64_t gbase = (int64_t)cd.win_off + dlo;
#pragma unroll
        for (int j = 0; j < WW; ++j) gw[j] = active ? fetch16(a.win_words, gbase + 16 * j) : 0u;
        uint32_t gnext = 0;   // stream of window codes entering the top of the window
        uint32_t rbuf = 0;    // stream of read codes
        const int64_t rbase = (int64_t)cd.read_off;

        // one DP column; last selects the free trailing verticals of column L
        uint32_t lastmis[WW];   // mismatch fields of column L
        auto column = [&](int c, bool last) {
            // window for column c (c >= 2): shift by one code, take the next from the stream
            if (c >= 2) {
                if (((c - 2) & 15) == 0) gnext = fetch16(a.win_words, gbase + 16 * WW + (c - 2));
#pragma unroll
                for (int j = 0; j < WW - 1; ++j)
                    gw[j] = __builtin_amdgcn_alignbit(gw[j + 1], gw[j], 2);
                gw[WW - 1] = __builtin_amdgcn_alignbit(gnext, gw[WW - 1], 2);
                gnext >>= 2;
            }
            if (((c - 1) & 15) == 0) rbuf = fetch16(a.read_words, rbase + (c - 1));
            const uint32_t rc = rbuf & 3u;
            rbuf >>= 2;
            const uint32_t r8 = rc * 0x5555u;          // 24-bit multiply (full rate)
            const uint32_t rrep = r8 | (r8 << 16);     // the read code in all 16 fields
            uint32_t mis[WW];
#pragma unroll
            for (int j = 0; j < WW; ++j) {
                const uint32_t x = gw[j] ^ rrep;
                mis[j] = x | (x >> 1);   // bit 2i set <=> window code i != read code
            }
            if (last) {
#pragma unroll
                for (int j = 0; j < WW; ++j) lastmis[j] = mis[j];
            }
            int ifree = BMAX;            // first band cell whose top move is free
            if (last) ifree = w - 2 * e + 1 - c - dlo;
            const uint32_t sh = (uint32_t)(c & 7) * 4u;
            int up = 1 << 20;
#pragma unroll
            for (int i = 0; i < BMAX; ++i) {
                const int bit = (int)((mis[i >> 4] >> (2 * (i & 15))) & 1u);
                const int lf = (i + 1 < BMAX) ? v[i + 1] : (1 << 20);
                const int dg = v[i] + (bit ? mm : ma);
                const int tcost = (last && i >= ifree) ? 0 : mm;
                int nv = imin(imin(dg, lf + mm), up + tcost);
                nv = imin(nv, K);
                v[i] = nv;
                up = nv;
            }
            if (c <= k) {
                // row 0 keeps its unit-step init M[0][c] = c (Q7) whatever the mismatch cost
                const int i0 = k - c;
#pragma unroll
                for (int i = 0; i < BMAX; ++i) v[i] = (i == i0) ? imin(c, K) : v[i];
            }
            // nibble c & 7 of every diagonal's word; a block is stored when complete
            // and its words start again from 0
#pragma unroll
            for (int i = 0; i < BMAX; ++i) acc[i] += (uint32_t)v[i] << sh;
            if ((c & 7) == 7 || last) {
                // streaming stores: the band is read back only by the walk, much
                // later; the code words the next columns fetch stay in L2
                uint32_t *__restrict__ blk = scr + (size_t)(c >> 3) * BMAX * 64 + lane;
#pragma unroll
                for (int i = 0; i < BMAX; ++i) __builtin_nontemporal_store(acc[i], &blk[i * 64]);
#pragma unroll
                for (int i = 0; i < BMAX; ++i) acc[i] = 0u;
            }
        };

        for (int c0 = 1; c0 < Lmax; ++c0) {
            // the column is the same on every lane: kept scalar, so the steps taken
            // every 8 and 16 columns are scalar branches, not masked ones
            const int c = __builtin_amdgcn_readfirstlane(c0);
            if (c < L) column(c, false);
        }
        // tail of the tile's scratch, for the walk's first (vertical) run: columns
        // L and L - 1 by diagonal (nibble j of word m is diagonal 8m + j), then the
        // equal-character bits of column L
        uint32_t *__restrict__ tail = scr + (size_t)((Lmax + 8) >> 3) * BMAX * 64 + lane;
        auto store_column = [&](int c, int word0) {
#pragma unroll
            for (int m = 0; m < NW; ++m) {
                uint32_t x = 0;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int i = 8 * m + j;
                    x |= ((uint32_t)v[i] & 15u) << (4 * j);
                }
                __builtin_nontemporal_store(x, &tail[(word0 + m) * 64]);
            }
        };
        if (active) {
            store_column(L - 1, NW);
            column(L, true);
            store_column(L, 0);
            uint32_t eqw[2] = {0u, 0u};
#pragma unroll
            for (int i = 0; i < BMAX; ++i)
                if (!((lastmis[i >> 4] >> (2 * (i & 15))) & 1u)) eqw[i >> 5] |= 1u << (i & 31);
            tail[(2 * NW) * 64] = eqw[0];
            tail[(2 * NW + 1) * 64] = eqw[1];
        }

        // ---- score (cuda-functions.cu:495-497); the walk runs in trace_kernel
        if (active) {
            int score = K;
            const int iend = w - L - dlo;
            if (iend >= 0 && iend < BMAX) {
                const uint32_t wd0 = band_word<BMAX>(scr, iend, L >> 3, lane);
                score = imin((int)((wd0 >> ((L & 7) * 4)) & 15u), K);
            }
            srm_result r;
            r.score = (int16_t)score;
            r.n_ops = 0;
            r.end_row = (uint16_t)w;
            r.end_col = (uint16_t)L;
            a.res[t] = r;
        }
    }
}

// word (q, i) of a lane's band, 0xFFFFFFFF (nibbles read as K) outside it
template <int BMAX>
__device__ __forceinline__ uint32_t band_word_or_k(const uint32_t *__restrict__ scr, int i, int q,
                                                  int lane)
{
    return (i >= 0 && i < BMAX && q >= 0) ? band_word<BMAX>(scr, i, q, lane) : 0xFFFFFFFFu;
}

// 8 two-bit fields (bits 2j..2j+1) moved to nibbles (bits 4j..4j+1)
__device__ __forceinline__ uint32_t spread2(uint32_t f)
{
    f = (f | (f << 8)) & 0x00FF00FFu;
    f = (f | (f << 4)) & 0x0F0F0F0Fu;
    return (f | (f << 2)) & 0x33333333u;
}

// Traceback of every hit of a band class (cuda-functions.cu:504-534, App. A.3),
// one wave per tile of band_kernel, reading the band the forward pass left in
// the tile's scratch.  Few registers and many resident waves: the walk's reloads
// (one block of 8 columns per reload on a diagonal walk) wait for memory while
// other waves run.  The walk goes block by block: a run of diagonal steps inside
// a block of 8 columns is found from the block's word at once, the other steps
// (gaps, the free rows of column L, the start of the read) one cell at a time;
// column 0 comes from its closed form max(0, h - 2s) (Q7) without loads.
template <int NW>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8)))
void trace_kernel(BandArgs a)
{
    constexpr int BMAX = NW * 8;
    const int lane = threadIdx.x & 63;
    const uint32_t tile = (uint32_t)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    if (tile >= a.ntiles) return;
    const uint4 td = a.tiles[tile];
    if ((uint32_t)lane >= td.y) return;
    const uint32_t t = a.perm[td.x + lane];
    const srm_cand cd = a.cand[t];
    srm_result r = a.res[t];
    const int k = cd.k, K = k + 1;
    if (r.score > k) return;
    const int L = cd.L, w = cd.w;
    const int s = flag_s(cd.flag, k), e = flag_e(cd.flag, k);
    const int dlo = -k;
    const int ma = a.ma, mm = a.mm;
    const uint32_t *__restrict__ scr = a.scratch + a.tile_scratch[tile];
    const bool stop_row0 = (cd.bits & SRM_CAND_STOP_ROW0) != 0;
    uint32_t *__restrict__ ops = a.ops + cd.ops_off;
    const int64_t rbase = (int64_t)cd.read_off;
    const int cap = w + k;
    int row = w, col = L, n = 0;
    uint32_t opw = 0;
    bool done = false;
    auto emit = [&](int op) {
        opw |= (uint32_t)op << ((n & 15) * 2);
        ++n;
        if ((n & 15) == 0) {
            ops[(n >> 4) - 1] = opw;
            opw = 0;
        }
        row -= (op != 2);
        col -= (op != 1);
    };
    {
        // the first run, up column L (its free tail rows), from the packed tail
        const uint32_t *__restrict__ tail =
            scr + (size_t)((td.z + 8) >> 3) * BMAX * 64 + lane;
        // vectors in registers: a lane's word pick (i >> 3) becomes a select
        // chain, never a private array in scratch memory
        typedef uint32_t TailWords __attribute__((ext_vector_type(NW)));
        TailWords tl, tp;
#pragma unroll
        for (int m = 0; m < NW; ++m) {
            tl[m] = tail[m * 64];
            tp[m] = tail[(NW + m) * 64];
        }
        const uint32_t eq0 = tail[(2 * NW) * 64], eq1 = tail[(2 * NW + 1) * 64];
        auto nib = [&](const TailWords &X, int i) -> int {
            if (i < 0 || i >= BMAX) return 15;
            return (int)((X[i >> 3] >> ((i & 7) * 4)) & 15u);
        };
        const int wd = L - 1;
        while (col == L) {
            if (stop_row0 && row == 0) {
                done = true;
                break;
            }
            const int i = row - L - dlo;
            const int vv = nib(tl, i);
            const int Dg = row >= 1 ? imin(nib(tp, i), K) : SRM_FOREIGN;
            const int Lf = imin(nib(tp, i + 1), K);
            const int Tp = row >= 1 ? imin(nib(tl, i - 1), K) : SRM_FOREIGN;
            const bool eq = i >= 0 && i < BMAX && ((i < 32 ? eq0 >> i : eq1 >> (i - 32)) & 1u);
            const int op = decide(ma, mm, vv, Dg, Lf, Tp, wd, imax(row - 1, 0), L, w, s, e, eq);
            if (op == 3 || n >= cap) {
                done = true;
                break;
            }
            emit(op);
        }
    }
    // the read's codes, staged in this wave's LDS (lane-interleaved words: the
    // lanes' reads of one word index hit different banks); word u holds the
    // read_words word ((rbase - 1) >> 4) + u
    extern __shared__ uint32_t trace_lds[];
    const int rwn = a.code_words;
    uint32_t *__restrict__ lcodes = trace_lds + (size_t)(threadIdx.x >> 6) * 64 * rwn + lane;
    if (rwn && !done) {
        const uint32_t *__restrict__ src = a.read_words + ((rbase - 1) >> 4);
        const int nwd = ((L + 15) >> 4) + 3;   // <= rwn: L <= the class's max L
        for (int u = 0; u < nwd; ++u) lcodes[u * 64] = src[u];
    }
    // the walk's block: words of diagonal ci in column blocks cq (W0) and cq-1
    // (P0) and the character comparison of its 8 cells; the neighbour diagonals'
    // words (Wm: ci-1, Wp and Pp: ci+1) are loaded only when a cell of the block
    // needs them (a step that is not diagonal)
    int ci = -1, cq = -1;
    bool side = false;
    uint32_t W0 = 0, P0 = 0, Wm = 0, Wp = 0, Pp = 0;
    uint32_t mis = 0;          // bit 4j set <=> the characters of cell 8cq+j differ
    const int64_t woff = (int64_t)cd.win_off;
    auto m0 = [&](int h, int i) -> int {   // stored nibble of column 0, row h, band cell i
        return (i >= 0 && i < BMAX) ? imin(imax(h - 2 * s, 0), K) : 15;
    };
    while (!done) {
        if (stop_row0 && row == 0) {
            done = true;
            break;
        }
        if (col == 0) {
            // column 0: M[h][0] = max(0, h - 2s) clipped at K, as the forward pass
            // stored it; a vertical step or the stop remain (Dg = Lf = COL0;
            // with match >= 0, as on every banded candidate, the characters
            // never decide)
            const int i = row - dlo;
            const int vv = m0(row, i);
            const int Tp = row >= 1 ? imin(m0(row - 1, i - 1), K) : SRM_FOREIGN;
            const int op = decide(ma, mm, vv, SRM_COL0, SRM_COL0, Tp, 0, imax(row - 1, 0), L,
                                  w, s, e, false);
            if (op == 3 || n >= cap) {
                done = true;
                break;
            }
            emit(op);
            continue;
        }
        const int i = row - col - dlo;
        const int q = col >> 3;
        if (i != ci || q != cq) {
            W0 = (i == ci && q == cq - 1) ? P0 : band_word_or_k<BMAX>(scr, i, q, lane);
            P0 = band_word_or_k<BMAX>(scr, i, q - 1, lane);
            // codes of the block's 8 cells: read column 8q+j-1, window row
            // (i + dlo + 8q + j) - 1; exact for rows and columns >= 1, the only
            // cells whose characters can decide (Dg is FOREIGN on row 0)
            uint32_t rc;
            if (rwn) {   // words wi, wi + 1 < ((L - 1) >> 4) + 3 <= nwd (8q <= L - 1)
                // code rbase + 8q - 1 sits at place o of the staged codes
                const int o = (int)((rbase - 1) & 15) + 8 * q;
                const int wi = o >> 4;
                rc = __builtin_amdgcn_alignbit(lcodes[(wi + 1) * 64], lcodes[wi * 64],
                                               (uint32_t)(o & 15) * 2u);
            } else {
                rc = fetch16(a.read_words, rbase + 8 * q - 1);
            }
            const uint32_t gc = fetch16(a.win_words, woff + (i + dlo + 8 * q) - 1);
            const uint32_t x = rc ^ gc;
            mis = spread2((x | (x >> 1)) & 0x5555u);
            ci = i;
            cq = q;
            side = false;
        }
        const int j0 = col & 7;
        // the first two rules of decide(), one cell at a time
        int rr;
        const int rmax = imin(j0 + 1, imin(imin(col, row), cap - n));
        rr = 0;
        while (rr < rmax) {
            const int j2 = j0 - rr;
            const int v2 = (int)((W0 >> (4 * j2)) & 15u);
            const int d2 =
                imin(j2 ? (int)((W0 >> (4 * j2 - 4)) & 15u) : (int)(P0 >> 28), K);
            if (d2 + mm != v2 && (d2 + ma != v2 || ((mis >> (4 * j2)) & 1u))) break;
            ++rr;
        }
        if (rr > 0) {
            // rr diagonal steps: 0 bits in the transcript
            if (((n & 15) + rr) >= 16) {   // the run completes the current ops word
                ops[n >> 4] = opw;
                opw = 0;
            }
            n += rr;
            row -= rr;
            col -= rr;
            continue;
        }
        // one cell that is not on a diagonal run (col >= 1)
        if (!side) {
            Wm = band_word_or_k<BMAX>(scr, i - 1, q, lane);
            Wp = band_word_or_k<BMAX>(scr, i + 1, q, lane);
            Pp = band_word_or_k<BMAX>(scr, i + 1, q - 1, lane);
            side = true;
        }
        const uint32_t sh = (uint32_t)j0 * 4u;
        // a nibble of 15 outside the band reads as K (never equal to v <= k)
        const int vv = (int)((W0 >> sh) & 15u);
        const bool in = j0 != 0;
        const uint32_t dgw = in ? W0 : P0, lfw = in ? Wp : Pp;
        const uint32_t psh = in ? sh - 4u : 28u;
        const int Dg = row >= 1 ? imin((int)((dgw >> psh) & 15u), K) : SRM_FOREIGN;
        const int Lf = imin((int)((lfw >> psh) & 15u), K);
        const int Tp = row >= 1 ? imin((int)((Wm >> sh) & 15u), K) : SRM_FOREIGN;
        const bool eq = ((mis >> sh) & 1u) == 0u;   // decides only with row >= 1
        const int op =
            decide(ma, mm, vv, Dg, Lf, Tp, col - 1, imax(row - 1, 0), L, w, s, e, eq);
        if (op == 3 || n >= cap) {
            done = true;
            break;
        }
        emit(op);
    }
    if (n & 15) ops[n >> 4] = opw;
    r.n_ops = (uint16_t)n;
    r.end_row = (uint16_t)row;
    r.end_col = (uint16_t)col;
    a.res[t] = r;
}

// streaming copy of n 16-byte words (calibrates the HBM traffic counters)
__global__ __launch_bounds__(256) void copy_kernel(const uint4 *__restrict__ src,
                                                   uint4 *__restrict__ dst, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

hipError_t launch_copy(const void *src, void *dst, size_t bytes, hipStream_t st)
{
    const size_t n = bytes / 16;
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, st, (const uint4 *)src, (uint4 *)dst,
                       n);
    return hipGetLastError();
}

// clears the queue counters of a launch sequence (a kernel, not hipMemsetAsync,
// so the sequence can be captured into a graph and replayed)
__global__ void reset_counters(uint32_t *c, int n)
{
    int i = threadIdx.x;
    if (i < n) c[i] = 0;
}

// ---------------------------------------------------------------------------
// host-side launch wrappers
// ---------------------------------------------------------------------------
hipError_t launch_plain(const PlainArgs &a, bool codes, hipStream_t st)
{
    if (a.n <= 0) return hipSuccess;
    dim3 block(256), grid((a.n + 255) / 256);
    if (codes)
        hipLaunchKernelGGL(plain_kernel<true>, grid, block, 0, st, a);
    else
        hipLaunchKernelGGL(plain_kernel<false>, grid, block, 0, st, a);
    return hipGetLastError();
}

template <int NW>
static hipError_t launch_band_nw(const BandArgs &a, int blocks, hipStream_t st)
{
    hipLaunchKernelGGL((band_kernel<NW>), dim3(blocks), dim3(256), 0, st, a);
    return hipGetLastError();
}

template <int NW>
static int occupancy_nw()
{
    int blocks = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &blocks, (const void *)band_kernel<NW>, 256, 0);
    return e == hipSuccess && blocks > 0 ? blocks : 1;
}

// resident 256-thread blocks per CU of the forward pass of a band class (unit
// costs: bits_band_kernel)
int band_occupancy(int nw, bool unit)
{
    if (unit) return bits_band_occupancy(nw);
    switch (nw) {
    case 1: return occupancy_nw<1>();
    case 2: return occupancy_nw<2>();
    case 3: return occupancy_nw<3>();
    case 4: return occupancy_nw<4>();
    case 5: return occupancy_nw<5>();
    case 6: return occupancy_nw<6>();
    case 7: return occupancy_nw<7>();
    case 8: return occupancy_nw<8>();
    default: return 1;
    }
}

hipError_t launch_band(const BandArgs &a, int nw, bool unit, int blocks, int maxL, hipStream_t st)
{
    if (unit) return launch_bits_band(a, nw, blocks, maxL, st);
    switch (nw) {
    case 1: return launch_band_nw<1>(a, blocks, st);
    case 2: return launch_band_nw<2>(a, blocks, st);
    case 3: return launch_band_nw<3>(a, blocks, st);
    case 4: return launch_band_nw<4>(a, blocks, st);
    case 5: return launch_band_nw<5>(a, blocks, st);
    case 6: return launch_band_nw<6>(a, blocks, st);
    case 7: return launch_band_nw<7>(a, blocks, st);
    case 8: return launch_band_nw<8>(a, blocks, st);
    default: return hipErrorInvalidValue;
    }
}

template <int NW>
static hipError_t launch_trace_nw(const BandArgs &a, size_t lds, hipStream_t st)
{
    hipLaunchKernelGGL((trace_kernel<NW>), dim3((a.ntiles + 3) / 4), dim3(256), lds, st, a);
    return hipGetLastError();
}

// words per lane that stage a read of length <= maxL in LDS: codes from one
// before the read's first to 16 past its last, from word boundaries
static int trace_code_words(int maxL) { return ((maxL + 15) >> 4) + 3; }

// one wave per tile of the class; unit (match 0, mismatch 1): bits_trace_kernel.
// trace_kernel stages reads of up to 600 codes in LDS (at most 40 KiB per block
// of 4 waves); longer ones are read from read_words.
hipError_t launch_trace(const BandArgs &args, int nw, bool unit, int maxL, hipStream_t st)
{
    if (args.ntiles == 0) return hipSuccess;
    if (unit) return launch_bits_trace(args, nw, maxL, st);
    BandArgs a = args;
    a.code_words = maxL <= 600 ? trace_code_words(maxL) : 0;
    const size_t lds = (size_t)a.code_words * 256 * 4;
    switch (nw) {
    case 1: return launch_trace_nw<1>(a, lds, st);
    case 2: return launch_trace_nw<2>(a, lds, st);
    case 3: return launch_trace_nw<3>(a, lds, st);
    case 4: return launch_trace_nw<4>(a, lds, st);
    case 5: return launch_trace_nw<5>(a, lds, st);
    case 6: return launch_trace_nw<6>(a, lds, st);
    case 7: return launch_trace_nw<7>(a, lds, st);
    case 8: return launch_trace_nw<8>(a, lds, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_reset(uint32_t *counters, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(reset_counters, dim3(1), dim3(64 * ((n + 63) / 64)), 0, st, counters, n);
    return hipGetLastError();
}

}  // namespace srm
