// band_bits.hip -- the banded kernels for unit costs (match 0, mismatch 1, gap 1),
// the costs of every SARUMAN run (src/saruman.c: match 0, mismatch 1).  Their
// results equal the plain kernel's (nw_kernels.hip) bit for bit; other costs run
// the nibble band of band_kernel / trace_kernel there.
//
// Forward pass (bits_band_kernel): with unit costs neighbouring DP cells differ by
// -1, 0 or +1, so a column of the band is two bit vectors (vertical deltas +1 and
// -1) and the next column follows from them and the column's equal-character bits
// in a couple of dozen word operations (Myers 1999; Hyyrö 2003 for the banded,
// diagonal-aligned form).  Bit i of a column is band cell (diagonal) i, row
// h = c + i - k.  Cells outside the band enter as upper bounds (the cell above
// the band as its left neighbour + 1, the cell below as the one above it + 1), so
// every band value is >= the reference's and equal to it wherever the reference's
// is <= k (the cells a hit's path and its comparisons use).  Rows above row 0
// continue M[h][c] = c - h, which the recurrence reproduces and which gives row 0
// its M[0][c] = c.
//
// What the walk needs is stored per block of 8 columns (columns 8q+1 .. 8q+8):
//   - D bits: for each diagonal i and column, whether the cell equals its
//     diagonal predecessor (D0).  8 bits per diagonal per block, the forward
//     pass's 8 column words transposed in registers;
//   - the block's boundary column 8q: the value of band cell 0 and the column's
//     delta vectors.  A band value is then boundary + deltas + D bits.
// That is 1 bit per cell plus 3-5 words per block, against 4 bits per cell for
// the nibble band.  Columns L - 1 and L go to the tile's tail as nibbles, in the
// format of trace_kernel's tail (column L has the free trailing rows).  The forward
// pass also leaves the read's and the band's window codes, 16 per word, lane-
// interleaved: the walk's lanes then read them a few cache lines per load instead
// of one line per lane.
//
// Walk (bits_trace_kernel): the decisions of cuda-functions.cu:121-276 with unit
// costs.  A cell steps diagonally iff D0 is 0 (Dg + mismatch == v) or the
// characters are equal (Dg == v), so a diagonal run is the run of cells with
// D0 & mismatch == 0: found for a whole 32-column segment from one word (the D
// bytes of its 4 blocks).  The walk carries the value v of its cell; the other
// cells a decision compares (Tp, Lf) are computed from the block's boundary and
// D bits.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "nw_device.h"
#include "nw_seq.h"

namespace srm {

template <int NW>
struct BitsShape {
    static constexpr int BMAX = 8 * NW;                  // band cells stored per column
    static constexpr int NWD = NW <= 4 ? 1 : 2;          // 32-bit words of a column vector
    // the transposed D bits: word j of a half holds diagonals j, 8+j, 16+j, 24+j
    // (one byte each, n of them used); a half is stored as the stream of its 8n
    // used bytes (byte n*j + g = diagonal 8g + j), 2n words
    static constexpr int NLO = NW <= 4 ? NW : 4;         // bytes used per word, diagonals 0-31
    static constexpr int NHI = NW <= 4 ? 0 : NW - 4;     // diagonals 32-63
    static constexpr int dwords(int n) { return 2 * n; }
    static constexpr int DWLO = dwords(NLO);
    static constexpr int DW = DWLO + dwords(NHI);        // D words per block
    static constexpr int WB = DW + 1 + 2 * NWD;          // + cell 0's value, +1 deltas, -1 deltas
    // a tile's scratch in rows of 64 words (word l of a row is lane l's), for a
    // tile whose longest read is Lmax: blocks (columns 1 .. Lmax-1), tail, read
    // codes (row u: codes 16u .. 16u+15 of the read), window codes (row r: codes
    // 16r .. 16r+15 from the code of band cell 0 in column 1)
    __host__ __device__ static constexpr int blocks(int Lmax) { return Lmax >= 2 ? ((Lmax - 2) >> 3) + 1 : 0; }
    __host__ __device__ static constexpr int tail_row(int Lmax) { return blocks(Lmax) * WB; }
    __host__ __device__ static constexpr int read_row(int Lmax) { return tail_row(Lmax) + 2 * NW + 2; }
    __host__ __device__ static constexpr int win_row(int Lmax) { return read_row(Lmax) + ((Lmax + 31) >> 4) + 1; }
    __host__ __device__ static constexpr int rows(int Lmax) { return win_row(Lmax) + 2 * (NWD + ((Lmax + 30) >> 5) + 2); }
    // checkpoint layout (the walk recomputes a 32-column segment from its first
    // column's state): per segment g the state of column 32g -- band cell 0's
    // value, the +1 and -1 delta vectors, WS words -- then the tail, read and
    // window code rows as above
    static constexpr int WS = 1 + 2 * NWD;
    __host__ __device__ static constexpr int segs(int Lmax) { return Lmax >= 2 ? ((Lmax - 2) >> 5) + 1 : 0; }
    __host__ __device__ static constexpr int ck_tail_row(int Lmax) { return segs(Lmax) * WS; }
    __host__ __device__ static constexpr int ck_read_row(int Lmax) { return ck_tail_row(Lmax) + 2 * NW + 2; }
    __host__ __device__ static constexpr int ck_win_row(int Lmax) { return ck_read_row(Lmax) + ((Lmax + 31) >> 4) + 1; }
    __host__ __device__ static constexpr int ck_rows(int Lmax) { return ck_win_row(Lmax) + 2 * (NWD + ((Lmax + 30) >> 5) + 2); }
    template <bool CK> __host__ __device__ static constexpr int tail_at(int Lmax) { return CK ? ck_tail_row(Lmax) : tail_row(Lmax); }
    template <bool CK> __host__ __device__ static constexpr int read_at(int Lmax) { return CK ? ck_read_row(Lmax) : read_row(Lmax); }
    template <bool CK> __host__ __device__ static constexpr int win_at(int Lmax) { return CK ? ck_win_row(Lmax) : win_row(Lmax); }
};

// word and bit offset of diagonal i's D byte among a block's D words
template <int NW>
__device__ __forceinline__ void dloc(int i, int &word, int &shift)
{
    using S = BitsShape<NW>;
    const bool hi = S::NHI > 0 && i >= 32;
    const int n = hi ? S::NHI : S::NLO;
    const int base = hi ? S::DWLO : 0;
    const int ii = hi ? i - 32 : i;
    const int t = n * (ii & 7) + (ii >> 3);   // place in the half's byte stream
    word = base + (t >> 2);
    shift = 8 * (t & 3);
}

// bits 0, 2, .., 30 of x moved to bits 0 .. 15
__device__ __forceinline__ uint32_t even_bits(uint32_t x)
{
    x &= 0x55555555u;
    x = (x | (x >> 1)) & 0x33333333u;
    x = (x | (x >> 2)) & 0x0F0F0F0Fu;
    x = (x | (x >> 4)) & 0x00FF00FFu;
    return (x | (x >> 8)) & 0x0000FFFFu;
}

// 32 codes (x: the first 16, y: the next 16) as two bit planes (p0: low bits,
// p1: high bits; bit b = code b)
__device__ __forceinline__ void planes_of(uint32_t x, uint32_t y, uint32_t &p0, uint32_t &p1)
{
    p0 = even_bits(x) | (even_bits(y) << 16);
    p1 = even_bits(x >> 1) | (even_bits(y >> 1) << 16);
}

// the 32 codes from code index idx as two bit planes
__device__ __forceinline__ void planes32(const uint32_t *__restrict__ words, int64_t idx,
                                         uint32_t &p0, uint32_t &p1)
{
    const int64_t wi = idx >> 4;
    const uint32_t sh = (uint32_t)(idx & 15) * 2u;
    const uint32_t a = words[wi], b = words[wi + 1], c = words[wi + 2];
    planes_of(__builtin_amdgcn_alignbit(b, a, sh), __builtin_amdgcn_alignbit(c, b, sh), p0, p1);
}

// 8x8 bit transposes in each byte of x[0..7]: afterwards bit j' of byte g of x[j]
// is what bit j of byte g of x[j'] was
__device__ __forceinline__ void transpose8(uint32_t (&x)[8])
{
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t t = ((x[r] >> 4) ^ x[r + 4]) & 0x0F0F0F0Fu;
        x[r + 4] ^= t;
        x[r] ^= t << 4;
    }
#pragma unroll
    for (int r = 0; r < 8; r += (r & 1) ? 3 : 1) {   // 0, 1, 4, 5
        const uint32_t t = ((x[r] >> 2) ^ x[r + 2]) & 0x33333333u;
        x[r + 2] ^= t;
        x[r] ^= t << 2;
    }
#pragma unroll
    for (int r = 0; r < 8; r += 2) {
        const uint32_t t = ((x[r] >> 1) ^ x[r + 1]) & 0x55555555u;
        x[r + 1] ^= t;
        x[r] ^= t << 1;
    }
}

// the band cells of column o + 1 whose characters differ: window code planes
// ch0/ch1 (chunks of 32 codes from the column's band cell 0, shifted by o & 31)
// against the read code o from rbuf (its 16-code word)
template <int NWD>
__device__ __forceinline__ void bits_ne(const uint32_t (&ch0)[NWD + 1], const uint32_t (&ch1)[NWD + 1],
                                        uint32_t rbuf, int o, uint32_t (&ne)[NWD])
{
    const uint32_t sh = (uint32_t)(o & 31);
    const uint32_t rsh = (uint32_t)(o & 15) * 2u;
    // the read code's bits as all-zero / all-one masks
    const uint32_t r0 = (uint32_t)__builtin_amdgcn_sbfe((int)rbuf, rsh, 1);
    const uint32_t r1 = (uint32_t)__builtin_amdgcn_sbfe((int)rbuf, rsh + 1u, 1);
#pragma unroll
    for (int m = 0; m < NWD; ++m) {
        const uint32_t g0 = __builtin_amdgcn_alignbit(ch0[m + 1], ch0[m], sh);
        const uint32_t g1 = __builtin_amdgcn_alignbit(ch1[m + 1], ch1[m], sh);
        ne[m] = (g0 ^ r0) | (g1 ^ r1);
    }
}

// one column of the recurrence: the previous column's deltas VP / VN and the
// cells whose characters differ give the column's D0 words and (when live) its
// deltas
template <int NWD>
__device__ __forceinline__ void bits_column(uint32_t (&VP)[NWD], uint32_t (&VN)[NWD],
                                            const uint32_t (&ne)[NWD], uint32_t (&d0)[NWD], bool live)
{
    // previous column's deltas moved to this column's rows (cell i of column c - 1
    // is cell i - 1 here); the cell below the band enters as +1
    uint32_t vp[NWD], vn[NWD];
    if (NWD == 1) {
        vp[0] = __builtin_amdgcn_alignbit(1u, VP[0], 1);
        vn[0] = VN[0] >> 1;
    } else {
        vp[0] = __builtin_amdgcn_alignbit(VP[NWD - 1], VP[0], 1);
        vp[NWD - 1] = __builtin_amdgcn_alignbit(1u, VP[NWD - 1], 1);
        vn[0] = __builtin_amdgcn_alignbit(VN[NWD - 1], VN[0], 1);
        vn[NWD - 1] = VN[NWD - 1] >> 1;
    }
    // D0 = (((Eq & vp) + vp) ^ vp) | Eq | vn, the add carrying across words
    if (NWD == 1) {
        const uint32_t eqp = vp[0] & ~ne[0];
        d0[0] = ((eqp + vp[0]) ^ vp[0]) | ~ne[0] | vn[0];
    } else {
        const uint64_t P = ((uint64_t)vp[NWD - 1] << 32) | vp[0];
        const uint64_t E = ~(((uint64_t)ne[NWD - 1] << 32) | ne[0]);
        const uint64_t N = ((uint64_t)vn[NWD - 1] << 32) | vn[0];
        const uint64_t D = (((E & P) + P) ^ P) | E | N;
        d0[0] = (uint32_t)D;
        d0[NWD - 1] = (uint32_t)(D >> 32);
    }
    // horizontal deltas, then this column's vertical deltas; the cell above the
    // band enters with a horizontal +1
    uint32_t hp[NWD], hn[NWD];
#pragma unroll
    for (int m = 0; m < NWD; ++m) {
        hp[m] = vn[m] | ~(d0[m] | vp[m]);
        hn[m] = d0[m] & vp[m];
    }
    uint32_t xh[NWD], hns[NWD];
    xh[0] = (hp[0] << 1) | 1u;
    hns[0] = hn[0] << 1;
    if (NWD == 2) {
        xh[NWD - 1] = __builtin_amdgcn_alignbit(hp[NWD - 1], hp[0], 31);
        hns[NWD - 1] = __builtin_amdgcn_alignbit(hn[NWD - 1], hn[0], 31);
    }
#pragma unroll
    for (int m = 0; m < NWD; ++m) {
        VN[m] = live ? xh[m] & d0[m] : VN[m];
        VP[m] = live ? hns[m] | ~(d0[m] | xh[m]) : VP[m];
    }
}

template <int NW>
struct BitsWaves {
    static constexpr int value = NW <= 4 ? 8 : 6;
};

// Forward pass of a unit-cost band class: one wave per 64-candidate tile, one
// candidate per lane (the tile format of band_kernel).  Not persistent: the
// column loop is short, and tiles taken from one global counter serialise on its
// atomics; the workgroup dispatcher hands out the tiles (longest first) instead.
template <int NW, bool CK>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(BitsWaves<NW>::value)))
void bits_band_kernel(BandArgs a)
{
    using S = BitsShape<NW>;
    constexpr int BMAX = S::BMAX, NWD = S::NWD, DW = S::DW, WB = S::WB;
    constexpr int WW = (BMAX + 15) / 16;   // code words of column L's band
    const int lane = threadIdx.x & 63;

    {
        const uint32_t tile =
            __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
        if (tile >= a.ntiles) return;
        uint4 td = a.tiles[tile];
        td.x = __builtin_amdgcn_readfirstlane(td.x);
        td.y = __builtin_amdgcn_readfirstlane(td.y);
        td.z = __builtin_amdgcn_readfirstlane(td.z);
        uint32_t *__restrict__ scr = a.scratch + a.tile_scratch[tile];
        const bool active = (uint32_t)lane < td.y;
        const uint32_t pos = td.x + (active ? (uint32_t)lane : 0u);   // tile order
        const srm_cand cd = a.cand_tile[pos];
        const int L = active ? (int)cd.L : 0;
        const int w = cd.w, k = cd.k;
        const int s = flag_s(cd.flag, k), e = flag_e(cd.flag, k);
        const int dlo = -k;
        const int K = k + 1;
        const int Lmax = (int)td.z;

        // ---- column 0: deltas of M[h][0] = max(0, h - 2s) (Q7) and of c - h above
        // row 0: -1 for h <= 0 (cells 0..k), 0 up to row 2s, +1 below
        uint32_t VP[NWD], VN[NWD];
#pragma unroll
        for (int m = 0; m < NWD; ++m) {
            const int p0 = k + 2 * s + 1 - 32 * m;   // first +1 cell, relative to word m
            VP[m] = p0 <= 0 ? ~0u : p0 >= 32 ? 0u : ~((1u << p0) - 1u);
            const int n0 = k + 1 - 32 * m;           // cells below n0 are -1
            VN[m] = n0 <= 0 ? 0u : n0 >= 32 ? ~0u : (1u << n0) - 1u;
        }
        int vt = k;   // value of band cell 0 (row -k: M = k)
        auto store_boundary = [&](int q) {
            // checkpoint layout: the first column of every 32-column segment only
            if (CK && (q & 3)) return;
            // (the stored layout keeps it after the block's D words)
            constexpr int o0 = CK ? 0 : DW;
            uint32_t *__restrict__ b = scr + (size_t)(CK ? (q >> 2) * S::WS : q * WB) * 64 + lane;
            __builtin_nontemporal_store((uint32_t)vt, &b[o0 * 64]);
#pragma unroll
            for (int m = 0; m < NWD; ++m) {
                __builtin_nontemporal_store(VP[m], &b[(o0 + 1 + m) * 64]);
                __builtin_nontemporal_store(VN[m], &b[(o0 + 1 + NWD + m) * 64]);
            }
        };

        // window codes of band cell 0 in column c: gbase + c - 1 (cell (h, c)
        // compares window row h - 1 with read column c - 1).  Bit planes of the
        // codes stream through chunks of 32: ch[m] holds codes gbase + 32(m0 + m).
        const int64_t gbase = (int64_t)cd.win_off + dlo;
        const int64_t rbase = (int64_t)cd.read_off;
        // the window codes also go to the tile's window rows (2 per chunk) for the walk
        uint32_t *__restrict__ wrow = scr + (size_t)S::template win_at<CK>(Lmax) * 64 + lane;
        uint32_t *__restrict__ rrow = scr + (size_t)S::template read_at<CK>(Lmax) * 64 + lane;
        uint32_t ch0[NWD + 1], ch1[NWD + 1];
#pragma unroll
        for (int m = 0; m <= NWD; ++m) {
            uint32_t x = 0u, y = 0u;
            if (active) {
                const int64_t idx = gbase + 32 * m;
                const uint32_t *__restrict__ wsrc = a.win_words + (idx >> 4);
                const uint32_t sh = (uint32_t)(idx & 15) * 2u;
                x = __builtin_amdgcn_alignbit(wsrc[1], wsrc[0], sh);
                y = __builtin_amdgcn_alignbit(wsrc[2], wsrc[1], sh);
            }
            planes_of(x, y, ch0[m], ch1[m]);
            __builtin_nontemporal_store(x, &wrow[(2 * m) * 64]);
            __builtin_nontemporal_store(y, &wrow[(2 * m + 1) * 64]);
        }
        // the read's codes, staged in this wave's LDS for reads of up to 250 codes
        // (lane-interleaved words; word u holds read_words word ((rbase - 1) >> 4) + u):
        // one pass over the read instead of a reload every 16 columns
        extern __shared__ uint32_t band_lds[];
        const int rwn = a.code_words;
        uint32_t *__restrict__ lcodes = band_lds + (size_t)(threadIdx.x >> 6) * 64 * rwn + lane;
        if (rwn && active) {
            const uint32_t *__restrict__ src = a.read_words + ((rbase - 1) >> 4);
            const int nwd = ((L + 15) >> 4) + 3;   // <= rwn: L <= the class's max L
            for (int u = 0; u < nwd; ++u) lcodes[u * 64] = src[u];
        }
        // 16 read codes from code offset o (staged: code rbase + o is at place
        // rp0 + o; every place read is inside the lane's rwn words)
        const int rp0 = (int)((rbase - 1) & 15) + 1;
        auto read16 = [&](int o) -> uint32_t {
            if (rwn) {
                const int p = rp0 + o;
                const int wi = p >> 4;
                return __builtin_amdgcn_alignbit(lcodes[(wi + 1) * 64], lcodes[wi * 64],
                                                 (uint32_t)(p & 15) * 2u);
            }
            // a lane past its read (o >= L in a tile of longer reads) never uses
            // these codes: it reads at its read's end, inside the padding
            return fetch16(a.read_words, rbase + imin(o, L));
        };
        uint32_t rbuf = 0;
        // the next chunk of window codes, loaded one block ahead (before the block's
        // stores, so waiting for them never waits for the stores)
        const uint32_t gsh0 = (uint32_t)(gbase & 15) * 2u;   // code offset of gbase + 32m
        // column L's window codes, for the tail: loaded now, used after the loop
        uint32_t gw[WW];
#pragma unroll
        for (int u = 0; u < WW; ++u)
            gw[u] = active ? fetch16(a.win_words, gbase + (L - 1) + 16 * u) : 0u;
        uint32_t pg0 = 0, pg1 = 0, pg2 = 0;
        uint32_t d[8][NWD];   // D0 words of the group's 8 columns
#pragma unroll
        for (int j = 0; j < 8; ++j)
#pragma unroll
            for (int m = 0; m < NWD; ++m) d[j][m] = 0u;

        for (int cb0 = 1; cb0 < Lmax; cb0 += 8) {
            // columns cb .. cb+7 = block q; cb is the same on every lane
            const int cb = __builtin_amdgcn_readfirstlane(cb0);
            const int q = (cb - 1) >> 3;
            // refills from the loads issued one block earlier, on every lane (lanes
            // past their read decode words they never use)
            if (((cb - 1) & 31) == 0 && cb > 1) {
#pragma unroll
                for (int m = 0; m < NWD; ++m) {
                    ch0[m] = ch0[m + 1];
                    ch1[m] = ch1[m + 1];
                }
                const uint32_t x = __builtin_amdgcn_alignbit(pg1, pg0, gsh0);
                const uint32_t y = __builtin_amdgcn_alignbit(pg2, pg1, gsh0);
                planes_of(x, y, ch0[NWD], ch1[NWD]);
                const int m = NWD + ((cb - 1) >> 5);
                __builtin_nontemporal_store(x, &wrow[(2 * m) * 64]);
                __builtin_nontemporal_store(y, &wrow[(2 * m + 1) * 64]);
            }
            if (((cb - 1) & 15) == 0) {
                rbuf = read16(cb - 1);
                __builtin_nontemporal_store(rbuf, &rrow[((cb - 1) >> 4) * 64]);
            }
            store_boundary(q);   // column cb - 1 = 8q (garbage on lanes past their read)
            // every lane steps every column of the tile (no branch per column); a
            // lane past its read keeps the deltas of its column L - 1.  Blocks where
            // every lane with a candidate is inside its read (most of them: a tile's
            // reads are of about one length) run a copy of the steps without that
            // select
            auto block_columns = [&](auto every_live) {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int c = cb + j;
                const bool live = decltype(every_live)::value || c < L;
                uint32_t ne[NWD], d0[NWD];
                bits_ne<NWD>(ch0, ch1, rbuf, c - 1, ne);
                bits_column<NWD>(VP, VN, ne, d0, live);
#pragma unroll
                for (int m = 0; m < NWD; ++m) d[j][m] = d0[m];
            }
            };
            if (__all(!active || cb + 8 <= L))
                block_columns(std::true_type{});
            else
                block_columns(std::false_type{});
            // loads for the next block's window refill (columns from cb + 8), issued
            // before this block's stores
            {
                const int o2 = cb + 7;
                if (cb + 8 < L) {
                    if ((o2 & 31) == 0) {
                        const int64_t wi = (gbase + o2 + 32 * NWD) >> 4;
                        pg0 = a.win_words[wi];
                        pg1 = a.win_words[wi + 1];
                        pg2 = a.win_words[wi + 2];
                    }
                }
            }
            // ---- block q: D bits by diagonal (the checkpoint layout stores none: the
            // walk recomputes them)
            uint32_t diag0 = 0;
            if (CK) {
#pragma unroll
                for (int j = 0; j < 8; ++j) diag0 |= (d[j][0] & 1u) << j;
            } else {
            uint32_t *__restrict__ blk = scr + (size_t)q * WB * 64 + lane;
#pragma unroll
            for (int m = 0; m < NWD; ++m) {
                uint32_t x[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) x[j] = d[j][m];
                transpose8(x);
                if (m == 0) diag0 = x[0] & 0xFFu;   // diagonal 0, columns cb .. cb+7
                const int n = m == 0 ? S::NLO : S::NHI;
                const int base = m == 0 ? 0 : S::DWLO;
#pragma unroll
                for (int u = 0; u < 2 * n; ++u) {
                    // bytes 4u .. 4u+3 of the stream (all indices are constants)
                    uint32_t y = 0;
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const int t = 4 * u + b;
                        y |= ((x[t / n] >> (8 * (t % n))) & 0xFFu) << (8 * b);
                    }
                    __builtin_nontemporal_store(y, &blk[(base + u) * 64]);
                }
            }
            }
            // band cell 0 grows by 1 per column where its D0 bit is 0
            const int cnt = imax(imin(L - cb, 8), 0);
            vt += cnt - __builtin_popcount(diag0 & ((1u << cnt) - 1u));
        }

        // ---- tail for the walk's first run: columns L - 1 and L as nibbles by
        // diagonal (nibble j of word m is cell 8m + j, values clipped at K), then
        // the equal-character bits of column L.  Column L - 1 from its deltas;
        // column L with the free top moves of its trailing rows.
        uint32_t *__restrict__ tail = scr + (size_t)S::template tail_at<CK>(Lmax) * 64 + lane;
        if (active) {
            const uint32_t rc = read16(L - 1) & 3u;
            const uint32_t rrep = rc * 0x55555555u;
            const int ifree = w - 2 * e + 1 - L - dlo;   // first cell with a free top move
            const int iend = w - L - dlo;                // cell of row w
            int score = K;
            int cur = vt;                                // column L - 1, cell i
            int up = 1 << 20;                            // column L, cell i - 1
            uint32_t eqw[2] = {0u, 0u};
#pragma unroll
            for (int m = 0; m < NW; ++m) {
                uint32_t xp = 0, xl = 0;
#pragma unroll
                for (int jj = 0; jj < 8; ++jj) {
                    const int i = 8 * m + jj;
                    int nxt = 1 << 20;                   // column L - 1, cell i + 1
                    if (i + 1 < 32 * NWD) {
                        const int b = i + 1;
                        nxt = cur + (int)((VP[b >> 5] >> (b & 31)) & 1u) -
                              (int)((VN[b >> 5] >> (b & 31)) & 1u);
                    }
                    const uint32_t x = (gw[i >> 4] ^ rrep) >> (2 * (i & 15));
                    const int mis = (int)((x | (x >> 1)) & 1u);
                    int nv = imin(cur + mis, nxt + 1);
                    nv = imin(nv, up + (i >= ifree ? 0 : 1));
                    xp |= (uint32_t)imin(cur, K) << (4 * jj);
                    xl |= (uint32_t)imin(nv, K) << (4 * jj);
                    if (!mis) eqw[i >> 5] |= 1u << (i & 31);
                    if (i == iend) score = imin(nv, K);
                    up = nv;
                    cur = nxt;
                }
                __builtin_nontemporal_store(xl, &tail[m * 64]);
                __builtin_nontemporal_store(xp, &tail[(NW + m) * 64]);
            }
            tail[(2 * NW) * 64] = eqw[0];
            tail[(2 * NW + 1) * 64] = eqw[1];
            // ---- score (cuda-functions.cu:495-497), in tile order: bits_trace_kernel
            // writes every candidate's result (one scattered write per candidate)
            a.tile_score[pos] = score;
        }
    }
}

// Traceback of every hit of a unit-cost band class (cuda-functions.cu:504-534),
// one wave per tile of bits_band_kernel.  The first run (column L's free rows)
// and column 0 are as in trace_kernel; elsewhere the walk goes block by block.
// occupancy target of the walk: the checkpoint layout holds a segment's 32 D0
// columns in registers (a lower target leaves room for them without spilling)
template <int NW, bool CK>
struct TraceWaves {
    static constexpr int value = CK ? (NW <= 4 ? 4 : 3) : (NW <= 4 ? 8 : 7);
};

template <int NW, bool CK>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TraceWaves<NW, CK>::value)))
void bits_trace_kernel(BandArgs a)
{
    using S = BitsShape<NW>;
    constexpr int BMAX = S::BMAX, NWD = S::NWD, DW = S::DW, WB = S::WB;
    const int lane = threadIdx.x & 63;
    const uint32_t tile = (uint32_t)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    if (tile >= a.ntiles) return;
    const uint4 td = a.tiles[tile];
    if ((uint32_t)lane >= td.y) return;
    const uint32_t pos = td.x + lane;   // tile order
    const srm_cand cd = a.cand_tile[pos];
    const int score = a.tile_score[pos];
    // the packed tail (the first run's columns L - 1 and L), loaded with the
    // descriptor rather than after the score test
    const uint32_t *__restrict__ tail =
        a.scratch + a.tile_scratch[tile] + (size_t)S::template tail_at<CK>((int)td.z) * 64 + lane;
    typedef uint32_t TailWords __attribute__((ext_vector_type(NW)));
    TailWords tl, tp;
#pragma unroll
    for (int m = 0; m < NW; ++m) {
        tl[m] = tail[m * 64];
        tp[m] = tail[(NW + m) * 64];
    }
    const uint32_t eq0 = tail[(2 * NW) * 64], eq1 = tail[(2 * NW + 1) * 64];
    const int k = cd.k, K = k + 1;
    const uint32_t t = a.perm[pos];
    srm_result r;
    r.score = (int16_t)score;
    const int L = cd.L, w = cd.w;
    if (score > k) {   // no walk: the score and the walk's start cell
        r.n_ops = 0;
        r.end_row = (uint16_t)w;
        r.end_col = (uint16_t)L;
        a.res[t] = r;
        return;
    }
    const int s = flag_s(cd.flag, k), e = flag_e(cd.flag, k);
    const int dlo = -k;
    const uint32_t *__restrict__ scr = a.scratch + a.tile_scratch[tile];
    const bool stop_row0 = (cd.bits & SRM_CAND_STOP_ROW0) != 0;
    uint32_t *__restrict__ ops = a.ops + cd.ops_off;
    // the transcript words go to LDS as they complete (lane-interleaved), and to
    // the slot in one burst at the end, so that the L2 merges each slot's sectors
    // (a.stage_words words per lane; 0: straight to the slot)
    extern __shared__ uint32_t walk_lds[];
    uint32_t *__restrict__ stage = walk_lds + (size_t)(threadIdx.x >> 6) * 64 * a.stage_words + lane;
    auto put_word = [&](int x, uint32_t v) {
        if (a.stage_words) stage[x * 64] = v;
        else ops[x] = v;
    };
    const int cap = w + k;
    int row = w, col = L, n = 0;
    uint32_t opw = 0;
    bool done = false;
    int v = 0;   // value of the walk's cell once it has left column L (<= k: exact)
    auto emit = [&](int op) {
        opw |= (uint32_t)op << ((n & 15) * 2);
        ++n;
        if ((n & 15) == 0) {
            put_word((n >> 4) - 1, opw);
            opw = 0;
        }
        row -= (op != 2);
        col -= (op != 1);
    };
    {
        // the first run, up column L (its free tail rows), from the packed tail
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
            const int op = decide(0, 1, vv, Dg, Lf, Tp, wd, imax(row - 1, 0), L, w, s, e, eq);
            if (op == 3 || n >= cap) {
                done = true;
                break;
            }
            emit(op);
        }
        if (!done) v = nib(tp, row - (L - 1) - dlo);   // the cell in column L - 1
    }
    // the read's codes and the band's window codes, as the forward pass left them
    // in the tile's rows: a segment's words are in the same few rows for the lanes
    const uint32_t *__restrict__ rrow = scr + (size_t)S::template read_at<CK>((int)td.z) * 64 + lane;
    const uint32_t *__restrict__ wrow = scr + (size_t)S::template win_at<CK>((int)td.z) * 64 + lane;
    // checkpoint layout: the D0 words of the 32 columns of segment rg, recomputed
    // from the segment's first column, and the state of its four blocks' first
    // columns (band cell 0's value, the delta vectors)
    uint32_t dcol[32][NWD];
    int bv[4] = {0, 0, 0, 0};
    uint32_t bp[4][NWD], bn[4][NWD];
    int rg = -1;
    auto recompute = [&](int g) {
        const uint32_t *__restrict__ sb = scr + (size_t)g * S::WS * 64 + lane;
        int v0 = (int)sb[0];
        uint32_t VPr[NWD], VNr[NWD];
#pragma unroll
        for (int m = 0; m < NWD; ++m) {
            VPr[m] = sb[(1 + m) * 64];
            VNr[m] = sb[(1 + NWD + m) * 64];
        }
        // window code planes of chunks g .. g + NWD and the read codes of the
        // segment, from the rows the forward pass left
        uint32_t c0[NWD + 1], c1[NWD + 1];
#pragma unroll
        for (int m = 0; m <= NWD; ++m)
            planes_of(wrow[(2 * (g + m)) * 64], wrow[(2 * (g + m) + 1) * 64], c0[m], c1[m]);
        const uint32_t rw0 = rrow[(2 * g) * 64], rw1 = rrow[(2 * g + 1) * 64];
#pragma unroll
        for (int j = 0; j < 32; ++j) {
            if ((j & 7) == 0) {
                bv[j >> 3] = v0;
#pragma unroll
                for (int m = 0; m < NWD; ++m) {
                    bp[j >> 3][m] = VPr[m];
                    bn[j >> 3][m] = VNr[m];
                }
            }
            uint32_t ne[NWD];
            bits_ne<NWD>(c0, c1, j < 16 ? rw0 : rw1, j, ne);
            bits_column<NWD>(VPr, VNr, ne, dcol[j], true);
            v0 += 1 - (int)(dcol[j][0] & 1u);
        }
    };
    // D0 bits of diagonal di over the segment's columns (bit j: column 32 rg + 1 + j);
    // 0 outside the band
    auto gather = [&](int di) -> uint32_t {
        if (di < 0 || di >= BMAX) return 0u;
        const uint32_t b = (uint32_t)(di & 31);
        const bool hi = NWD == 2 && di >= 32;
        // bit b of each column word into bit j: a bit-field extract and a shift-or
        // per column
        uint32_t x = 0;
#pragma unroll
        for (int j = 0; j < 32; ++j)
            x = (__builtin_amdgcn_ubfe(hi ? dcol[j][NWD - 1] : dcol[j][0], b, 1u) << j) | x;
        return x;
    };
    // the walk goes by segments of 32 columns (4 blocks) along its diagonal: the D0
    // bits of the diagonal in columns 32g+1 .. 32g+32 and the mismatches of those
    // cells, bit j for column 32g + 1 + j; a diagonal run crosses a segment in one
    // step and waits for memory once per segment
    int ci = -1, cg = -1;
    uint32_t dseg = 0, mseg = 0;
    uint32_t r0 = 0, r1 = 0;   // read codes of segment cg (rows 2cg, 2cg + 1)
    int wl0 = INT32_MIN;       // window row of wg0 (wg0 .. wg2: three window rows from it); none yet
    uint32_t wg0 = 0, wg1 = 0, wg2 = 0;
    const int qlast = (L - 2) >> 3;   // last block holding columns of this read
    // the side cells of a non-diagonal step (block sq of diagonal si): D0 bits of
    // diagonals si - 1 and si + 1 and the block's boundary
    int si = -1, sq = -1;
    uint32_t dm = 0, dp = 0;
    int bvt = 0;                // value of band cell 0 in the boundary column
    uint32_t bvp[NWD], bvn[NWD];
#pragma unroll
    for (int m = 0; m < NWD; ++m) bvp[m] = bvn[m] = 0u;
    auto dbyte = [&](const uint32_t *__restrict__ blk, int i) -> uint32_t {
        if (i < 0 || i >= BMAX) return 0u;
        int wd, sh;
        dloc<NW>(i, wd, sh);
        return (blk[wd * 64] >> sh) & 0xFFu;
    };
    // value of band cell i in column 8sq + nb (nb = 0 .. 8) of the side block;
    // K outside the band (never equal to a value <= k)
    auto cell = [&](int i, uint32_t dbits, int nb) -> int {
        if (i < 0 || i >= BMAX) return K;
        int base = bvt;
#pragma unroll
        for (int m = 0; m < NWD; ++m) {
            // cells 1 .. i of word m
            const int top = i - 32 * m;
            uint32_t msk = top < 0 ? 0u : top >= 31 ? ~0u : (2u << top) - 1u;
            if (m == 0) msk &= ~1u;
            base += __builtin_popcount(bvp[m] & msk) - __builtin_popcount(bvn[m] & msk);
        }
        return base + nb - __builtin_popcount(dbits & ((1u << nb) - 1u));
    };
    while (!done) {
        if (stop_row0 && row == 0) break;
        if (col == 0 && L >= 2) {
            // column 0 (Q7, Q4, Q5) in one step: decide() takes the top move while
            // Tp + 1 - corr == M[row][0] = max(0, row - 2s), with corr = 1 for rows
            // <= 2s + 3 when s > 0.  So it climbs to row 0 from rows <= 2s (every row
            // when s = 0), stops at once on rows 2s+1 .. 2s+3, climbs to row 2s + 3
            // from lower rows; then it stops (row 0 has no top cell)
            int T = (s == 0 || row <= 2 * s) ? row : imax(row - (2 * s + 3), 0);
            T = imin(T, cap - n);
            while (T > 0) {
                const int c = imin(T, 16 - (n & 15));
                const uint32_t tops = c == 16 ? 0x55555555u : (1u << (2 * c)) / 3u;
                opw |= tops << ((n & 15) * 2);
                n += c;
                row -= c;
                T -= c;
                if ((n & 15) == 0) {
                    put_word((n >> 4) - 1, opw);
                    opw = 0;
                }
            }
            break;
        }
        if (col == 0) {
            // column 0: M[h][0] = max(0, h - 2s) (Q7); a vertical step or the stop
            // remain (Dg = Lf = COL0)
            const int vv = imax(row - 2 * s, 0);
            const int Tp = row >= 1 ? imax(row - 1 - 2 * s, 0) : SRM_FOREIGN;
            const int op = decide(0, 1, vv, SRM_COL0, SRM_COL0, Tp, 0, imax(row - 1, 0), L, w,
                                  s, e, false);
            if (op == 3 || n >= cap) break;
            emit(op);
            continue;
        }
        const int i = row - col - dlo;
        const int g = (col - 1) >> 5;
        int j0 = (col - 1) & 31;
        if (i != ci || g != cg) {
            // characters: read codes 32g .. 32g+31 (rows 2g, 2g+1) against window
            // codes from band cell 0's code + i + 32g (exact for rows and columns
            // >= 1, the only cells whose characters can decide)
            if (g != cg) {   // a new diagonal in the same segment keeps its read codes
                r0 = rrow[(2 * g) * 64];
                r1 = rrow[(2 * g + 1) * 64];
            }
            const int wp = imax(i, 0) + 32 * g;
            if ((wp >> 4) != wl0) {   // a diagonal next to the last one mostly shares them
                const uint32_t *__restrict__ ww = wrow + (size_t)(wp >> 4) * 64;
                wg0 = ww[0];
                wg1 = ww[64];
                wg2 = ww[128];
                wl0 = wp >> 4;
            }
            const uint32_t gsh = (uint32_t)(wp & 15) * 2u;
            const uint32_t x0 = r0 ^ __builtin_amdgcn_alignbit(wg1, wg0, gsh);
            const uint32_t x1 = r1 ^ __builtin_amdgcn_alignbit(wg2, wg1, gsh);
            mseg = even_bits(x0 | (x0 >> 1)) | (even_bits(x1 | (x1 >> 1)) << 16);
            // the D0 bits of the diagonal over the segment decide only where the run
            // below the walk's column meets differing characters: at equal ones D0
            // is 1 (a unit-cost cell is its diagonal predecessor or one more, and
            // equal characters give the predecessor).  A run without differing
            // characters is walked as all ones, nothing loaded or recomputed.
            if (i < 0 || i >= BMAX) {
                dseg = 0u;
            } else if (!(mseg & ((2u << j0) - 1u))) {
                dseg = ~0u;
            } else if (CK) {
                if (g != rg) {
                    recompute(g);
                    rg = g;
                }
                dseg = gather(i);
            } else {
                // D words of blocks 4g .. 4g+3 (clamped to the read's last block: bits
                // of columns past the walk's cell are never looked at)
                int wd = 0, sh = 0;
                dloc<NW>(i, wd, sh);
                dseg = 0u;
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    dseg |= ((scr[((size_t)imin(4 * g + u, qlast) * WB + wd) * 64 + lane] >> sh) &
                             0xFFu) << (8 * u);
            }
            ci = i;
            cg = g;
        }
        // cells j0, j0-1, .. of the segment step diagonally until one has D0 set
        // and differing characters
        const uint32_t below = (2u << j0) - 1u;   // bits 0 .. j0 (j0 = 31: all)
        const uint32_t fail = dseg & mseg & below;
        const int rf = fail ? j0 - (31 - __builtin_clz(fail)) : j0 + 1;
        const int rr = imin(rf, imin(row, cap - n));
        if (rr > 0) {
            // rr diagonal steps: 0 bits in the transcript; v drops by the mismatches
            const uint32_t run = below & ~((1u << (j0 + 1 - rr)) - 1u);
            v -= __builtin_popcount(~dseg & run);
            const int n2 = n + rr;
            if ((n2 >> 4) != (n >> 4)) {   // the run completes the current ops word
                put_word(n >> 4, opw);
                opw = 0;
                if ((n2 >> 4) > (n >> 4) + 1) put_word((n >> 4) + 1, 0u);   // and a whole one
            }
            n = n2;
            row -= rr;
            col -= rr;
            // a run that ends on its segment's failing cell takes that cell's step
            // now (the loop's next pass would come back here with the same words)
            if (!fail || rr != rf || (stop_row0 && row == 0)) continue;
            j0 -= rr;
        }
        // one cell that is not on a diagonal run (col >= 1)
        const int q = (col - 1) >> 3, jb = (col - 1) & 7;
        if (i != si || q != sq) {
            if (CK) {
                // block q is block u of the walk's segment, recomputed if it is not
                // the one held
                const int gq = q >> 2;
                if (gq != rg) {
                    recompute(gq);
                    rg = gq;
                }
                const int u = q - 4 * rg;
                dm = (gather(i - 1) >> (8 * u)) & 0xFFu;
                dp = (gather(i + 1) >> (8 * u)) & 0xFFu;
                bvt = u == 0 ? bv[0] : u == 1 ? bv[1] : u == 2 ? bv[2] : bv[3];
#pragma unroll
                for (int m = 0; m < NWD; ++m) {
                    bvp[m] = u == 0 ? bp[0][m] : u == 1 ? bp[1][m] : u == 2 ? bp[2][m] : bp[3][m];
                    bvn[m] = u == 0 ? bn[0][m] : u == 1 ? bn[1][m] : u == 2 ? bn[2][m] : bn[3][m];
                }
            } else {
                const uint32_t *__restrict__ blk = scr + (size_t)q * WB * 64 + lane;
                dm = dbyte(blk, i - 1);
                dp = dbyte(blk, i + 1);
                if (q != sq) {   // the boundary belongs to the block, not the diagonal
                    bvt = (int)blk[DW * 64];
#pragma unroll
                    for (int m = 0; m < NWD; ++m) {
                        bvp[m] = blk[(DW + 1 + m) * 64];
                        bvn[m] = blk[(DW + 1 + NWD + m) * 64];
                    }
                }
            }
            si = i;
            sq = q;
        }
        const int Dg = row >= 1 ? v - 1 + (int)((dseg >> j0) & 1u) : SRM_FOREIGN;
        const int Tp = row >= 1 ? cell(i - 1, dm, jb + 1) : SRM_FOREIGN;
        const int Lf = cell(i + 1, dp, jb);
        const bool eq = ((mseg >> j0) & 1u) == 0u;   // decides only with row >= 1
        const int op = decide(0, 1, v, Dg, Lf, Tp, col - 1, imax(row - 1, 0), L, w, s, e, eq);
        if (op == 3 || n >= cap) break;
        v = op == 0 ? Dg : op == 1 ? Tp : Lf;
        emit(op);
    }
    if (n & 15) put_word(n >> 4, opw);
    if (a.stage_words)
        for (int x = 0; x < ((n + 15) >> 4); ++x) ops[x] = stage[x * 64];
    r.n_ops = (uint16_t)n;
    r.end_row = (uint16_t)row;
    r.end_col = (uint16_t)col;
    a.res[t] = r;
}

// ---------------------------------------------------------------------------
// host-side launch wrappers
// ---------------------------------------------------------------------------
// words of one tile's scratch for the unit-cost kernels
// words of one tile's scratch for the unit-cost kernels (ck: checkpoint layout)
template <int NW>
static size_t bits_tile_words_nw(int maxL, bool ck)
{
    return (size_t)(ck ? BitsShape<NW>::ck_rows(maxL) : BitsShape<NW>::rows(maxL)) * 64;
}

size_t bits_tile_words(int nw, int maxL, bool ck)
{
    switch (nw) {
    case 1: return bits_tile_words_nw<1>(maxL, ck);
    case 2: return bits_tile_words_nw<2>(maxL, ck);
    case 3: return bits_tile_words_nw<3>(maxL, ck);
    case 4: return bits_tile_words_nw<4>(maxL, ck);
    case 5: return bits_tile_words_nw<5>(maxL, ck);
    case 6: return bits_tile_words_nw<6>(maxL, ck);
    case 7: return bits_tile_words_nw<7>(maxL, ck);
    default: return bits_tile_words_nw<8>(maxL, ck);
    }
}

// waves per block of the unit-cost kernels: 1, 2 or 4 (anything else is 4)
static int bits_block_waves(const BandArgs &a)
{
    return a.block_waves == 1 || a.block_waves == 2 ? a.block_waves : 4;
}

template <int NW>
static hipError_t launch_bits_band_nw(const BandArgs &a, int blocks, size_t lds, hipStream_t st)
{
    if (a.walk_ck)
        hipLaunchKernelGGL((bits_band_kernel<NW, true>), dim3(blocks), dim3(64 * bits_block_waves(a)), lds, st, a);
    else
        hipLaunchKernelGGL((bits_band_kernel<NW, false>), dim3(blocks), dim3(64 * bits_block_waves(a)), lds, st, a);
    return hipGetLastError();
}

// reads of up to 250 codes stage their codes in LDS (19 KiB per block of 4 waves,
// 8 blocks per CU); longer ones are read from read_words.  blocks: of
// a.block_waves waves, one wave per tile
hipError_t launch_bits_band(const BandArgs &args, int nw, int blocks, int maxL, hipStream_t st)
{
    BandArgs a = args;
    a.code_words = maxL <= 250 ? ((maxL + 15) >> 4) + 3 : 0;
    const size_t lds = (size_t)a.code_words * 64 * bits_block_waves(a) * 4;
    switch (nw) {
    case 1: return launch_bits_band_nw<1>(a, blocks, lds, st);
    case 2: return launch_bits_band_nw<2>(a, blocks, lds, st);
    case 3: return launch_bits_band_nw<3>(a, blocks, lds, st);
    case 4: return launch_bits_band_nw<4>(a, blocks, lds, st);
    case 5: return launch_bits_band_nw<5>(a, blocks, lds, st);
    case 6: return launch_bits_band_nw<6>(a, blocks, lds, st);
    case 7: return launch_bits_band_nw<7>(a, blocks, lds, st);
    case 8: return launch_bits_band_nw<8>(a, blocks, lds, st);
    default: return hipErrorInvalidValue;
    }
}

template <int NW>
static int bits_occupancy_nw()
{
    int blocks = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &blocks, (const void *)(bits_band_kernel<NW, false>), 256, 0);
    return e == hipSuccess && blocks > 0 ? blocks : 1;
}

int bits_band_occupancy(int nw)
{
    switch (nw) {
    case 1: return bits_occupancy_nw<1>();
    case 2: return bits_occupancy_nw<2>();
    case 3: return bits_occupancy_nw<3>();
    case 4: return bits_occupancy_nw<4>();
    case 5: return bits_occupancy_nw<5>();
    case 6: return bits_occupancy_nw<6>();
    case 7: return bits_occupancy_nw<7>();
    case 8: return bits_occupancy_nw<8>();
    default: return 1;
    }
}

template <int NW>
static hipError_t launch_bits_trace_nw(const BandArgs &a, hipStream_t st)
{
    const int bw = bits_block_waves(a);
    const size_t lds = (size_t)a.stage_words * 64 * bw * 4;
    if (a.walk_ck)
        hipLaunchKernelGGL((bits_trace_kernel<NW, true>), dim3((a.ntiles + bw - 1) / bw), dim3(64 * bw), lds, st, a);
    else
        hipLaunchKernelGGL((bits_trace_kernel<NW, false>), dim3((a.ntiles + bw - 1) / bw), dim3(64 * bw), lds, st, a);
    return hipGetLastError();
}

// the walk reads the codes from the tile's rows: no LDS
// transcript words per lane the walk stages in LDS for reads of up to maxL codes:
// a band candidate's transcript holds at most w + k <= L + 63 ops (band cells
// k + max(2s + k, w - L) + 1 <= 64), one more word for a run's lookahead write;
// none for longer reads (the LDS would cut the walk's occupancy)
int bits_trace_stage_words(int maxL) { return maxL <= 320 ? ((maxL + 63) >> 4) + 2 : 0; }

hipError_t launch_bits_trace(const BandArgs &args, int nw, int maxL, hipStream_t st)
{
    BandArgs a = args;
    a.stage_words = bits_trace_stage_words(maxL);
    switch (nw) {
    case 1: return launch_bits_trace_nw<1>(a, st);
    case 2: return launch_bits_trace_nw<2>(a, st);
    case 3: return launch_bits_trace_nw<3>(a, st);
    case 4: return launch_bits_trace_nw<4>(a, st);
    case 5: return launch_bits_trace_nw<5>(a, st);
    case 6: return launch_bits_trace_nw<6>(a, st);
    case 7: return launch_bits_trace_nw<7>(a, st);
    case 8: return launch_bits_trace_nw<8>(a, st);
    default: return hipErrorInvalidValue;
    }
}

} // namespace srm
