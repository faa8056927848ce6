// records_gpu.hip -- the output records formatted on the device.
//
// After a job has run, every hit (score <= k) becomes the TSV line of the
// reference's post-processing (src/cuda-functions.cu:828-994), byte for byte the
// line srm_format_record writes on the host:
//   rec_measure      one thread per candidate derives, from the transcript
//                    alone, the leading and closing gaps of the read row, the
//                    indel scan, the printed length and the start/stop, and the
//                    line length;
//   (scan)           exclusive sum of the line lengths = each line's offset;
//   rec_write_lines  eight lines per wave, eight lanes per line, for the lines of
//                    up to 320 printed positions: name, coordinates, both rows and
//                    the score, built in LDS and copied out by whole words;
//   rec_write        one wave per line for the rest (long lines, raw-byte
//                    candidates).
// The text lands in candidate order in one HBM buffer that the host copies out
// in one transfer.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "nw_device.h"
#include "rec_ops.h"

namespace srm {

namespace {

constexpr int RB = 256;

// characters %d prints for v
__device__ __forceinline__ int ndig(int v)
{
    const uint32_t u = v < 0 ? 0u - (uint32_t)v : (uint32_t)v;
    return 1 + (v < 0) + (u >= 10u) + (u >= 100u) + (u >= 1000u) + (u >= 10000u) +
           (u >= 100000u) + (u >= 1000000u) + (u >= 10000000u) + (u >= 100000000u) +
           (u >= 1000000000u);
}

__device__ __forceinline__ int name_len(const RecArgs &a, int32_t id)
{
    if (a.names) {
        const char *s = a.names + (int64_t)id * a.name_stride;
        int n = 0;
        while (n < a.name_stride - 1 && s[n]) ++n;
        return n;
    }
    return a.prefix_len + ndig(id);
}

// line sizes of rec_write_lines (below), which writes the lines grouped_line() accepts
constexpr int QL = 8;                // lines per wave
constexpr int TPL = 8;               // lanes per line
constexpr int CHUNK = 40;            // printed positions per lane at most
constexpr int MAXP = TPL * CHUNK;    // printed positions per line (320)
constexpr int LCAP = 704;            // text bytes per line
constexpr int WPL = MAXP / 16 + 3;   // staged words per line: transcript, read codes, window codes
constexpr int NAMEW = 64;            // staged name words per line (name_stride <= 256)

__device__ __forceinline__ bool grouped_line(uint64_t len, uint32_t lead_rn, uint32_t bits)
{
    const int endp = (int)(lead_rn & 0xFFFFu) + (int)(lead_rn >> 16);
    return len > 0 && len <= (uint64_t)(LCAP - 4) && endp <= MAXP && !(bits & SRM_CAND_BYTES);
}

// leading gaps, closing gaps, indel scan and printed length of one hit
// (cuda-functions.cu:866-980), a transcript word at a time; returns the line length
// and whether rec_write_lines writes the line (else rec_write does)
__device__ uint64_t measure_one(const RecArgs &a, int t, bool &grouped)
{
    const srm_cand c = a.cand[t];
    const srm_result r = a.res[t];
    const int n = r.n_ops, w = c.w, L = c.L;
    const uint32_t *ops = a.ops + c.ops_off;
    const srm_rec::Measure m = srm_rec::measure_ops_by([ops](int W) { return ops[W]; }, n, w);
    const int lead = m.lead, del = m.del, ins = m.ins, eor = m.eor, trail = m.trail;
    const short chars = eor > 0 ? (short)(eor - trail) : (short)(L + del);   // :972-980 (Q8)
    const int cap = w + (int)c.k;
    const int lim = chars < 0 ? cap : chars;
    int rn = n - lead < lim ? n - lead : lim;                    // %.*s stops at NUL (Q9)
    if (rn < 0) rn = 0;
    const uint32_t start = (uint32_t)a.pos[t] + (uint32_t)lead;
    const int32_t stop = (int32_t)((uint32_t)(int32_t)chars - (uint32_t)ins + start - 1u);
    a.info[t] = make_uint4(start, (uint32_t)stop, (uint32_t)lead | ((uint32_t)rn << 16), 0u);
    const int32_t id = a.read_id ? a.read_id[t] : t;
    const uint64_t len = (uint64_t)(name_len(a, id) + ndig((int32_t)start) + ndig(stop) + 2 +
                                    2 * rn + ndig((int)r.score) + 7);
    grouped = grouped_line(len, (uint32_t)lead | ((uint32_t)rn << 16), c.bits);
    return len;
}

// a grid that strides over the candidates; the hits (hits[0]) and the lines left
// to rec_write (hits[1]) are summed per block and added with one atomic each per
// block (the counters are words that every block would otherwise update once per
// wave)
__global__ void __launch_bounds__(RB) rec_measure(RecArgs a)
{
    __shared__ uint32_t s_hits[RB / 64][2];
    uint32_t hits = 0, longs = 0;
    for (int t = blockIdx.x * RB + threadIdx.x; t <= a.n; t += gridDim.x * RB) {
        if (t == a.n) {
            a.len[t] = 0;                           // the scan's tail element
            continue;
        }
        const bool hit = a.res[t].score <= (int)a.cand[t].k;
        bool grouped = true;
        a.len[t] = hit ? measure_one(a, t, grouped) : 0;
        hits += hit;
        longs += hit && !grouped;
    }
    // block sums: waves by reduction in registers, then the block's waves in LDS
    for (int d = 32; d >= 1; d >>= 1) {
        hits += (uint32_t)__shfl_xor((int)hits, d, 64);
        longs += (uint32_t)__shfl_xor((int)longs, d, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        s_hits[threadIdx.x >> 6][0] = hits;
        s_hits[threadIdx.x >> 6][1] = longs;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t b = 0, l = 0;
        for (int w = 0; w < RB / 64; ++w) {
            b += s_hits[w][0];
            l += s_hits[w][1];
        }
        if (b) atomicAdd(a.hits, b);
        if (l) atomicAdd(a.hits + 1, l);
    }
}

// ---------------------------------------------------------------------------
// rec_write: one wave per line.  The line is assembled in LDS at the output's
// byte alignment (so that LDS dword i and output dword i cover the same bytes),
// then copied out with dword stores, byte stores only where the first and last
// dwords are shared with the neighbouring lines.
//   header  name \t start \t stop \t << or >> \t   one character per lane
//   rows    printed positions in chunks of 16 per lane: the chunk's row and column
//           advances, a wave scan of them, then the chunk's characters from two
//           words of read codes and two of window codes
//   tail    \t score \n after the genome row
// The longest line: a 266-byte name, 2 x 11 coordinate characters, 2 x (w + k)
// <= 2 x 1155 row characters, a 6-character score and 9 separators -- under
// LINE_LDS with the alignment offset.
constexpr int LINE_LDS = 3072;

// 10^e for e in 0..9, in registers (a table would be a memory load per digit)
__device__ __forceinline__ uint32_t pow10u(int e)
{
    uint32_t p = 1u;
    for (int i = 0; i < e; ++i) p *= 10u;
    return p;
}

// character i of v printed with nd characters (%d)
__device__ __forceinline__ char digit_at(int32_t v, int nd, int i)
{
    if (v < 0) {
        if (i == 0) return '-';
        return (char)('0' + ((0u - (uint32_t)v) / pow10u(nd - 1 - i)) % 10u);
    }
    return (char)('0' + ((uint32_t)v / pow10u(nd - 1 - i)) % 10u);
}

// 'A', 'C', 'G', 'T' for codes 0..3 from one constant (no table in memory)
__device__ __forceinline__ char base_char(uint32_t code)
{
    return (char)((0x54474341u >> (8u * code)) & 0xFFu);
}

__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t x, int lane, uint32_t &total)
{
    uint32_t s = x;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(s, d, 64);
        if (lane >= d) s += y;
    }
    total = __shfl(s, 63, 64);
    return s - x;
}

__device__ __forceinline__ uint32_t lane_u32(uint32_t v, int j)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, j);
}

__global__ void __launch_bounds__(RB) rec_write(RecArgs a)
{
    __shared__ uint32_t lds[RB / 64][LINE_LDS / 4];
    if (a.hits[1] == 0) return;   // no line of this kind (counted by rec_measure)
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    char *buf = (char *)lds[wv];
    const int nwaves = gridDim.x * (RB / 64);
    const int ntile = (a.n + 63) / 64;
    // a wave takes 64 consecutive candidates at a time: their metadata in one
    // coalesced load per field (a candidate per lane), then their lines one by one
    for (int tile = __builtin_amdgcn_readfirstlane(blockIdx.x * (RB / 64) + wv); tile < ntile;
         tile += nwaves) {
        const int tl = tile * 64 + lane;
        const bool in = tl < a.n;
        const uint64_t len_l = in ? a.len[tl] : 0;
        if (!__ballot(len_l != 0)) continue;
        const uint4 inf_l = len_l ? a.info[tl] : make_uint4(0u, 0u, 0u, 0u);
        uint32_t cw_l[5] = {0u, 0u, 0u, 0u, 0u}, rw_l[2] = {0u, 0u};
        if (len_l) {
            const uint32_t *cp = (const uint32_t *)(a.cand + tl);
            const uint32_t *rp = (const uint32_t *)(a.res + tl);
            for (int i = 0; i < 5; ++i) cw_l[i] = cp[i];
            rw_l[0] = rp[0];
            rw_l[1] = rp[1];
        }
        // the lines rec_write_lines leaves: long ones and raw-byte candidates
        uint64_t hits = __ballot(len_l != 0 && !grouped_line(len_l, inf_l.z, cw_l[4] >> 16));
        if (!hits) continue;
        const uint64_t off_l = len_l ? a.off[tl] : 0;
        const int32_t id_l = in ? (a.read_id ? a.read_id[tl] : tl) : 0;
        while (hits) {
            const int j = (int)__builtin_ctzll(hits);
            hits &= hits - 1;
            const uint64_t len = (uint64_t)lane_u32((uint32_t)len_l, j) |
                                 ((uint64_t)lane_u32((uint32_t)(len_l >> 32), j) << 32);
            const uint64_t dst = (uint64_t)lane_u32((uint32_t)off_l, j) |
                                 ((uint64_t)lane_u32((uint32_t)(off_l >> 32), j) << 32);
            const int al = (int)(dst & 3u);
            srm_cand c;
            {
                uint32_t cw[5];
                for (int i = 0; i < 5; ++i) cw[i] = lane_u32(cw_l[i], j);
                __builtin_memcpy(&c, cw, sizeof c);
            }
            srm_result r;
            {
                uint32_t rw2[2] = {lane_u32(rw_l[0], j), lane_u32(rw_l[1], j)};
                __builtin_memcpy(&r, rw2, sizeof r);
            }
            const uint4 inf = make_uint4(lane_u32(inf_l.x, j), lane_u32(inf_l.y, j),
                                         lane_u32(inf_l.z, j), 0u);
            const int lead = (int)(inf.z & 0xFFFFu), rn = (int)(inf.z >> 16);
            const int32_t id = (int32_t)lane_u32((uint32_t)id_l, j);
            const int32_t start = (int32_t)inf.x, stop = (int32_t)inf.y;
            const int nds = ndig(start), ndt = ndig(stop), ndi = ndig(id), ndsc = ndig((int)r.score);
            const int name_n = a.names ? name_len(a, id) : a.prefix_len + ndi;
            const int hdr = name_n + 1 + nds + 1 + ndt + 1 + 2 + 1;
            char *line = buf + al;
            // ---- header
            for (int h = lane; h < hdr; h += 64) {
                char ch;
                int q = h;
                if (q < name_n) {
                    if (a.names)
                        ch = a.names[(int64_t)id * a.name_stride + q];
                    else
                        ch = q < a.prefix_len ? a.prefix[q] : digit_at(id, ndi, q - a.prefix_len);
                } else if ((q -= name_n) == 0) {
                    ch = '\t';
                } else if ((q -= 1) < nds) {
                    ch = digit_at(start, nds, q);
                } else if ((q -= nds) == 0) {
                    ch = '\t';
                } else if ((q -= 1) < ndt) {
                    ch = digit_at(stop, ndt, q);
                } else if ((q -= ndt) == 0 || q == 3) {
                    ch = '\t';
                } else {
                    ch = c.flag > 3 ? '<' : '>';
                }
                line[h] = ch;
            }
            // ---- both rows: printed position p holds transcript op n - 1 - p
            char *orow = line + hdr;
            char *grow = orow + rn + 1;
            const uint32_t *ops = a.ops + c.ops_off;
            const bool bytes = (c.bits & SRM_CAND_BYTES) != 0;
            const int n = r.n_ops;
            const int endp = lead + rn;
            uint32_t row0 = r.end_row, col0 = r.end_col;
            for (int base = 0; base < endp; base += 64 * 16) {
                const int p0 = base + lane * 16;
                const int cnt = endp - p0 < 0 ? 0 : (endp - p0 > 16 ? 16 : endp - p0);
                // the chunk's ops, position-ordered, 2 bits each: ops j0 = n-1-p0 down to
                // j0-cnt+1, from the one or two transcript words that hold them
                uint32_t chunk = 0;
                if (cnt > 0) {
                    const int j0 = n - 1 - p0, jm = j0 - (cnt - 1);
                    const int wlo = jm >> 4;
                    uint64_t v = ops[wlo];
                    if ((j0 >> 4) > wlo) v |= (uint64_t)ops[wlo + 1] << 32;
                    for (int i = 0; i < cnt; ++i)
                        chunk |= (uint32_t)((v >> (2 * (j0 - i - 16 * wlo))) & 3u) << (2 * i);
                }
                uint32_t drow = 0, dcol = 0;
                for (int i = 0; i < cnt; ++i) {
                    const uint32_t op = (chunk >> (2 * i)) & 3u;
                    drow += op != SRM_OP_LEFT;
                    dcol += op != SRM_OP_TOP;
                }
                uint32_t trow, tcol;
                uint32_t row = row0 + wave_excl_scan(drow, lane, trow);
                uint32_t col = col0 + wave_excl_scan(dcol, lane, tcol);
                if (cnt > 0) {
                    uint64_t rw = 0, gw = 0;
                    int rsh = 0, gsh = 0;
                    if (!bytes) {   // the chunk's codes span at most two words each
                        const int64_t ri = (int64_t)c.read_off + col, gi = (int64_t)c.win_off + row;
                        rw = (uint64_t)a.read_words[ri >> 4] | ((uint64_t)a.read_words[(ri >> 4) + 1] << 32);
                        gw = (uint64_t)a.win_words[gi >> 4] | ((uint64_t)a.win_words[(gi >> 4) + 1] << 32);
                        rsh = (int)(ri & 15) * 2;
                        gsh = (int)(gi & 15) * 2;
                    }
                    for (int i = 0; i < cnt; ++i) {
                        const uint32_t op = (chunk >> (2 * i)) & 3u;
                        const int p = p0 + i;
                        if (p >= lead) {
                            char rc = '_', gc = '_';
                            if (op != SRM_OP_TOP)
                                rc = bytes ? a.read_bytes[(int64_t)c.read_off + col]
                                           : base_char((uint32_t)(rw >> rsh) & 3u);
                            if (op != SRM_OP_LEFT)
                                gc = bytes ? a.win_bytes[(int64_t)c.win_off + row]
                                           : base_char((uint32_t)(gw >> gsh) & 3u);
                            orow[p - lead] = rc;
                            grow[p - lead] = gc;
                        }
                        if (op != SRM_OP_LEFT) {
                            ++row;
                            gsh += 2;
                        }
                        if (op != SRM_OP_TOP) {
                            ++col;
                            rsh += 2;
                        }
                    }
                }
                row0 += trow;
                col0 += tcol;
            }
            // ---- tail: \t between the rows, then \t score \n
            if (lane == 0) orow[rn] = '\t';
            char *tail = grow + rn;
            if (lane <= ndsc + 1)
                tail[lane] = lane == 0 ? '\t' : lane == ndsc + 1 ? '\n' : digit_at((int)r.score, ndsc, lane - 1);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // ---- copy out: LDS dword i is output dword (dst - al) / 4 + i
            uint32_t *o32 = (uint32_t *)(a.out + (dst - (uint64_t)al));
            const int end = al + (int)len;
            const int nd = (end + 3) >> 2;
            for (int i = lane; i < nd; i += 64) {
                const uint32_t v = lds[wv][i];
                const int b0 = 4 * i;
                if (b0 >= al && b0 + 4 <= end) {
                    o32[i] = v;
                } else {
                    char *o8 = (char *)(o32 + i);
                    for (int b = 0; b < 4; ++b)
                        if (b0 + b >= al && b0 + b < end) o8[b] = (char)(v >> (8 * b));
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
}

// ---------------------------------------------------------------------------
// rec_write_lines: QL = 8 lines per wave at a time, TPL = 8 lanes per line.  A
// wave takes 64 consecutive candidates (their length, offset and metadata in one
// coalesced load per field, a candidate per lane) and writes the lines
// grouped_line() accepts, eight at a time:
//   stage   each line's lanes load the transcript words and the read and window
//           code words its rows need (and a name from the table) into the line's
//           LDS slot: the eight lines' loads are in flight together
//   build   header and tail (the four numbers a lane each), then the rows into
//           the slot at the output's byte alignment: the positions split over the
//           line's lanes in runs of a multiple of 4 (at most CHUNK each), placed by
//           a scan of the runs' row and column advances, four characters per row
//           at a time
//   copy    the line's lanes store it with dwords, bytes only in the first and last
//           dwords, which the neighbouring lines share
// Every step of a line stays within its wave: no block barrier.  Long lines and
// raw-byte candidates are rec_write's.
constexpr int SLOT_TEXT = LCAP / 4 + 1;    // text words per slot (a line at its alignment)

__device__ __forceinline__ uint32_t from_lane(uint32_t v, int src)
{
    return (uint32_t)__shfl((int)v, src, 64);
}

__device__ __forceinline__ uint64_t bitrange64(int lo, int hi)   // bits [lo, hi) of 0..63
{
    if (hi <= lo) return 0ull;
    const uint64_t top = hi >= 64 ? ~0ull : ((1ull << hi) - 1ull);
    return top & ~((1ull << lo) - 1ull);
}

// four 2-bit fields from field f of a staged word array (word f >> 4 and the next)
__device__ __forceinline__ uint32_t fields4(const uint32_t *w, int f)
{
    const int x = f >> 4;
    const uint64_t v = (uint64_t)w[x] | ((uint64_t)w[x + 1] << 32);
    return (uint32_t)(v >> (2 * (f & 15))) & 0xFFu;
}

// four 2-bit fields (x < 256) spread into the four bytes, field t in byte t
__device__ __forceinline__ uint32_t spread4(uint32_t x)
{
    const uint32_t t = (x | (x << 12)) & 0x000F000Fu;
    return (t | (t << 6)) & 0x03030303u;
}

// the four characters of one row at four positions: gaps has 1 in byte t when
// position t is a gap in this row ('_'), codes8 the row's next four codes (code
// k in field k); position t prints code t - (gaps before t)
__device__ __forceinline__ uint32_t row_chars4(uint32_t codes8, uint32_t gaps)
{
    const uint32_t ch = __builtin_amdgcn_perm(0u, 0x54474341u, spread4(codes8));   // "ACGT"
    // gaps before byte t, t <= 3: bytes 0..2 of gaps summed into the bytes above
    const uint32_t before = __umul24(gaps, 0x010101u) << 8;
    // selector byte t: code byte t - before (0..3) or, for a gap, byte 4 of the
    // pair = '_'
    const uint32_t sel = 0x03020100u + (gaps << 2) - before;
    return __builtin_amdgcn_perm(0x5F5F5F5Fu, ch, sel);
}

// a row's four characters c that start at slot byte x into the slot's words:
// the word of x gets c's first 4 - (x & 3) bytes, after the last (x & 3) bytes
// of prev (the four characters before) when pfull, else alone as bytes
__device__ __forceinline__ void put_head(uint32_t *s32, int x, uint32_t c, uint32_t prev, bool pfull)
{
    const int d = x >> 2, sh = x & 3;
    if (sh == 0 || pfull) {
        s32[d] = sh == 0 ? c : __builtin_amdgcn_alignbyte(c, prev, (uint32_t)(4 - sh));
    } else {
        char *s8 = (char *)s32 + x;
        for (int b = 0; b < 4 - sh; ++b) s8[b] = (char)(c >> (8 * b));
    }
}

// the last (x & 3) bytes of the four characters c that start at slot byte x
__device__ __forceinline__ void put_tail(uint32_t *s32, int x, uint32_t c)
{
    const int sh = x & 3;
    char *s8 = (char *)s32 + (x - sh + 4);
    for (int b = 0; b < sh; ++b) s8[b] = (char)(c >> (8 * (4 - sh + b)));
}

// v as %d prints it, nd = ndig(v) characters at o
__device__ __forceinline__ void put_dec(char *o, int32_t v, int nd)
{
    uint32_t x = v < 0 ? 0u - (uint32_t)v : (uint32_t)v;
    int e = nd;
    // two digits a step: r = x mod 100, its tens digit (r * 103) >> 10 (exact for r < 100)
    while (x >= 100u) {
        const uint32_t d = x / 100u, r = x - 100u * d, t = (r * 103u) >> 10;
        o[--e] = (char)('0' + (r - 10u * t));
        o[--e] = (char)('0' + t);
        x = d;
    }
    if (x >= 10u) {
        const uint32_t t = (x * 103u) >> 10;
        o[--e] = (char)('0' + (x - 10u * t));
        o[--e] = (char)('0' + t);
    } else {
        o[--e] = (char)('0' + x);
    }
    if (v < 0) o[0] = '-';
}

template <bool NAMES>   // names from a table (else prefix + id)
__global__ void __launch_bounds__(RB) __attribute__((amdgpu_waves_per_eu(5))) rec_write_lines(RecArgs a)
{
    __shared__ uint32_t s_text[RB / 64][QL][SLOT_TEXT];
    // transcript words staged after one zero word (fields below the staged ones read 0)
    __shared__ uint32_t s_ops[RB / 64][QL][WPL + 1], s_rc[RB / 64][QL][WPL], s_gc[RB / 64][QL][WPL];
    // a name table's names staged per line (not needed for prefix + id: the smaller
    // block leaves room for a fifth block on a CU)
    __shared__ uint32_t s_name[NAMES ? RB / 64 : 1][NAMES ? QL : 1][NAMES ? NAMEW : 1];
    __shared__ char s_prefix[256];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int q = lane / TPL, u = lane % TPL;      // this lane's line slot and place in it
    const int sb = lane & ~(TPL - 1);              // the slot's first lane
    if (!NAMES)
        for (int i = tid; i < a.prefix_len; i += RB) s_prefix[i] = a.prefix[i];
    if (tid < (RB / 64) * QL) s_ops[tid / QL][tid % QL][0] = 0u;
    __syncthreads();
    const int nwaves = gridDim.x * (RB / 64);
    const int ntile = (a.n + 63) / 64;
    const uint64_t e64 = 0x5555555555555555ull;
    for (int tile = __builtin_amdgcn_readfirstlane(blockIdx.x * (RB / 64) + wv); tile < ntile;
         tile += nwaves) {
        // ---- the tile's candidates, one per lane
        const int tl = tile * 64 + lane;
        const bool in = tl < a.n;
        const uint64_t len_l = in ? a.len[tl] : 0;
        if (!__ballot(len_l != 0)) continue;
        uint32_t cw[5] = {0u, 0u, 0u, 0u, 0u}, rw[2] = {0u, 0u};
        uint4 inf = make_uint4(0u, 0u, 0u, 0u);
        uint64_t off_l = 0;
        int32_t id_l = 0;
        if (len_l) {
            const uint32_t *cp = (const uint32_t *)(a.cand + tl);
            const uint32_t *rp = (const uint32_t *)(a.res + tl);
            for (int i = 0; i < 5; ++i) cw[i] = cp[i];
            rw[0] = rp[0];
            rw[1] = rp[1];
            inf = a.info[tl];
            off_l = a.off[tl];
            id_l = a.read_id ? a.read_id[tl] : tl;
        }
        const bool mine = len_l != 0 && grouped_line(len_l, inf.z, cw[4] >> 16);
        const uint64_t todo = __ballot(mine);
        const int count = (int)__popcll(todo);
        // the tile's lines in order of printed positions, so that the QL lines of a
        // pass take about as many row steps and copy words each (a pass lasts as
        // long as its longest line): a bitonic sort of (positions << 6 | lane) over
        // the wave, the other candidates last
        const uint32_t endp_l = (inf.z & 0xFFFFu) + (inf.z >> 16);
        uint32_t order = ((mine ? (endp_l < 1022u ? endp_l : 1022u) : 1023u) << 6) | (uint32_t)lane;
        if (count > QL) {
            for (int k = 2; k <= 64; k <<= 1)
                for (int jx = k >> 1; jx > 0; jx >>= 1) {
                    const uint32_t o = (uint32_t)__shfl_xor((int)order, jx, 64);
                    const bool keep_min = ((lane & jx) == 0) == ((lane & k) == 0);
                    order = keep_min ? (o < order ? o : order) : (o > order ? o : order);
                }
        }
        for (int base = 0; base < count; base += QL) {
            // the next QL lines: the owner lane of this lane's slot
            int owner = -1;
            if (count > QL) {
                const int si = base + q;
                const uint32_t pick = from_lane(order, si);
                owner = si < count ? (int)(pick & 63u) : -1;
            } else {
                uint64_t t = todo;   // one pass, in lane order
                for (int g = 0; g < QL && t; ++g) {
                    const int h = (int)__builtin_ctzll(t);
                    t &= t - 1;
                    if (g == q) owner = h;
                }
            }
            const bool act = owner >= 0;
            const int src = act ? owner : 0;
            // the line's metadata from its owner
            const uint32_t read_off = from_lane(cw[0], src), win_off = from_lane(cw[1], src);
            const uint32_t ops_off = from_lane(cw[2], src), lw = from_lane(cw[3], src);
            const uint32_t fl = from_lane(cw[4], src);
            const uint32_t r0w = from_lane(rw[0], src), r1w = from_lane(rw[1], src);
            const int32_t start = (int32_t)from_lane(inf.x, src), stop = (int32_t)from_lane(inf.y, src);
            const uint32_t leadrn = from_lane(inf.z, src);
            const uint64_t loff = (uint64_t)from_lane((uint32_t)off_l, src) |
                                  ((uint64_t)from_lane((uint32_t)(off_l >> 32), src) << 32);
            const uint32_t llen = from_lane((uint32_t)len_l, src);
            const int32_t lid = (int32_t)from_lane((uint32_t)id_l, src);
            const int L = (int)(lw & 0xFFFFu), w = (int)(lw >> 16);
            const int flag = (int)(fl & 0xFFu);
            const int score = (int)(int16_t)(r0w & 0xFFFFu), n = (int)(r0w >> 16);
            const int end_row = (int)(r1w & 0xFFFFu), end_col = (int)(r1w >> 16);
            const int lead = (int)(leadrn & 0xFFFFu), rn = (int)(leadrn >> 16);
            const int endp = lead + rn;
            const int wlo = endp > 0 ? (n - endp) >> 4 : 0;
            const int nw_ops = endp > 0 ? ((n - 1) >> 4) - wlo + 1 : 0;
            const int rlo = (int)((read_off + (uint32_t)end_col) >> 4);
            const int glo = (int)((win_off + (uint32_t)end_row) >> 4);
            const int rnw = (int)((read_off + (uint32_t)min(end_col + endp, L)) >> 4) - rlo + 1;
            const int gnw = (int)((win_off + (uint32_t)min(end_row + endp, w)) >> 4) - glo + 1;
            uint32_t *sops = s_ops[wv][q] + 1, *src_ = s_rc[wv][q], *sgc = s_gc[wv][q];
            // ---- stage: the loads of the line's transcript, read and window words,
            // all issued at once; they land in LDS after the header is built, which
            // needs none of them (each array holds at most 21 words for a line of at
            // most MAXP positions: three words per lane)
            const bool staged = act && endp > 0;
            uint32_t xo[3] = {0u, 0u, 0u}, xr[3] = {0u, 0u, 0u}, xg[3] = {0u, 0u, 0u};
            if (staged) {
                const uint32_t *ops = a.ops + ops_off;
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    const int k = u + TPL * t;
                    xo[t] = k < nw_ops ? ops[wlo + k] : 0u;
                    xr[t] = k < rnw ? a.read_words[rlo + k] : 0u;
                    xg[t] = k < gnw ? a.win_words[glo + k] : 0u;
                }
            }
            if (NAMES) {   // a table name goes to LDS first: the header copies it
                if (act) {
                    const char *nm = a.names + (int64_t)lid * a.name_stride;
                    char *sn = (char *)s_name[NAMES ? wv : 0][NAMES ? q : 0];
                    for (int i = u; i < a.name_stride; i += TPL) sn[i] = nm[i];
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
            // ---- build: the header
            const int al = (int)(loff & 3u);
            char *line = (char *)s_text[wv][q] + al;
            int hdr = 0;
            if (act) {
                // the line's four numbers, a lane each: start, stop, the read id (when
                // names are prefix + id) and the score; their widths to every lane
                const int32_t num = u == 0 ? start : u == 1 ? stop : u == 2 ? lid : score;
                const int ndn = ndig(num);
                const int nds = __shfl(ndn, sb, 64), ndt = __shfl(ndn, sb + 1, 64);
                const int ndi = __shfl(ndn, sb + 2, 64), ndsc = __shfl(ndn, sb + 3, 64);
                const char *sname = (const char *)s_name[NAMES ? wv : 0][NAMES ? q : 0];
                int name_n;
                if (NAMES) {
                    name_n = 0;
                    while (name_n < a.name_stride - 1 && sname[name_n]) ++name_n;
                } else {
                    name_n = a.prefix_len + ndi;
                }
                hdr = name_n + 1 + nds + 1 + ndt + 1 + 2 + 1;
                // header: name \t start \t stop \t << or >> \t: the name's characters
                // (those of the prefix when names are prefix + id) over the lanes, the
                // separators and direction characters a lane each, the digits below
                const int nn = NAMES ? name_n : a.prefix_len;
                for (int h = u; h < nn; h += TPL) line[h] = NAMES ? sname[h] : s_prefix[h];
                {
                    const char dir = flag > 3 ? '<' : '>';
                    const int x = name_n + 3 + nds + ndt;   // the first direction character
                    if (u == 4) {
                        line[name_n] = '\t';
                        line[hdr + rn] = '\t';                    // between the rows
                    } else if (u == 5) {
                        line[name_n + 1 + nds] = '\t';
                        line[hdr + 2 * rn + 1] = '\t';            // before the score
                    } else if (u == 6) {
                        line[name_n + 2 + nds + ndt] = '\t';
                        line[hdr + 2 * rn + 2 + ndsc] = '\n';
                    } else if (u == 7) {
                        line[x] = dir;
                        line[x + 1] = dir;
                        line[x + 2] = '\t';
                    }
                }
                // the digits: each of the four lanes prints its number as %d does
                {
                    int at = -1;
                    if (u == 0) at = name_n + 1;
                    else if (u == 1) at = name_n + 2 + nds;
                    else if (u == 2 && !NAMES) at = a.prefix_len;
                    else if (u == 3) at = hdr + 2 * rn + 2;
                    if (at >= 0) put_dec(line + at, num, ndn);
                }
            }
            // ---- stage: the staged words into LDS
            if (staged) {
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    const int k = u + TPL * t;
                    if (k < nw_ops) sops[k] = xo[t];
                    if (k < rnw) src_[k] = xr[t];
                    if (k < gnw) sgc[k] = xg[t];
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // ---- build: the rows
            if (act) {
                // the line's positions split over its lanes in runs of a multiple of 4
                // (at most CHUNK each)
                const int chunk = (((endp + TPL - 1) / TPL) + 3) & ~3;
                const int p0 = u * chunk;
                const int cnt = endp - p0 < 0 ? 0 : (endp - p0 > chunk ? chunk : endp - p0);
                uint32_t drow = 0, dcol = 0;
                if (cnt > 0) {
                    // advances of the run: ops other than LEFT move the row, ops other
                    // than TOP the column, counted over its ops [jm, jm + cnt), at most
                    // 40 fields from field jm & 15 of staged word (jm >> 4) - wlo
                    const int jm = n - 1 - p0 - (cnt - 1);
                    const int x = (jm >> 4) - wlo;
                    const uint32_t o0 = sops[x];
                    const uint32_t o1 = x + 1 < nw_ops ? sops[x + 1] : 0u;
                    const uint32_t o2 = x + 2 < nw_ops ? sops[x + 2] : 0u;
                    const uint32_t o3 = x + 3 < nw_ops ? sops[x + 3] : 0u;
                    const int b0 = 2 * (jm & 15), b1 = b0 + 2 * cnt;
                    const uint64_t X = (uint64_t)o0 | ((uint64_t)o1 << 32);
                    const uint64_t Y = (uint64_t)o2 | ((uint64_t)o3 << 32);
                    const uint64_t mx = bitrange64(b0 < 64 ? b0 : 64, b1 < 64 ? b1 : 64);
                    const uint64_t my = bitrange64(b0 > 64 ? b0 - 64 : 0, b1 > 64 ? b1 - 64 : 0);
                    const uint64_t xl = X & e64, xh = (X >> 1) & e64;
                    const uint64_t yl = Y & e64, yh = (Y >> 1) & e64;
                    const int tops = __popcll(xl & ~xh & mx) + __popcll(yl & ~yh & my);
                    const int lefts = __popcll(xh & ~xl & mx) + __popcll(yh & ~yl & my);
                    drow = (uint32_t)(cnt - lefts);
                    dcol = (uint32_t)(cnt - tops);
                }
                uint32_t sr = drow, sc = dcol;
                for (int d = 1; d < TPL; d <<= 1) {
                    const uint32_t yr = __shfl_up(sr, d, TPL), yc = __shfl_up(sc, d, TPL);
                    if (u >= d) {
                        sr += yr;
                        sc += yc;
                    }
                }
                const int row = end_row + (int)(sr - drow), col = end_col + (int)(sc - dcol);
                if (cnt > 0) {
                    // code fields of the run's first cell in the staged read and window words
                    int rf = (int)(read_off + (uint32_t)col) - 16 * rlo;
                    int gf = (int)(win_off + (uint32_t)row) - 16 * glo;
                    const int jlo = 16 * wlo;
                    // slot bytes of position 0 in the two rows (place = position - lead)
                    uint32_t *s32 = s_text[wv][q];
                    const int xo = al + hdr - lead, xg = xo + rn + 1;
                    uint32_t prev_r = 0u, prev_g = 0u;
                    bool pfull = false;   // the group before was written whole but its tail
                    // four positions at a time: their ops, then four characters per row;
                    // a whole group goes into the slot as words (joined with the group
                    // before at the slot's word boundaries), the first and last groups
                    // of a run and the group that crosses lead as bytes
                    for (int i = 0; i < cnt; i += 4) {
                        // ops of positions p0 + i .. p0 + i + 3: fields j .. j + 3 with
                        // j = n - 4 - p0 - i (descending positions); fields below the
                        // staged words belong to positions past the line (taken as 0)
                        const int j = n - 4 - p0 - i;   // j - jlo >= -3
                        const uint32_t og = fields4(sops, j - jlo);
                        // op of position p0 + i + t in byte t
                        const uint32_t ob = __builtin_amdgcn_perm(0u, spread4(og), 0x00010203u);
                        const uint32_t lo = ob & 0x01010101u, hi = (ob >> 1) & 0x01010101u;
                        const uint32_t tops = lo & ~hi, lefts = hi & ~lo;   // TOP 1, LEFT 2
                        const uint32_t rch = row_chars4(fields4(src_, rf), tops);
                        const uint32_t gch = row_chars4(fields4(sgc, gf), lefts);
                        rf += 4 - (int)__builtin_popcount(tops);
                        gf += 4 - (int)__builtin_popcount(lefts);
                        const int P = p0 + i;
                        if (i + 4 <= cnt && P >= lead) {
                            put_head(s32, xo + P, rch, prev_r, pfull);
                            put_head(s32, xg + P, gch, prev_g, pfull);
                            prev_r = rch;
                            prev_g = gch;
                            pfull = true;
                        } else {
                            if (pfull) {
                                put_tail(s32, xo + P - 4, prev_r);
                                put_tail(s32, xg + P - 4, prev_g);
                            }
                            pfull = false;
                            char *s8 = (char *)s32;
#pragma unroll
                            for (int t = 0; t < 4; ++t) {
                                if (i + t < cnt && P + t >= lead) {
                                    s8[xo + P + t] = (char)(rch >> (8 * t));
                                    s8[xg + P + t] = (char)(gch >> (8 * t));
                                }
                            }
                        }
                    }
                    if (pfull) {
                        const int P = p0 + ((cnt - 1) & ~3);   // the last group
                        put_tail(s32, xo + P, prev_r);
                        put_tail(s32, xg + P, prev_g);
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // ---- copy the line out
            if (act) {
                uint32_t *o32 = (uint32_t *)(a.out + (loff - (uint64_t)al));
                const int end = al + (int)llen;
                const int last = end >> 2;   // the word of the line's end
                const uint32_t *t32 = s_text[wv][q];
                // whole words [al ? 1 : 0, last), four per lane at a time (the LDS
                // reads before the stores)
                int i = (al ? 1 : 0) + u;
                for (; i + 3 * TPL < last; i += 4 * TPL) {
                    const uint32_t v0 = t32[i], v1 = t32[i + TPL], v2 = t32[i + 2 * TPL],
                                   v3 = t32[i + 3 * TPL];
                    o32[i] = v0;
                    o32[i + TPL] = v1;
                    o32[i + 2 * TPL] = v2;
                    o32[i + 3 * TPL] = v3;
                }
                for (; i < last; i += TPL) o32[i] = t32[i];
                // the first and last words, shared with the neighbouring lines: bytes
                // [al, 4) of word 0 when al > 0, bytes of word last below end
                const bool first = u == 0 && al;
                const bool tail = u == 1 && (end & 3) && !(al && last == 0);
                if (first || tail) {
                    const int w = first ? 0 : last;
                    const int b0 = first ? al : 0, b1 = first ? (end < 4 ? end : 4) : (end & 3);
                    const uint32_t v = t32[w];
                    char *o8 = (char *)(o32 + w);
                    for (int b = b0; b < b1; ++b) o8[b] = (char)(v >> (8 * b));
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
}

}  // namespace

hipError_t launch_rec_measure(const RecArgs &a, hipStream_t st)
{
    if (a.n < 0) return hipErrorInvalidValue;
    const int blocks = a.n / RB + 1;
    hipLaunchKernelGGL(rec_measure, dim3(blocks < 2048 ? blocks : 2048), dim3(RB), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_rec_write(const RecArgs &a, hipStream_t st)
{
    if (a.n <= 0) return hipSuccess;
    // a wave per tile of 64 candidates for both kernels: the short lines eight at
    // a time, then the others one at a time
    const int tiles = (a.n + 63) / 64;
    const int waves = tiles < 65536 * 4 ? tiles : 65536 * 4;
    if (a.names)
        hipLaunchKernelGGL(rec_write_lines<true>, dim3((waves + RB / 64 - 1) / (RB / 64)), dim3(RB), 0,
                           st, a);
    else
        hipLaunchKernelGGL(rec_write_lines<false>, dim3((waves + RB / 64 - 1) / (RB / 64)), dim3(RB),
                           0, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(rec_write, dim3((waves + RB / 64 - 1) / (RB / 64)), dim3(RB), 0, st, a);
    return hipGetLastError();
}

// exclusive sum of n + 1 line lengths (the last one 0): off[n] is the total
hipError_t rec_scan(void *tmp, size_t &tmp_bytes, const uint64_t *len, uint64_t *off, int n,
                    hipStream_t st)
{
    return hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, len, off, n + 1, st);
}

}  // namespace srm
