// records_gpu.hip -- the output records formatted on the device.
//
// After a job has run, every hit (score <= k) becomes the TSV line of the
// reference's post-processing (src/cuda-functions.cu:828-994), byte for byte the
// line srm_format_record writes on the host:
//   rec_measure  one thread per candidate derives, from the transcript alone,
//                the leading and closing gaps of the read row, the indel scan,
//                the printed length and the start/stop, and the line length;
//   (scan)       exclusive sum of the line lengths = each line's offset;
//   rec_write    the same thread replays the transcript from the walk's end cell
//                and writes name, coordinates, both rows and the score.
// The text lands in candidate order in one HBM buffer that the host copies out
// in one transfer.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "nw_device.h"

namespace srm {

namespace {

constexpr int RB = 256;

__device__ __forceinline__ uint32_t rcode(const uint32_t *w, int64_t idx)
{
    return (w[idx >> 4] >> ((uint32_t)(idx & 15) * 2u)) & 3u;
}

__device__ __forceinline__ uint32_t op_at(const uint32_t *ops, int j)
{
    return (ops[j >> 4] >> ((j & 15) * 2)) & 3u;
}

// characters %d prints for v
__device__ __forceinline__ int ndig(int v)
{
    uint32_t u = v < 0 ? 0u - (uint32_t)v : (uint32_t)v;
    int n = v < 0 ? 2 : 1;
    while (u >= 10) {
        u /= 10;
        ++n;
    }
    return n;
}

// writes v as %d does; returns the end
__device__ __forceinline__ char *put_int(char *o, int v)
{
    const int nd = ndig(v);
    char *e = o + nd;
    uint32_t u = v < 0 ? 0u - (uint32_t)v : (uint32_t)v;
    char *q = e;
    do {
        *--q = (char)('0' + u % 10);
        u /= 10;
    } while (u);
    if (v < 0) *o = '-';
    return e;
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

// a transcript read op by op, one word load per 16 ops
struct OpCursor {
    const uint32_t *ops;
    int wi = -1;
    uint32_t word = 0;
    __device__ uint32_t at(int j)
    {
        const int w = j >> 4;
        if (w != wi) {
            wi = w;
            word = ops[w];
        }
        return (word >> ((j & 15) * 2)) & 3u;
    }
};

// leading gaps, closing gaps, indel scan and printed length of one hit
// (cuda-functions.cu:866-980); returns the line length
__device__ uint64_t measure_one(const RecArgs &a, int t)
{
    const srm_cand c = a.cand[t];
    const srm_result r = a.res[t];
    const uint32_t *ops = a.ops + c.ops_off;
    const int n = r.n_ops, w = c.w, L = c.L;
    // slot position p holds transcript op n-1-p (the reference reverses the rows);
    // the forward scans and the closing-gap scan each keep their current word
    OpCursor fw{ops}, bw{ops};
    auto R_gap = [&](int p) { return fw.at(n - 1 - p) == SRM_OP_TOP; };
    auto G_gap = [&](int p) { return fw.at(n - 1 - p) == SRM_OP_LEFT; };
    int lead = 0;
    while (lead < n && R_gap(lead)) ++lead;                      // :878-886
    int trail = 0;
    for (int p = n - 1; p >= 0 && bw.at(n - 1 - p) == SRM_OP_TOP; --p) ++trail;   // :897-907
    int del = 0, ins = 0, i = 0, z = 0, eor = 0;                 // :910-959
    while (i < w - trail - lead + z) {
        const int p = lead + i;
        if (p >= n) {          // NUL: the zero padding after the emitted row
            eor = i;
            break;
        }
        if (R_gap(p)) {
            ++del;
            ++z;
        } else if (G_gap(p)) {
            ++ins;
            ++z;
        }
        ++i;
    }
    const short chars = eor > 0 ? (short)(eor - trail) : (short)(L + del);   // :972-980 (Q8)
    const int cap = w + (int)c.k;
    const int lim = chars < 0 ? cap : chars;
    int rn = n - lead < lim ? n - lead : lim;                    // %.*s stops at NUL (Q9)
    if (rn < 0) rn = 0;
    const uint32_t start = (uint32_t)a.pos[t] + (uint32_t)lead;
    const int32_t stop = (int32_t)((uint32_t)(int32_t)chars - (uint32_t)ins + start - 1u);
    a.info[t] = make_uint4(start, (uint32_t)stop, (uint32_t)lead | ((uint32_t)rn << 16), 0u);
    const int32_t id = a.read_id ? a.read_id[t] : t;
    return (uint64_t)(name_len(a, id) + ndig((int32_t)start) + ndig(stop) + 2 + 2 * rn +
                      ndig((int)r.score) + 7);
}

__global__ void __launch_bounds__(RB) rec_measure(RecArgs a)
{
    const int t = blockIdx.x * RB + threadIdx.x;
    const bool hit = t < a.n && a.res[t].score <= (int)a.cand[t].k;
    if (t < a.n) a.len[t] = hit ? measure_one(a, t) : 0;
    if (t == a.n) a.len[t] = 0;                     // the scan's tail element
    const uint64_t b = __ballot(hit);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(a.hits, (uint32_t)__popcll(b));
}

// run sizes of rec_write_group (below), which writes the lines grouped_line() accepts
constexpr int GB = 256;              // threads per block = candidates per tile
constexpr int GL = 16;               // lines per run
constexpr int TPL = GB / GL;         // threads per line
constexpr int CHUNK = 32;            // printed positions per thread
constexpr int MAXP = TPL * CHUNK;    // printed positions per line (512)
constexpr int LCAP = 1024;           // text bytes per line
constexpr int GTEXT = GL * LCAP;     // text bytes per run
constexpr int WPL = MAXP / 16 + 3;   // staged words per line: transcript, read codes, window codes
constexpr int NAMEW = 64;            // staged name words per line (name_stride <= 256)

__device__ __forceinline__ bool grouped_line(uint64_t len, uint32_t lead_rn, uint32_t bits)
{
    const int endp = (int)(lead_rn & 0xFFFFu) + (int)(lead_rn >> 16);
    return len > 0 && len <= (uint64_t)(LCAP - 4) && endp <= MAXP && !(bits & SRM_CAND_BYTES);
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
        // the lines rec_write_group leaves: long ones and raw-byte candidates
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
// rec_write_group: the lines of up to 16 consecutive hits at a time per block.
// A block takes 256 candidates (one per thread: its length, offset and metadata in
// one coalesced load per field) and cuts its hits, in candidate order, into runs
// of up to GL lines whose text is contiguous in the output.  For each run:
//   stage   the transcript words and the read and window code words each line's
//           rows need, and a name from the table, into LDS (TPL threads per line,
//           every load of the run in flight at once)
//   build   header, rows and tail into one LDS text buffer at the output's byte
//           alignment; a line's rows in chunks of CHUNK printed positions per
//           thread, placed by a scan of the chunks' row and column advances
//   copy    the run's text to the output with dword stores (bytes only where the
//           first and last dwords are shared with the neighbouring text)
// Lines longer than LCAP bytes or MAXP printed positions, and raw-byte
// candidates, are left to rec_write, which writes only those.
// 2-bit field idx (0..47) of three consecutive words
__device__ __forceinline__ uint32_t bits2(uint32_t w0, uint32_t w1, uint32_t w2, int idx)
{
    const uint32_t w = idx < 16 ? w0 : (idx < 32 ? w1 : w2);
    return (w >> ((idx & 15) * 2)) & 3u;
}

// per-line metadata of a run, in LDS
enum { M_ROFF, M_WOFF, M_OPSOFF, M_LW, M_FLAG, M_RES0, M_RES1, M_START, M_STOP, M_LEADRN,
       M_OFFLO, M_OFFHI, M_ID, M_LEN, M_WORDS };

__global__ void __launch_bounds__(GB) rec_write_group(RecArgs a)
{
    __shared__ uint32_t s_text[GTEXT / 4 + 2];
    __shared__ uint32_t s_ops[GL][WPL], s_rc[GL][WPL], s_gc[GL][WPL];
    __shared__ uint32_t s_name[GL][NAMEW];
    __shared__ uint32_t s_meta[GL][M_WORDS];
    __shared__ int s_list[GB];             // the tile's hits, candidate order: thread, -1 - thread if not grouped
    __shared__ uint32_t s_hlen[GB];        // their text lengths
    __shared__ int s_slot[GB];             // per thread: run << 8 | line in run, or -1
    __shared__ int s_rstart[GB], s_rcount[GB];
    __shared__ int s_nruns, s_nhits;
    __shared__ int s_wsum[GB / 64];
    __shared__ char s_prefix[256];

    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (!a.names)
        for (int q = tid; q < a.prefix_len; q += GB) s_prefix[q] = a.prefix[q];
    const int ntile = (a.n + GB - 1) / GB;
    for (int tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        // ---- the tile's candidates, one per thread
        const int t = tile * GB + tid;
        const bool in = t < a.n;
        const uint64_t len = in ? a.len[t] : 0;
        const bool hit = len > 0;
        uint32_t cw[5] = {0u, 0u, 0u, 0u, 0u}, rw[2] = {0u, 0u};
        uint4 inf = make_uint4(0u, 0u, 0u, 0u);
        uint64_t off = 0;
        int32_t id = 0;
        if (hit) {
            const uint32_t *cp = (const uint32_t *)(a.cand + t);
            const uint32_t *rp = (const uint32_t *)(a.res + t);
            for (int i = 0; i < 5; ++i) cw[i] = cp[i];
            rw[0] = rp[0];
            rw[1] = rp[1];
            inf = a.info[t];
            off = a.off[t];
            id = a.read_id ? a.read_id[t] : t;
        }
        const bool grp = hit && grouped_line(len, inf.z, cw[4] >> 16);
        // ---- compact the hits in candidate order
        const uint64_t hb = __ballot(hit);
        const int before = __popcll(hb & ((1ull << lane) - 1ull));
        if (lane == 0) s_wsum[wv] = __popcll(hb);
        __syncthreads();
        int base = 0, nh = 0;
        for (int x = 0; x < GB / 64; ++x) {
            if (x < wv) base += s_wsum[x];
            nh += s_wsum[x];
        }
        const int kk = base + before;
        if (hit) {
            s_list[kk] = grp ? tid : -1 - tid;
            s_hlen[kk] = (uint32_t)len;
        }
        s_slot[tid] = -1;
        __syncthreads();
        // ---- runs of consecutive grouped hits (at most GL lines, GTEXT - 4 bytes)
        if (tid == 0) {
            int nr = 0, k = 0;
            while (k < nh) {
                if (s_list[k] < 0) {
                    ++k;
                    continue;
                }
                s_rstart[nr] = k;
                int m = 0;
                uint32_t tot = 0;
                while (k < nh && s_list[k] >= 0 && m < GL && tot + s_hlen[k] <= (uint32_t)(GTEXT - 4)) {
                    tot += s_hlen[k];
                    s_slot[s_list[k]] = (nr << 8) | m;
                    ++m;
                    ++k;
                }
                s_rcount[nr] = m;
                ++nr;
            }
            s_nruns = nr;
            s_nhits = nh;
        }
        __syncthreads();
        const int nruns = s_nruns;
        const int myslot = s_slot[tid];
        for (int g = 0; g < nruns; ++g) {
            // ---- the run's metadata from its owners
            if (grp && (myslot >> 8) == g) {
                uint32_t *m = s_meta[myslot & 0xFF];
                m[M_ROFF] = cw[0];
                m[M_WOFF] = cw[1];
                m[M_OPSOFF] = cw[2];
                m[M_LW] = cw[3];
                m[M_FLAG] = cw[4];
                m[M_RES0] = rw[0];
                m[M_RES1] = rw[1];
                m[M_START] = inf.x;
                m[M_STOP] = inf.y;
                m[M_LEADRN] = inf.z;
                m[M_OFFLO] = (uint32_t)off;
                m[M_OFFHI] = (uint32_t)(off >> 32);
                m[M_ID] = (uint32_t)id;
                m[M_LEN] = (uint32_t)len;
            }
            __syncthreads();
            const int nl = s_rcount[g];
            const int li = tid / TPL, u = tid % TPL;
            const bool act = li < nl;
            const uint32_t *m = s_meta[act ? li : 0];
            const uint32_t read_off = m[M_ROFF], win_off = m[M_WOFF];
            const int L = (int)(m[M_LW] & 0xFFFFu), w = (int)(m[M_LW] >> 16);
            const int flag = (int)(m[M_FLAG] & 0xFFu);
            const int score = (int)(int16_t)(m[M_RES0] & 0xFFFFu), n = (int)(m[M_RES0] >> 16);
            const int end_row = (int)(m[M_RES1] & 0xFFFFu), end_col = (int)(m[M_RES1] >> 16);
            const int lead = (int)(m[M_LEADRN] & 0xFFFFu), rn = (int)(m[M_LEADRN] >> 16);
            const int endp = lead + rn;
            const int32_t lid = (int32_t)m[M_ID];
            const uint64_t loff = (uint64_t)m[M_OFFLO] | ((uint64_t)m[M_OFFHI] << 32);
            const int wlo = endp > 0 ? (n - endp) >> 4 : 0;
            const int rlo = (int)((read_off + (uint32_t)end_col) >> 4);
            const int glo = (int)((win_off + (uint32_t)end_row) >> 4);
            // ---- stage: transcript, codes and name of every line of the run
            if (act) {
                if (endp > 0) {
                    const uint32_t *ops = a.ops + m[M_OPSOFF];
                    const int nw_ops = ((n - 1) >> 4) - wlo + 1;
                    for (int k = u; k < nw_ops; k += TPL) s_ops[li][k] = ops[wlo + k];
                    const int rhi = (int)((read_off + (uint32_t)min(end_col + endp, L)) >> 4);
                    for (int k = u; k <= rhi - rlo; k += TPL) s_rc[li][k] = a.read_words[rlo + k];
                    const int ghi = (int)((win_off + (uint32_t)min(end_row + endp, w)) >> 4);
                    for (int k = u; k <= ghi - glo; k += TPL) s_gc[li][k] = a.win_words[glo + k];
                }
                if (a.names) {
                    const char *nm = a.names + (int64_t)lid * a.name_stride;
                    char *sn = (char *)s_name[li];
                    for (int q = u; q < a.name_stride; q += TPL) sn[q] = nm[q];
                }
            }
            __syncthreads();
            // ---- build the run's text in LDS
            const uint64_t off0 = (uint64_t)s_meta[0][M_OFFLO] | ((uint64_t)s_meta[0][M_OFFHI] << 32);
            const int al0 = (int)(off0 & 3u);
            if (act) {
                char *line = (char *)s_text + al0 + (int)(loff - off0);
                const int32_t start = (int32_t)m[M_START], stop = (int32_t)m[M_STOP];
                const int nds = ndig(start), ndt = ndig(stop), ndi = ndig(lid), ndsc = ndig(score);
                int name_n;
                if (a.names) {
                    const char *sn = (const char *)s_name[li];
                    name_n = 0;
                    while (name_n < a.name_stride - 1 && sn[name_n]) ++name_n;
                } else {
                    name_n = a.prefix_len + ndi;
                }
                const int hdr = name_n + 1 + nds + 1 + ndt + 1 + 2 + 1;
                for (int h = u; h < hdr; h += TPL) {
                    char ch;
                    int q = h;
                    if (q < name_n) {
                        if (a.names)
                            ch = ((const char *)s_name[li])[q];
                        else
                            ch = q < a.prefix_len ? s_prefix[q] : digit_at(lid, ndi, q - a.prefix_len);
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
                        ch = flag > 3 ? '<' : '>';
                    }
                    line[h] = ch;
                }
                char *orow = line + hdr;
                char *grow = orow + rn + 1;
                // rows: positions p0 .. p0 + cnt - 1 of this thread; position p holds op n-1-p
                const int p0 = u * CHUNK;
                const int cnt = endp - p0 < 0 ? 0 : (endp - p0 > CHUNK ? CHUNK : endp - p0);
                uint32_t o0 = 0, o1 = 0, o2 = 0;
                int jb = 0;
                if (cnt > 0) {
                    const int j0 = n - 1 - p0, jm = j0 - (cnt - 1);
                    jb = jm >> 4;
                    const int nw_ops = ((n - 1) >> 4) - wlo + 1;
                    const int x = jb - wlo;
                    o0 = s_ops[li][x];
                    o1 = x + 1 < nw_ops ? s_ops[li][x + 1] : 0u;
                    o2 = x + 2 < nw_ops ? s_ops[li][x + 2] : 0u;
                }
                uint32_t drow = 0, dcol = 0;
                for (int i = 0; i < cnt; ++i) {
                    const uint32_t op = bits2(o0, o1, o2, n - 1 - (p0 + i) - 16 * jb);
                    drow += op != SRM_OP_LEFT;
                    dcol += op != SRM_OP_TOP;
                }
                // exclusive scan of the advances over the line's TPL threads
                uint32_t sr = drow, sc = dcol;
                for (int d = 1; d < TPL; d <<= 1) {
                    const uint32_t yr = __shfl_up(sr, d, TPL), yc = __shfl_up(sc, d, TPL);
                    if (u >= d) {
                        sr += yr;
                        sc += yc;
                    }
                }
                int row = end_row + (int)(sr - drow), col = end_col + (int)(sc - dcol);
                if (cnt > 0) {
                    const int rnw = (int)((read_off + (uint32_t)min(end_col + endp, L)) >> 4) - rlo + 1;
                    const int gnw = (int)((win_off + (uint32_t)min(end_row + endp, w)) >> 4) - glo + 1;
                    const int ri = (int)(read_off + (uint32_t)col) - 16 * rlo;
                    const int gi = (int)(win_off + (uint32_t)row) - 16 * glo;
                    const int rx = ri >> 4, gx = gi >> 4;
                    const uint32_t r0 = rx < rnw ? s_rc[li][rx] : 0u;
                    const uint32_t r1 = rx + 1 < rnw ? s_rc[li][rx + 1] : 0u;
                    const uint32_t r2 = rx + 2 < rnw ? s_rc[li][rx + 2] : 0u;
                    const uint32_t g0 = gx < gnw ? s_gc[li][gx] : 0u;
                    const uint32_t g1 = gx + 1 < gnw ? s_gc[li][gx + 1] : 0u;
                    const uint32_t g2 = gx + 2 < gnw ? s_gc[li][gx + 2] : 0u;
                    int rk = ri & 15, gk = gi & 15;
                    for (int i = 0; i < cnt; ++i) {
                        const int p = p0 + i;
                        const uint32_t op = bits2(o0, o1, o2, n - 1 - p - 16 * jb);
                        if (p >= lead) {
                            orow[p - lead] = op != SRM_OP_TOP ? base_char(bits2(r0, r1, r2, rk)) : '_';
                            grow[p - lead] = op != SRM_OP_LEFT ? base_char(bits2(g0, g1, g2, gk)) : '_';
                        }
                        if (op != SRM_OP_LEFT) ++gk;
                        if (op != SRM_OP_TOP) ++rk;
                    }
                }
                if (u == 0) orow[rn] = '\t';
                char *tail = grow + rn;
                if (u <= ndsc + 1)
                    tail[u] = u == 0 ? '\t' : (u == ndsc + 1 ? '\n' : digit_at(score, ndsc, u - 1));
            }
            __syncthreads();
            // ---- copy the run's text out
            {
                const uint32_t *ml = s_meta[nl - 1];
                const uint64_t end_off = ((uint64_t)ml[M_OFFLO] | ((uint64_t)ml[M_OFFHI] << 32)) + ml[M_LEN];
                const int end = al0 + (int)(end_off - off0);
                const int nd = (end + 3) >> 2;
                uint32_t *o32 = (uint32_t *)(a.out + (off0 - (uint64_t)al0));
                for (int i = tid; i < nd; i += GB) {
                    const uint32_t v = s_text[i];
                    const int b0 = 4 * i;
                    if (b0 >= al0 && b0 + 4 <= end) {
                        o32[i] = v;
                    } else {
                        char *o8 = (char *)(o32 + i);
                        for (int b = 0; b < 4; ++b)
                            if (b0 + b >= al0 && b0 + b < end) o8[b] = (char)(v >> (8 * b));
                    }
                }
            }
            __syncthreads();
        }
    }
}

}  // namespace

hipError_t launch_rec_measure(const RecArgs &a, hipStream_t st)
{
    if (a.n < 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(rec_measure, dim3(a.n / RB + 1), dim3(RB), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_rec_write(const RecArgs &a, hipStream_t st)
{
    if (a.n <= 0) return hipSuccess;
    // runs of short lines: a block per tile of 256 candidates, a grid that strides
    const int gtiles = (a.n + GB - 1) / GB;
    hipLaunchKernelGGL(rec_write_group, dim3(gtiles < 4096 ? gtiles : 4096), dim3(GB), 0, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // the other lines: a wave per tile of 64 candidates
    const int tiles = (a.n + 63) / 64;
    const int waves = tiles < 65536 * 4 ? tiles : 65536 * 4;
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
