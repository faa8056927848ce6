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

// leading gaps, closing gaps, indel scan and printed length of one hit
// (cuda-functions.cu:866-980); returns the line length
__device__ uint64_t measure_one(const RecArgs &a, int t)
{
    const srm_cand c = a.cand[t];
    const srm_result r = a.res[t];
    const uint32_t *ops = a.ops + c.ops_off;
    const int n = r.n_ops, w = c.w, L = c.L;
    // slot position p holds transcript op n-1-p (the reference reverses the rows)
    auto R_gap = [&](int p) { return op_at(ops, n - 1 - p) == SRM_OP_TOP; };
    auto G_gap = [&](int p) { return op_at(ops, n - 1 - p) == SRM_OP_LEFT; };
    int lead = 0;
    while (lead < n && R_gap(lead)) ++lead;                      // :878-886
    int trail = 0;
    for (int p = n - 1; p >= 0 && R_gap(p); --p) ++trail;        // :897-907
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

__global__ void __launch_bounds__(RB) rec_write(RecArgs a)
{
    const int t = blockIdx.x * RB + threadIdx.x;
    if (t >= a.n || a.len[t] == 0) return;
    const srm_cand c = a.cand[t];
    const srm_result r = a.res[t];
    const uint4 inf = a.info[t];
    const int lead = (int)(inf.z & 0xFFFFu), rn = (int)(inf.z >> 16);
    char *o = a.out + a.off[t];
    const int32_t id = a.read_id ? a.read_id[t] : t;
    if (a.names) {
        const char *s = a.names + (int64_t)id * a.name_stride;
        for (int j = 0; j < a.name_stride - 1 && s[j]; ++j) *o++ = s[j];
    } else {
        for (int j = 0; j < a.prefix_len; ++j) *o++ = a.prefix[j];
        o = put_int(o, id);
    }
    *o++ = '\t';
    o = put_int(o, (int32_t)inf.x);
    *o++ = '\t';
    o = put_int(o, (int32_t)inf.y);
    *o++ = '\t';
    *o++ = c.flag > 3 ? '<' : '>';
    *o++ = c.flag > 3 ? '<' : '>';
    *o++ = '\t';
    char *orow = o;              // read row: rn characters
    char *grow = o + rn + 1;     // genome row after the tab
    const uint32_t *ops = a.ops + c.ops_off;
    const bool bytes = (c.bits & SRM_CAND_BYTES) != 0;
    const int n = r.n_ops;
    int row = r.end_row, col = r.end_col;
    const int end = lead + rn;
    for (int p = 0; p < end; ++p) {
        const uint32_t op = op_at(ops, n - 1 - p);
        char rc = '_', gc = '_';
        if (op != SRM_OP_TOP)
            rc = bytes ? a.read_bytes[(int64_t)c.read_off + col]
                       : "ACGT"[rcode(a.read_words, (int64_t)c.read_off + col)];
        if (op != SRM_OP_LEFT)
            gc = bytes ? a.win_bytes[(int64_t)c.win_off + row]
                       : "ACGT"[rcode(a.win_words, (int64_t)c.win_off + row)];
        if (p >= lead) {
            orow[p - lead] = rc;
            grow[p - lead] = gc;
        }
        if (op != SRM_OP_LEFT) ++row;
        if (op != SRM_OP_TOP) ++col;
    }
    o = grow - 1;
    *o = '\t';
    o = grow + rn;
    *o++ = '\t';
    o = put_int(o, (int)r.score);
    *o = '\n';
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
    hipLaunchKernelGGL(rec_write, dim3((a.n + RB - 1) / RB), dim3(RB), 0, st, a);
    return hipGetLastError();
}

// exclusive sum of n + 1 line lengths (the last one 0): off[n] is the total
hipError_t rec_scan(void *tmp, size_t &tmp_bytes, const uint64_t *len, uint64_t *off, int n,
                    hipStream_t st)
{
    return hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, len, off, n + 1, st);
}

}  // namespace srm
