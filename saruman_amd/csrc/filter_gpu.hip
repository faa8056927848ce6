// filter_gpu.hip -- the q-gram filter ("SEASHORE") on the GPU.
//
// Same candidates, flags, perfect hits and batch boundaries as the host filter
// (filter.cpp); only the per-read work moves to the device:
//   1. count   one thread per read looks up each of its q-grams in the device
//              copy of the q-gram index and writes the list length + 1 and
//              the index of the list's first key;
//   2. scan    exclusive sum of the lengths = the offset of every 0-terminated
//              list in one flat array, reads in order (the reference's
//              data_array, hash.c:396-461, for all chunks at once);
//   3. fill    the same threads write the lists: a copy of the key run the
//              count found, or the hash-table lookup again for a q-gram
//              holding a byte other than ACGT;
//   4. pair    one thread per read runs run_seashore (hash.c:502-834) and
//              add_start_position (hash.c:40-178) over the flat array.  A list
//              index beyond the list runs on into the following lists and the
//              following reads' lists up to the end of the read's chunk of
//              chunk_value reads (Q12), exactly as a walk of data_array does.
//              The read's result-hash entries live in LDS (16 entries, 4 perfect
//              hits per read); a read that needs more is marked and run again
//              with slots in HBM (1024, then 65534 entries);
//   5. compact every wave reserves space for its reads' entries with one atomic
//              and writes them contiguously; the host downloads 12 bytes per
//              read plus the entries.
// The sequential batch pass (Q10 leak, flush boundaries) then runs on the host
// over the downloaded entries (filter_internal.h).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "filter_internal.h"
#include "filter_pair.h"
#include "nw_device.h"

using namespace srm_filter_impl;
using namespace srm_pair;

namespace {

constexpr int FB = 256;      // threads per block
constexpr int ECAP = 16;     // result-hash entries per read kept in LDS
constexpr int PCAP = 4;      // perfect hits per read kept in LDS

__global__ void __launch_bounds__(FB) fl_count(FArgs a)
{
    const uint32_t r = blockIdx.x * FB + threadIdx.x;
    if (r >= a.n_reads) return;
    const DParam P = a.par[a.pid[r]];
    const DQ X = a.qix[P.q];
    const char *rd = a.reads + a.roff[r];
    const uint32_t lb = a.list_base[r];
    for (int x = 0; x < P.n; ++x) {
        uint32_t lo;
        a.list_cnt[lb + x] =
            lookup<false>(X, a.genome, rd, P.L, a.dr, a.gram[P.gram_off + x], nullptr, &lo) + 1;
        a.list_lo[lb + x] = lo;
    }
}

__global__ void __launch_bounds__(FB) fl_fill(FArgs a)
{
    const uint32_t r = blockIdx.x * FB + threadIdx.x;
    if (r >= a.n_reads) return;
    const DParam P = a.par[a.pid[r]];
    const DQ X = a.qix[P.q];
    const char *rd = a.reads + a.roff[r];
    const uint32_t lb = a.list_base[r];
    const uint64_t mask = ((uint64_t)1 << X.pos_bits) - 1;
    for (int x = 0; x < P.n; ++x) {
        uint32_t *dst = a.flat + a.list_start[lb + x];
        const uint32_t lo = a.list_lo[lb + x];
        uint32_t c;
        if (lo != NO_RANGE) {   // the positions are the keys fl_count found
            c = a.list_cnt[lb + x] - 1;
            for (uint32_t i = 0; i < c; ++i) dst[i] = (uint32_t)(X.keys[lo + i] & mask);
        } else {
            c = lookup<true>(X, a.genome, rd, P.L, a.dr, a.gram[P.gram_off + x], dst);
        }
        dst[c] = 0;
    }
}

__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

template <bool GLOBAL>
__global__ void __launch_bounds__(FB) fl_pair(FArgs a)
{
    __shared__ int32_t s_pos[GLOBAL ? 1 : ECAP * FB];
    __shared__ int8_t s_flag[GLOBAL ? 1 : ECAP * FB];
    __shared__ int32_t s_perf[GLOBAL ? 1 : PCAP * FB];
    const uint32_t j = blockIdx.x * FB + threadIdx.x;
    const bool active = j < a.n_todo;
    Store s;
    if (GLOBAL) {
        s.pos = a.slot_pos + (size_t)j * a.cap;
        s.flag = a.slot_flag + (size_t)j * a.cap;
        s.perf = a.slot_perf + (size_t)j * a.pcap;
        s.stride = 1;
        s.cap = a.cap;
        s.pcap = a.pcap;
    } else {
        s.pos = s_pos + threadIdx.x;
        s.flag = s_flag + threadIdx.x;
        s.perf = s_perf + threadIdx.x;
        s.stride = FB;
        s.cap = ECAP;
        s.pcap = PCAP;
    }
    if (active) seashore(a, a.todo ? a.todo[j] : j, s);
    const uint32_t ne = (active && !s.ovf) ? s.n : 0, np = (active && !s.ovf) ? s.np : 0;
    // one atomic pair per block reserves the block's output (the two counters are
    // words that every wave would otherwise update); the waves take consecutive
    // parts of it
    __shared__ uint32_t s_sum[FB / 64][2], s_base[2];
    const uint32_t ie = wave_incl_sum(ne), ip = wave_incl_sum(np);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 63) {
        s_sum[wv][0] = ie;
        s_sum[wv][1] = ip;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t te = 0, tp = 0;
        for (int v = 0; v < FB / 64; ++v) {
            te += s_sum[v][0];
            tp += s_sum[v][1];
        }
        s_base[0] = te ? atomicAdd(&a.counters[0], te) : 0u;
        s_base[1] = tp ? atomicAdd(&a.counters[1], tp) : 0u;
    }
    __syncthreads();
    uint32_t be = s_base[0] + ie - ne, bp = s_base[1] + ip - np;
    for (int v = 0; v < wv; ++v) {
        be += s_sum[v][0];
        bp += s_sum[v][1];
    }
    if (!active) return;
    for (uint32_t e = 0; e < ne; ++e) {
        a.out_pos[be + e] = s.P(e);
        a.out_flag[be + e] = s.F(e);
    }
    for (uint32_t e = 0; e < np; ++e) a.out_perf[bp + e] = s.Q(e);
    a.meta[3 * (size_t)j + 0] = be;
    a.meta[3 * (size_t)j + 1] = bp;
    a.meta[3 * (size_t)j + 2] = s.ovf ? OVERFLOW_MARK : (ne | (np << 16));
}

struct GBuf {
    void *p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes)
    {
        if (bytes <= cap) return SRM_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        size_t want = (std::max(bytes, (size_t)256) + 4095) & ~(size_t)4095;
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            p = nullptr;
            return srm::set_error(SRM_ENOMEM, "hipMalloc(%zu) failed: %s", want,
                                  hipGetErrorString(e));
        }
        cap = want;
        return SRM_OK;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    template <class T> T *as() const { return (T *)p; }
};

#define FL_TRY(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return srm::set_error(SRM_EDEVICE, "%s failed: %s (%s:%d)", #expr,              \
                                  hipGetErrorString(e_), __FILE__, __LINE__);               \
    } while (0)

#define FL_RC(expr)                      \
    do {                                 \
        int rc_ = (expr);                \
        if (rc_ != SRM_OK) return rc_;   \
    } while (0)

}  // namespace

struct srm_dfilter {
    srm_ctx *ctx = nullptr;
    int device = 0;
    hipStream_t st = nullptr;
    const srm_index *ix = nullptr;
    size_t glen = 0;
    GBuf genome, dq;
    std::vector<GBuf> keys, pre, odd_hash, odd_pos;
    GBuf par, gram, reads, roff, rlen, pid, list_base, list_cnt, list_lo, list_start, flat, cub_tmp;
    GBuf slot_pos, slot_flag, slot_perf, out_pos, out_flag, out_perf, meta, counters, todo;
    hipEvent_t ev[2] = {nullptr, nullptr};
    float last_ms = 0.f;
    size_t last_lists = 0, last_flat = 0, last_overflow = 0;

    ~srm_dfilter()
    {
        for (GBuf *b : {&genome, &dq, &par, &gram, &reads, &roff, &rlen, &pid, &list_base,
                        &list_cnt, &list_lo, &list_start, &flat, &cub_tmp, &slot_pos, &slot_flag,
                        &slot_perf, &out_pos, &out_flag, &out_perf, &meta, &counters, &todo})
            b->release();
        for (auto *v : {&keys, &pre, &odd_hash, &odd_pos})
            for (auto &b : *v) b.release();
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
    }

    int upload(GBuf &b, const void *src, size_t bytes)
    {
        FL_RC(b.ensure(bytes));
        if (bytes) FL_TRY(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, st));
        return SRM_OK;
    }

    // runs the pairing kernel over n reads (todo == null: all reads of the call)
    // and downloads meta and the compact output
    int pair(FArgs a, bool global, const uint32_t *todo_host, uint32_t n, uint32_t cap,
             uint32_t pcap, std::vector<uint32_t> &meta_h, std::vector<int32_t> &pos_h,
             std::vector<int8_t> &flag_h, std::vector<int32_t> &perf_h)
    {
        a.n_todo = n;
        a.todo = nullptr;
        if (todo_host) {
            FL_RC(upload(todo, todo_host, (size_t)n * 4));
            a.todo = todo.as<uint32_t>();
        }
        if (global) {
            FL_RC(slot_pos.ensure((size_t)n * cap * 4));
            FL_RC(slot_flag.ensure((size_t)n * cap));
            FL_RC(slot_perf.ensure((size_t)n * pcap * 4));
            a.slot_pos = slot_pos.as<int32_t>();
            a.slot_flag = slot_flag.as<int8_t>();
            a.slot_perf = slot_perf.as<int32_t>();
        }
        a.cap = cap;
        a.pcap = pcap;
        FL_RC(out_pos.ensure((size_t)n * cap * 4));
        FL_RC(out_flag.ensure((size_t)n * cap));
        FL_RC(out_perf.ensure((size_t)n * pcap * 4));
        FL_RC(meta.ensure((size_t)n * 12));
        FL_RC(counters.ensure(8));
        a.out_pos = out_pos.as<int32_t>();
        a.out_flag = out_flag.as<int8_t>();
        a.out_perf = out_perf.as<int32_t>();
        a.meta = meta.as<uint32_t>();
        a.counters = counters.as<uint32_t>();
        FL_TRY(hipMemsetAsync(a.counters, 0, 8, st));
        const unsigned grid = (n + FB - 1) / FB;
        if (global) hipLaunchKernelGGL(fl_pair<true>, dim3(grid), dim3(FB), 0, st, a);
        else hipLaunchKernelGGL(fl_pair<false>, dim3(grid), dim3(FB), 0, st, a);
        FL_TRY(hipGetLastError());
        uint32_t tot[2];
        FL_TRY(hipMemcpyAsync(tot, a.counters, 8, hipMemcpyDeviceToHost, st));
        meta_h.resize((size_t)n * 3);
        FL_TRY(hipMemcpyAsync(meta_h.data(), a.meta, (size_t)n * 12, hipMemcpyDeviceToHost, st));
        FL_TRY(hipStreamSynchronize(st));
        pos_h.resize(tot[0]);
        flag_h.resize(tot[0]);
        perf_h.resize(tot[1]);
        if (tot[0]) {
            FL_TRY(hipMemcpyAsync(pos_h.data(), a.out_pos, (size_t)tot[0] * 4,
                                  hipMemcpyDeviceToHost, st));
            FL_TRY(hipMemcpyAsync(flag_h.data(), a.out_flag, tot[0], hipMemcpyDeviceToHost, st));
        }
        if (tot[1])
            FL_TRY(hipMemcpyAsync(perf_h.data(), a.out_perf, (size_t)tot[1] * 4,
                                  hipMemcpyDeviceToHost, st));
        FL_TRY(hipStreamSynchronize(st));
        return SRM_OK;
    }
};

namespace {

// one or both strands (strands: bit 0 forward, bit 1 reverse) over one upload of
// the reads
int run_device(srm_dfilter *G, const uint8_t *blocked, const char *reads,
               const uint64_t *read_off, const uint16_t *read_len, const uint8_t *read_k,
               size_t n_reads, int strands, int chunk_value, srm_filter **out_fwd,
               srm_filter **out_rev)
{
    ReadPlan plan;
    std::string err;
    if (int rc = plan_reads(G->ix, read_off, read_len, read_k, n_reads, plan, err))
        return srm::set_error(rc, "%s", err.c_str());
    std::vector<DParam> dpar;
    std::vector<uint16_t> gram;
    device_params(plan.params, dpar, gram);
    const auto &pid = plan.pid;
    const auto &list_base = plan.list_base;
    const uint64_t nl = list_base[n_reads];
    const uint64_t rbytes = plan.read_bytes;
    G->last_lists = 0;
    G->last_flat = 0;
    G->last_overflow = 0;
    FL_TRY(hipSetDevice(G->device));
    FL_TRY(hipEventRecord(G->ev[0], G->st));
    FL_RC(G->upload(G->par, dpar.data(), dpar.size() * sizeof(DParam)));
    FL_RC(G->upload(G->gram, gram.data(), gram.size() * 2));
    FL_RC(G->upload(G->reads, reads, rbytes));
    FL_RC(G->upload(G->roff, read_off, n_reads * 8));
    FL_RC(G->upload(G->rlen, read_len, n_reads * 2));
    FL_RC(G->upload(G->pid, pid.data(), n_reads * 2));
    FL_RC(G->upload(G->list_base, list_base.data(), (n_reads + 1) * 4));
    FL_RC(G->list_cnt.ensure((nl + 1) * 4));
    FL_RC(G->list_lo.ensure((nl + 1) * 4));
    FL_RC(G->list_start.ensure((nl + 1) * 4));
    FArgs a{};
    a.qix = G->dq.as<DQ>();
    a.par = G->par.as<DParam>();
    a.gram = G->gram.as<uint16_t>();
    a.genome = G->genome.as<char>();
    a.glen = G->glen;
    a.reads = G->reads.as<char>();
    a.roff = G->roff.as<uint64_t>();
    a.rlen = G->rlen.as<uint16_t>();
    a.pid = G->pid.as<uint16_t>();
    a.n_reads = (uint32_t)n_reads;
    a.cv = (uint32_t)chunk_value;
    a.list_base = G->list_base.as<uint32_t>();
    a.list_cnt = G->list_cnt.as<uint32_t>();
    a.list_lo = G->list_lo.as<uint32_t>();
    a.list_start = G->list_start.as<uint32_t>();
    // per strand: the device's entries and perfect hits for the batch pass, which
    // runs on the host once both strands' device work is done (the two strands'
    // passes side by side)
    std::vector<uint32_t> meta_s[2];
    std::vector<int32_t> pos_s[2], perf_s[2];
    std::vector<int8_t> flag_s[2];
    for (int dr : {0, 4}) {
        if (!(strands & (dr ? 2 : 1))) continue;
        a.dr = dr;
        G->last_lists += nl;
        const unsigned grid = (unsigned)((n_reads + FB - 1) / FB);
        FL_TRY(hipMemsetAsync(a.list_cnt + nl, 0, 4, G->st));
        hipLaunchKernelGGL(fl_count, dim3(grid), dim3(FB), 0, G->st, a);
        FL_TRY(hipGetLastError());
        size_t tmp = 0;
        FL_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, a.list_cnt, G->list_start.as<uint32_t>(),
                                                (int)(nl + 1), G->st));
        FL_RC(G->cub_tmp.ensure(tmp));
        FL_TRY(hipcub::DeviceScan::ExclusiveSum(G->cub_tmp.p, tmp, a.list_cnt,
                                                G->list_start.as<uint32_t>(), (int)(nl + 1), G->st));
        uint32_t total = 0;
        FL_TRY(hipMemcpyAsync(&total, G->list_start.as<uint32_t>() + nl, 4, hipMemcpyDeviceToHost,
                              G->st));
        FL_TRY(hipStreamSynchronize(G->st));
        G->last_flat += total;
        FL_RC(G->flat.ensure((size_t)total * 4 + 4));
        a.flat = G->flat.as<uint32_t>();
        hipLaunchKernelGGL(fl_fill, dim3(grid), dim3(FB), 0, G->st, a);
        FL_TRY(hipGetLastError());

        std::vector<uint32_t> &meta = meta_s[dr ? 1 : 0];
        std::vector<int32_t> &pos = pos_s[dr ? 1 : 0], &perf = perf_s[dr ? 1 : 0];
        std::vector<int8_t> &flag = flag_s[dr ? 1 : 0];
        FL_RC(G->pair(a, false, nullptr, (uint32_t)n_reads, ECAP, PCAP, meta, pos, flag, perf));
        // reads that outgrew the LDS slots run again with slots in HBM
        std::vector<uint32_t> over;
        for (size_t r = 0; r < n_reads; ++r)
            if (meta[3 * r + 2] == OVERFLOW_MARK) over.push_back((uint32_t)r);
        G->last_overflow += over.size();
        // (65534: a count word of 0xFFFF | 0xFFFF << 16 would read as the overflow mark)
        for (uint32_t cap : {1024u, 65534u}) {
            if (over.empty()) break;
            const size_t per = std::max<size_t>(1, ((size_t)256 << 20) / ((size_t)cap * 9));
            std::vector<uint32_t> still;
            for (size_t b = 0; b < over.size(); b += per) {
                const uint32_t m = (uint32_t)std::min(per, over.size() - b);
                std::vector<uint32_t> m2;
                std::vector<int32_t> p2, q2;
                std::vector<int8_t> f2;
                FL_RC(G->pair(a, true, over.data() + b, m, cap, cap, m2, p2, f2, q2));
                const uint32_t es = (uint32_t)pos.size(), ps = (uint32_t)perf.size();
                pos.insert(pos.end(), p2.begin(), p2.end());
                flag.insert(flag.end(), f2.begin(), f2.end());
                perf.insert(perf.end(), q2.begin(), q2.end());
                for (uint32_t t = 0; t < m; ++t) {
                    const uint32_t r = over[b + t];
                    if (m2[3 * t + 2] == OVERFLOW_MARK) {
                        still.push_back(r);
                        continue;
                    }
                    meta[3 * r + 0] = es + m2[3 * t + 0];
                    meta[3 * r + 1] = ps + m2[3 * t + 1];
                    meta[3 * r + 2] = m2[3 * t + 2];
                }
            }
            over.swap(still);
        }
        if (!over.empty())
            return srm::set_error(SRM_ENOMEM, "read %u holds more than 65534 filter entries",
                                  over[0]);
        FL_TRY(hipEventRecord(G->ev[1], G->st));
        FL_TRY(hipEventSynchronize(G->ev[1]));
        FL_TRY(hipEventElapsedTime(&G->last_ms, G->ev[0], G->ev[1]));
    }
    auto pass = [&](int x) {
        *(x ? out_rev : out_fwd) =
            batch_pass(MetaAcc{meta_s[x], pos_s[x], flag_s[x], perf_s[x]}, n_reads, read_len,
                       read_k, G->glen, blocked, chunk_value);
    };
    if ((strands & 3) == 3) {
        std::thread other;
        try {
            other = std::thread(pass, 1);
        } catch (...) {   // no thread to be had: both passes on this one
            pass(1);
        }
        pass(0);
        if (other.joinable()) other.join();
    } else {
        pass((strands & 1) ? 0 : 1);
    }
    return SRM_OK;
}

}  // namespace

extern "C" {

int srm_dfilter_create(srm_ctx *ctx, const srm_index *ix, const char *genome, size_t glen,
                       srm_dfilter **out)
{
    *out = nullptr;
    if (!ctx || !ix || glen != ix->glen)
        return srm::set_error(SRM_EINVAL, "genome length does not match the index");
    if (ix->base != 0)
        return srm::set_error(SRM_EINVAL, "a -G chunk's index runs on the host filter");
    auto *G = new srm_dfilter();
    G->ctx = ctx;
    G->device = srm::ctx_device(ctx);
    G->st = srm::ctx_stream(ctx);
    G->ix = ix;
    G->glen = glen;
    auto fail = [&](int rc) {
        delete G;
        return rc;
    };
    if (hipSetDevice(G->device) != hipSuccess ||
        hipEventCreate(&G->ev[0]) != hipSuccess || hipEventCreate(&G->ev[1]) != hipSuccess)
        return fail(srm::set_error(SRM_EDEVICE, "event creation failed"));
    int rc = G->upload(G->genome, genome, glen);
    if (rc) return fail(rc);
    std::vector<DQ> dq(32);
    std::memset(dq.data(), 0, dq.size() * sizeof(DQ));
    G->keys.resize(ix->idx.size());
    G->pre.resize(ix->idx.size());
    G->odd_hash.resize(ix->idx.size());
    G->odd_pos.resize(ix->idx.size());
    for (size_t i = 0; i < ix->idx.size(); ++i) {
        const QIndex &X = ix->idx[i];
        std::vector<uint64_t> oh;
        std::vector<uint32_t> op;
        odd_table(X, oh, op);
        if ((rc = G->upload(G->keys[i], X.keys.data(), X.keys.size() * 8)) ||
            (rc = G->upload(G->pre[i], X.pre.data(), X.pre.size() * 4)) ||
            (rc = G->upload(G->odd_hash[i], oh.data(), oh.size() * 8)) ||
            (rc = G->upload(G->odd_pos[i], op.data(), op.size() * 4)))
            return fail(rc);
        DQ &d = dq[(size_t)X.q];
        d.keys = G->keys[i].as<uint64_t>();
        d.pre = G->pre[i].as<uint32_t>();
        d.odd_hash = G->odd_hash[i].as<uint64_t>();
        d.odd_pos = G->odd_pos[i].as<uint32_t>();
        d.n_odd = (uint32_t)oh.size();
        d.q = X.q;
        d.pos_bits = X.pos_bits;
        d.pre_bases = X.pre_bases;
        // the upload above is asynchronous; the host vectors must outlive it
        if (hipStreamSynchronize(G->st) != hipSuccess)
            return fail(srm::set_error(SRM_EDEVICE, "index upload failed"));
    }
    if ((rc = G->upload(G->dq, dq.data(), dq.size() * sizeof(DQ)))) return fail(rc);
    if (hipStreamSynchronize(G->st) != hipSuccess)
        return fail(srm::set_error(SRM_EDEVICE, "index upload failed"));
    *out = G;
    return SRM_OK;
}

int srm_dfilter_run(srm_dfilter *G, const uint8_t *blocked, const char *reads,
                    const uint64_t *read_off, const uint16_t *read_len, const uint8_t *read_k,
                    size_t n_reads, int do_reverse, int chunk_value, srm_filter **out)
{
    *out = nullptr;
    if (!G || chunk_value < 1 || (do_reverse != 0 && do_reverse != 4))
        return srm::set_error(SRM_EINVAL, "bad filter arguments");
    if (n_reads == 0) {
        *out = new srm_filter();
        return SRM_OK;
    }
    srm_filter *F = nullptr;
    const int rc = run_device(G, blocked, reads, read_off, read_len, read_k, n_reads,
                              do_reverse ? 2 : 1, chunk_value, &F, &F);
    if (rc != SRM_OK) {
        delete F;
        return rc;
    }
    *out = F;
    return SRM_OK;
}

int srm_dfilter_run2(srm_dfilter *G, const uint8_t *blocked, const char *reads,
                     const uint64_t *read_off, const uint16_t *read_len, const uint8_t *read_k,
                     size_t n_reads, int chunk_value, srm_filter **out_fwd, srm_filter **out_rev)
{
    *out_fwd = nullptr;
    *out_rev = nullptr;
    if (!G || chunk_value < 1) return srm::set_error(SRM_EINVAL, "bad filter arguments");
    if (n_reads == 0) {
        *out_fwd = new srm_filter();
        *out_rev = new srm_filter();
        return SRM_OK;
    }
    srm_filter *Ff = nullptr, *Fr = nullptr;
    const int rc = run_device(G, blocked, reads, read_off, read_len, read_k, n_reads, 3,
                              chunk_value, &Ff, &Fr);
    if (rc != SRM_OK) {
        delete Ff;
        delete Fr;
        return rc;
    }
    *out_fwd = Ff;
    *out_rev = Fr;
    return SRM_OK;
}

int srm_dfilter_stats(const srm_dfilter *G, float *device_ms, uint64_t *n_lists,
                      uint64_t *n_list_entries, uint64_t *n_overflow)
{
    if (device_ms) *device_ms = G->last_ms;
    if (n_lists) *n_lists = G->last_lists;
    if (n_list_entries) *n_list_entries = G->last_flat;
    if (n_overflow) *n_overflow = G->last_overflow;
    return SRM_OK;
}

void srm_dfilter_destroy(srm_dfilter *G)
{
    if (!G) return;
    (void)hipSetDevice(G->device);
    (void)hipStreamSynchronize(G->st);
    delete G;
}

}  // extern "C"
