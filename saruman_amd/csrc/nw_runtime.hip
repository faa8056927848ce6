// nw_runtime.hip -- host runtime of libsaruman_hip.so: contexts (one HIP stream and
// device pool each), jobs (one uploaded candidate batch), the launch sequence of
// the banded and plain kernels, hipGraph replay, and the reference's four C entry
// points (src/cuda-functions.cu:541-1120) on top of them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <thread>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "nw_device.h"
#include "saruman_hip.h"

// ---------------------------------------------------------------------------
// globals of the drop-in contract (see saruman_hip.h)
// ---------------------------------------------------------------------------
extern "C" {
struct read_name_hash *read_names = nullptr;
unsigned short int read_length_value = 0;
unsigned short int be_verbose_value = 0;
unsigned short int use_gpu_value = 5000;   // util.c:192-194: 5000 = pick automatically
unsigned long int cuda_memory_per_thread = 0;
char *forward_output_control = nullptr;
char *reverse_output_control = nullptr;
unsigned int forward_alignments_done = 0;
unsigned int reverse_alignments_done = 0;
unsigned int forward_alignment_candidates = 0;
unsigned int reverse_alignment_candidates = 0;
char *main_alloc_ptr_for = nullptr;
char *main_alloc_ptr_rev = nullptr;
float alignment_time[2] = {0.f, 0.f};
}

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(SRM_EDEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                \
    } while (0)

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

// allocations that are made to fail (srm_debug_fail_allocs), process-wide
std::atomic<int> g_fail_allocs{0};

}  // namespace

struct srm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t pool_cap = 0;      // 0 = unlimited
    size_t pool_used = 0;     // bytes held by this context's buffers
    int num_cus = 0;
    DevBuf genome;            // resident packed genome (owned)
    const uint32_t *genome_borrowed = nullptr;   // or another context's (same device)
    size_t genome_words = 0;
    const uint32_t *genome_ptr() const
    {
        return genome_borrowed ? genome_borrowed : (const uint32_t *)genome.p;
    }
    std::mutex mu;
    srm_job *dropin_job = nullptr;
    srm_tuning tuning{};      // zero = defaults (occupancy-sized grids, no merging)

    int ensure(DevBuf &b, size_t bytes)
    {
        if (bytes <= b.cap) return SRM_OK;
        for (int f = g_fail_allocs.load(); f > 0;)
            if (g_fail_allocs.compare_exchange_weak(f, f - 1))
                return fail(SRM_ENOMEM, "injected allocation failure (%zu bytes)", bytes);
        // under a pool cap a buffer grows to exactly its 4 KiB-rounded need (what
        // srm_job_bytes_bound counts); otherwise by half again, for fewer regrowths
        size_t want = pool_cap ? bytes : std::max(bytes, b.cap + b.cap / 2);
        want = (want + 4095) & ~(size_t)4095;
        if (pool_cap && pool_used - b.cap + want > pool_cap) {
            want = bytes;
            if (pool_used - b.cap + want > pool_cap)
                return fail(SRM_ENOMEM, "device pool cap %zu bytes exceeded (need %zu more)",
                            pool_cap, want);
        }
        if (b.p) {
            (void)hipFree(b.p);
            pool_used -= b.cap;
            b.p = nullptr;
            b.cap = 0;
        }
        hipError_t e = hipMalloc(&b.p, want);
        if (e != hipSuccess) {
            b.p = nullptr;
            return fail(SRM_ENOMEM, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
        }
        b.cap = want;
        pool_used += want;
        return SRM_OK;
    }
    void release(DevBuf &b)
    {
        if (b.p) {
            (void)hipFree(b.p);
            pool_used -= b.cap;
        }
        b.p = nullptr;
        b.cap = 0;
    }
};

namespace {
struct BandClass {
    int nw = 0;
    uint32_t tile_begin = 0, ntiles = 0;
    int maxL = 0;
    int blocks = 0;
};
struct PlainChunk {
    uint32_t begin = 0, count = 0;   // range in the plain index list
    size_t mat_elems = 0;
};
// one host thread's share of an upload's validation and classification
struct UploadPart {
    std::vector<uint32_t> band, rest, bytes;
    size_t ops_end = 0, rows = 0;
    uint32_t maxL = 0;   // over the banded candidates
    int bad = -1, why = 0;
    void reset()
    {
        band.clear();
        rest.clear();
        bytes.clear();
        ops_end = rows = 0;
        maxL = 0;
        bad = -1;
        why = 0;
    }
};
// the host arrays of an upload's plan, kept by the job from one upload to the
// next: a batch of about the same size then reuses their pages (a fresh
// multi-megabyte allocation is faulted in page by page, which cost more than the
// planning itself)
struct HostPlan {
    std::vector<UploadPart> part;
    std::vector<uint32_t> band_idx, rest_idx, bytes_idx, all_idx, tiles;
    std::vector<uint64_t> kv, kv_tmp, tile_off, off_rest, off_all, off_bytes;
    std::vector<uint32_t> mat;   // per candidate: its plain-kernel matrix elements
};
}  // namespace

struct srm_job {
    srm_ctx *ctx = nullptr;
    int n = 0;
    HostPlan plan;                   // host arrays of the last upload's plan
    bool all_planned = false;        // the plain-kernel plan of every candidate made
    int ma = 0, mm = 1;
    bool unit = true;
    bool resident_genome = false;
    // device buffers
    DevBuf d_cand, d_read_words, d_win_words, d_read_bytes, d_win_bytes;
    DevBuf d_perm, d_cand_tile, d_tile_score, d_tiles, d_tile_scratch, d_queue, d_scratch;
    DevBuf d_idx_rest, d_idx_all, d_idx_bytes;      // plain index lists
    DevBuf d_off_rest, d_off_all, d_off_bytes;      // matrix offsets (per list, chunk-relative)
    DevBuf d_mat, d_res, d_ops;
    DevBuf d_rec_len, d_rec_off, d_rec_info, d_rec_tmp, d_rec_out, d_rec_pos, d_rec_id,
        d_rec_names, d_rec_hits;                    // device record formatting
    size_t ops_words = 0;
    // record stage of the launch sequence (srm_job_set_records)
    bool rec_stage = false;
    bool rec_names = false, rec_ids = false;
    int rec_stride = 0, rec_prefix_len = 0;
    size_t rec_tmp = 0;               // scan temporary storage
    size_t rec_cap = 0;               // output bytes: a bound on the text of every line
    size_t rec_rows_bound = 0;        // sum over candidates of 2 (w + k): both rows at most
    bool rec_done = false;            // a run since the stage was set up formatted records
    // host-side plan
    std::vector<BandClass> classes;
    std::vector<PlainChunk> chunks_rest, chunks_all, chunks_bytes;
    size_t n_band = 0, n_rest = 0, n_bytes = 0;
    size_t mat_budget_elems = 0;
    size_t mat_elems_band = 0, mat_elems_plain = 0;   // matrix elements a run of each path needs
    // timing
    hipEvent_t ev[4] = {};               // start, after band, after plain, after records
    hipEvent_t cs[17] = {}, ce[17] = {};   // around each band-class launch (a class
                                           // may be split in two by read length)
    hipEvent_t fork_ev = nullptr;        // concurrent classes: fork and joins
    hipEvent_t join_ev[8] = {};
    std::vector<hipStream_t> side;       // streams of concurrent band classes
    int class_streams = 0;               // at the last upload
    float last_ms = 0.f, band_ms = 0.f, plain_ms = 0.f, rec_ms = 0.f;
    float class_ms[9] = {};          // by NW (1..8), last eager run
    int block_waves = 4;             // waves per block of the unit-cost kernels (tuning)
    int walk_ck = 0;                 // unit costs: checkpoint scratch layout (tuning)
    // graph
    bool graph_enabled = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_kernel = -1;
    bool uploaded = false;
};

namespace {

using srm::band_cells;
using srm::tile_scratch_words;

void destroy_graph(srm_job *j)
{
    if (j->graph_exec) (void)hipGraphExecDestroy(j->graph_exec);
    if (j->graph) (void)hipGraphDestroy(j->graph);
    j->graph_exec = nullptr;
    j->graph = nullptr;
    j->graph_kernel = -1;
}

template <class T>
int upload(srm_ctx *ctx, DevBuf &b, const T *src, size_t count)
{
    size_t bytes = count * sizeof(T);
    // 64 bytes of slack: a w == 0 window still reads its row-0 byte
    int rc = ctx->ensure(b, bytes + 64);
    if (rc) return rc;
    if (bytes) HIP_TRY(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    return SRM_OK;
}

// builds chunked matrix offsets for a plain index list
// the plain kernel's int16 matrix of a candidate, in elements, 128-byte aligned
inline uint32_t plain_mat_elems(const srm_cand &c)
{
    return (uint32_t)((((size_t)c.w + 1) * ((size_t)c.L + 1) + 63) & ~(size_t)63);
}

// matrix offsets and chunks of the candidates idx (mat: every candidate's matrix
// elements)
void plan_plain(const std::vector<uint32_t> &idx, const uint32_t *mat, size_t budget_elems,
                std::vector<uint64_t> &off, std::vector<PlainChunk> &chunks)
{
    off.resize(idx.size());
    chunks.clear();
    PlainChunk cur;
    size_t acc = 0;
    for (size_t j = 0; j < idx.size(); ++j) {
        const size_t m = mat[idx[j]];
        if (cur.count && acc + m > budget_elems) {
            cur.mat_elems = acc;
            chunks.push_back(cur);
            cur = PlainChunk();
            cur.begin = (uint32_t)j;
            acc = 0;
        }
        off[j] = acc;
        acc += m;
        cur.count++;
    }
    if (cur.count) {
        cur.mat_elems = acc;
        chunks.push_back(cur);
    }
}

size_t max_chunk(std::initializer_list<const std::vector<PlainChunk> *> vs)
{
    size_t m = 0;
    for (auto *v : vs)
        for (auto &ch : *v) m = std::max(m, ch.mat_elems);
    return m;
}

// the plain kernel's plan over every candidate not on the raw-byte path, in index
// order, and its device lists: made for the first plain run after an upload (a
// banded run never needs it)
int plan_all(srm_job *j)
{
    if (j->all_planned) return SRM_OK;
    HostPlan &hp = j->plan;
    std::vector<uint32_t> &all_idx = hp.all_idx;
    all_idx.clear();
    all_idx.reserve((size_t)j->n - hp.bytes_idx.size());
    size_t b = 0;
    for (uint32_t t = 0; t < (uint32_t)j->n; ++t) {
        if (b < hp.bytes_idx.size() && hp.bytes_idx[b] == t) {
            ++b;
            continue;
        }
        all_idx.push_back(t);
    }
    plan_plain(all_idx, hp.mat.data(), j->mat_budget_elems, hp.off_all, j->chunks_all);
    int rc;
    if ((rc = upload(j->ctx, j->d_idx_all, all_idx.data(), all_idx.size()))) return rc;
    if ((rc = upload(j->ctx, j->d_off_all, hp.off_all.data(), hp.off_all.size()))) return rc;
    HIP_TRY(hipStreamSynchronize(j->ctx->stream));
    j->mat_elems_plain = max_chunk({&j->chunks_all, &j->chunks_bytes});
    j->all_planned = true;
    return SRM_OK;
}

// the record kernels' arguments over the job's buffers (set_up: the record stage
// of the launch sequence, names and ids as srm_job_set_records stored them)
srm::RecArgs rec_args(srm_job *j, bool set_up)
{
    srm_ctx *ctx = j->ctx;
    srm::RecArgs a{};
    a.n = j->n;
    a.cand = (const srm_cand *)j->d_cand.p;
    a.res = (const srm_result *)j->d_res.p;
    a.ops = (const uint32_t *)j->d_ops.p;
    a.read_words = (const uint32_t *)j->d_read_words.p + SRM_CODE_PAD / 16;
    a.win_words = j->resident_genome ? ctx->genome_ptr()
                                     : (const uint32_t *)j->d_win_words.p + SRM_CODE_PAD / 16;
    a.read_bytes = (const char *)j->d_read_bytes.p;
    a.win_bytes = (const char *)j->d_win_bytes.p;
    a.pos = (const int64_t *)j->d_rec_pos.p;
    a.len = (uint64_t *)j->d_rec_len.p;
    a.info = (uint4 *)j->d_rec_info.p;
    a.hits = (uint32_t *)j->d_rec_hits.p;
    a.off = (const uint64_t *)j->d_rec_off.p;
    a.out = (char *)j->d_rec_out.p;
    if (set_up) {
        a.read_id = j->rec_ids ? (const int32_t *)j->d_rec_id.p : nullptr;
        a.names = j->rec_names ? (const char *)j->d_rec_names.p : nullptr;
        a.name_stride = j->rec_stride;
        a.prefix = j->rec_names ? nullptr : (const char *)j->d_rec_names.p;
        a.prefix_len = j->rec_prefix_len;
    }
    return a;
}

int run_launches(srm_job *j, int kernel, bool timed)
{
    srm_ctx *ctx = j->ctx;
    hipStream_t st = ctx->stream;
    const bool use_band = kernel != SRM_KERNEL_PLAIN;
    if (timed) HIP_TRY(hipEventRecord(j->ev[0], st));
    if (use_band && !j->classes.empty()) {
        HIP_TRY(srm::launch_reset((uint32_t *)j->d_queue.p, (int)j->classes.size(), st));
        // with class_streams > 1 the classes run side by side on their own streams
        // (each with its own scratch), widest band first; otherwise in order on st
        const int S = std::min<int>(j->class_streams, (int)j->classes.size());
        std::vector<size_t> order(j->classes.size());
        for (size_t c = 0; c < order.size(); ++c) order[c] = c;
        if (S > 1) {
            std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
                return j->classes[x].nw > j->classes[y].nw;
            });
            HIP_TRY(hipEventRecord(j->fork_ev, st));
            for (int q = 0; q < S; ++q) HIP_TRY(hipStreamWaitEvent(j->side[(size_t)q], j->fork_ev, 0));
        }
        for (size_t o = 0; o < order.size(); ++o) {
            const size_t c = order[o];
            const BandClass &bc = j->classes[c];
            hipStream_t cst = S > 1 ? j->side[o % (size_t)S] : st;
            srm::BandArgs a;
            a.ma = j->ma;
            a.mm = j->mm;
            a.cand = (const srm_cand *)j->d_cand.p;
            a.perm = (const uint32_t *)j->d_perm.p;
            a.cand_tile = (const srm_cand *)j->d_cand_tile.p;
            a.tile_score = (int32_t *)j->d_tile_score.p;
            a.tiles = (const uint4 *)j->d_tiles.p + bc.tile_begin;
            a.ntiles = bc.ntiles;
            a.queue = (uint32_t *)j->d_queue.p + c;
            a.read_words = (const uint32_t *)j->d_read_words.p + SRM_CODE_PAD / 16;
            a.win_words = j->resident_genome
                              ? ctx->genome_ptr()
                              : (const uint32_t *)j->d_win_words.p + SRM_CODE_PAD / 16;
            a.scratch = (uint32_t *)j->d_scratch.p;
            a.tile_scratch = (const uint64_t *)j->d_tile_scratch.p + bc.tile_begin;
            a.res = (srm_result *)j->d_res.p;
            a.ops = (uint32_t *)j->d_ops.p;
            a.block_waves = j->block_waves;
            a.walk_ck = j->walk_ck;
            if (timed) HIP_TRY(hipEventRecord(j->cs[c], cst));
            HIP_TRY(srm::launch_band(a, bc.nw, j->unit, bc.blocks, bc.maxL, cst));
            HIP_TRY(srm::launch_trace(a, bc.nw, j->unit, bc.maxL, cst));
            if (timed) HIP_TRY(hipEventRecord(j->ce[c], cst));
        }
        if (S > 1)
            for (int q = 0; q < S; ++q) {
                HIP_TRY(hipEventRecord(j->join_ev[q], j->side[(size_t)q]));
                HIP_TRY(hipStreamWaitEvent(st, j->join_ev[q], 0));
            }
    }
    if (timed) HIP_TRY(hipEventRecord(j->ev[1], st));
    auto plain = [&](const std::vector<PlainChunk> &chunks, DevBuf &idx, DevBuf &off,
                     bool codes) -> int {
        for (const PlainChunk &ch : chunks) {
            srm::PlainArgs a;
            a.ma = j->ma;
            a.mm = j->mm;
            a.n = (int)ch.count;
            a.idx = (const uint32_t *)idx.p + ch.begin;
            a.cand = (const srm_cand *)j->d_cand.p;
            a.read_words = (const uint32_t *)j->d_read_words.p + SRM_CODE_PAD / 16;
            a.win_words = j->resident_genome
                              ? ctx->genome_ptr()
                              : (const uint32_t *)j->d_win_words.p + SRM_CODE_PAD / 16;
            a.read_bytes = (const char *)j->d_read_bytes.p;
            a.win_bytes = (const char *)j->d_win_bytes.p;
            a.mat = (int16_t *)j->d_mat.p;
            a.mat_off = (const uint64_t *)off.p + ch.begin;
            a.res = (srm_result *)j->d_res.p;
            a.ops = (uint32_t *)j->d_ops.p;
            HIP_TRY(srm::launch_plain(a, codes, st));
        }
        return SRM_OK;
    };
    int rc;
    if (use_band)
        rc = plain(j->chunks_rest, j->d_idx_rest, j->d_off_rest, true);
    else
        rc = plain(j->chunks_all, j->d_idx_all, j->d_off_all, true);
    if (rc) return rc;
    rc = plain(j->chunks_bytes, j->d_idx_bytes, j->d_off_bytes, false);
    if (rc) return rc;
    if (timed) HIP_TRY(hipEventRecord(j->ev[2], st));
    if (j->rec_stage) {
        // records of every hit into the output sized at set-up: counter reset (a
        // kernel, so that a graph can hold it), line lengths, their scan, the text
        srm::RecArgs a = rec_args(j, true);
        HIP_TRY(srm::launch_reset(a.hits, 2, st));
        HIP_TRY(srm::launch_rec_measure(a, st));
        size_t tmp = j->rec_tmp;
        HIP_TRY(srm::rec_scan(j->d_rec_tmp.p, tmp, a.len, (uint64_t *)j->d_rec_off.p, j->n, st));
        HIP_TRY(srm::launch_rec_write(a, st));
    }
    if (timed) HIP_TRY(hipEventRecord(j->ev[3], st));
    j->rec_done = j->rec_stage;
    return SRM_OK;
}

// out[p] = cand[perm[p]], p < n: the banded descriptors in tile order, gathered
// from the uploaded descriptors (perm holds indexes below their count)
__global__ void gather_cand_kernel(const srm_cand *__restrict__ cand,
                                   const uint32_t *__restrict__ perm, srm_cand *__restrict__ out,
                                   size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride)
        out[p] = cand[perm[p]];
}

hipError_t launch_gather_cand(const srm_cand *cand, const uint32_t *perm, srm_cand *out, size_t n,
                              hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(gather_cand_kernel, dim3((unsigned)blocks), dim3(256), 0, st, cand, perm,
                       out, n);
    return hipGetLastError();
}

// host threads for a planning pass over n items: one below 64k, else up to 8
unsigned plan_threads(size_t n)
{
    if (n < 65536) return 1;
    return std::min(8u, std::max(1u, std::thread::hardware_concurrency()));
}

// fn(u, lo, hi) on T contiguous ranges of [0, n), range u on thread u
template <class Fn>
void par_ranges(unsigned T, size_t n, Fn fn)
{
    if (T <= 1) {
        fn(0u, (size_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(T);
    for (unsigned u = 0; u < T; ++u) {
        try {
            th.emplace_back(fn, u, n * u / T, n * (u + 1) / T);
        } catch (...) {   // no thread to be had: the range on this one
            fn(u, n * u / T, n * (u + 1) / T);
        }
    }
    for (auto &x : th) x.join();
}

// stable sort of kv by bits [lo, lo + bits) of each element: least-significant-
// digit counting passes of at most 11 bits, each split over T threads (a thread
// counts its range's digits, then scatters that range to the offsets its counts
// give, so equal digits keep their order)
void radix_sort_kv(std::vector<uint64_t> &kv, std::vector<uint64_t> &tmp, int lo, int bits,
                   unsigned T)
{
    const size_t n = kv.size();
    const int passes = (bits + 10) / 11;
    if (n < 2 || passes == 0) return;
    const int db = (bits + passes - 1) / passes;
    const size_t B = (size_t)1 << db;
    tmp.resize(n);
    std::vector<size_t> cnt((size_t)T * B);
    for (int pass = 0; pass < passes; ++pass) {
        const int sh = lo + pass * db;
        const uint64_t mask = B - 1;
        std::fill(cnt.begin(), cnt.end(), 0);
        par_ranges(T, n, [&](unsigned u, size_t a, size_t b) {
            size_t *c = cnt.data() + (size_t)u * B;
            for (size_t i = a; i < b; ++i) ++c[(kv[i] >> sh) & mask];
        });
        size_t run = 0;
        for (size_t d = 0; d < B; ++d)
            for (unsigned u = 0; u < T; ++u) {
                const size_t c = cnt[(size_t)u * B + d];
                cnt[(size_t)u * B + d] = run;
                run += c;
            }
        par_ranges(T, n, [&](unsigned u, size_t a, size_t b) {
            size_t *c = cnt.data() + (size_t)u * B;
            for (size_t i = a; i < b; ++i) tmp[c[(kv[i] >> sh) & mask]++] = kv[i];
        });
        kv.swap(tmp);
    }
}
}  // namespace

namespace srm {

int set_error(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int ctx_device(const srm_ctx *c) { return c->device; }

hipStream_t ctx_stream(const srm_ctx *c) { return c->stream; }

}  // namespace srm

// ---------------------------------------------------------------------------
// extended API
// ---------------------------------------------------------------------------
extern "C" {

const char *srm_version(void) { return "saruman-mi355x 0.1 (gfx950)"; }

const char *srm_last_error(void) { return g_err.c_str(); }

int srm_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

void *srm_host_alloc(size_t bytes)
{
    void *p = nullptr;
    if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
        fail(SRM_ENOMEM, "hipHostMalloc(%zu) failed", bytes);
        return nullptr;
    }
    return p;
}

void srm_host_free(void *p)
{
    if (p) (void)hipHostFree(p);
}

int srm_device_sync(int device)
{
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipDeviceSynchronize());
    return SRM_OK;
}

int srm_ctx_create(int device, size_t pool_cap_bytes, srm_ctx **out)
{
    *out = nullptr;
    int n = srm_device_count();
    if (n <= 0) return fail(SRM_ENODEV, "no HIP device visible");
    if (device < 0 || device >= n) return fail(SRM_EINVAL, "device %d out of range", device);
    HIP_TRY(hipSetDevice(device));
    srm_ctx *c = new srm_ctx();
    c->device = device;
    c->pool_cap = pool_cap_bytes;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
        delete c;
        return fail(SRM_EDEVICE, "hipGetDeviceProperties failed");
    }
    c->num_cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(SRM_EDEVICE, "hipStreamCreate failed");
    }
    *out = c;
    return SRM_OK;
}

void srm_job_destroy(srm_job *j);

void srm_ctx_destroy(srm_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->dropin_job) srm_job_destroy(c->dropin_job);
    (void)hipStreamSynchronize(c->stream);
    c->release(c->genome);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

int srm_ctx_set_genome(srm_ctx *c, const uint32_t *words, size_t n_words)
{
    HIP_TRY(hipSetDevice(c->device));
    int rc = c->ensure(c->genome, n_words * 4);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(c->genome.p, words, n_words * 4, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->genome_words = n_words;
    c->genome_borrowed = nullptr;
    return SRM_OK;
}

int srm_ctx_set_tuning(srm_ctx *ctx, const srm_tuning *t)
{
    if (!t) {
        ctx->tuning = srm_tuning{};
        return SRM_OK;
    }
    for (int i = 0; i < 9; ++i)
        if (t->blocks_per_cu[i] < 0 || t->blocks_per_cu[i] > 64) return fail(SRM_EINVAL, "blocks_per_cu");
    if (t->merge_below < 0 || t->merge_below > 8) return fail(SRM_EINVAL, "merge_below");
    if (t->class_streams < 0 || t->class_streams > 8) return fail(SRM_EINVAL, "class_streams");
    if (t->tile_cands != 0 && t->tile_cands != 32 && t->tile_cands != 64)
        return fail(SRM_EINVAL, "tile_cands");
    if (t->block_waves != 0 && t->block_waves != 1 && t->block_waves != 2 && t->block_waves != 4)
        return fail(SRM_EINVAL, "block_waves");
    if (t->split_len < 0 || t->split_len > 65535) return fail(SRM_EINVAL, "split_len");
    if (t->walk_layout != 0 && t->walk_layout != 1) return fail(SRM_EINVAL, "walk_layout");
    ctx->tuning = *t;
    return SRM_OK;
}

int srm_job_band_info(srm_job *j, int nw, int *blocks, int *ntiles, float *ms)
{
    if (nw < 1 || nw > 8) return fail(SRM_EINVAL, "nw");
    *blocks = 0;
    *ntiles = 0;
    *ms = j->class_ms[nw];
    for (const auto &bc : j->classes)
        if (bc.nw == nw) {   // the two halves of a class split by read length add up
            *blocks += bc.blocks;
            *ntiles += (int)bc.ntiles;
        }
    return SRM_OK;
}

int srm_band_occupancy(int nw, int unit) { return srm::band_occupancy(nw, unit != 0); }

int srm_ctx_borrow_genome(srm_ctx *dst, const srm_ctx *src)
{
    if (!src->genome_ptr()) return fail(SRM_ESTATE, "source context has no genome");
    if (dst->device != src->device) return fail(SRM_EINVAL, "contexts on different devices");
    dst->release(dst->genome);
    dst->genome_borrowed = src->genome_ptr();
    dst->genome_words = src->genome_words;
    return SRM_OK;
}

int srm_job_create(srm_ctx *ctx, srm_job **out)
{
    HIP_TRY(hipSetDevice(ctx->device));
    srm_job *j = new srm_job();
    j->ctx = ctx;
    bool ok = true;
    for (auto &e : j->ev) ok = ok && hipEventCreate(&e) == hipSuccess;
    for (auto &e : j->cs) ok = ok && hipEventCreate(&e) == hipSuccess;
    for (auto &e : j->ce) ok = ok && hipEventCreate(&e) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&j->fork_ev, hipEventDisableTiming) == hipSuccess;
    for (auto &e : j->join_ev) ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        srm_job_destroy(j);
        return fail(SRM_EDEVICE, "hipEventCreate failed");
    }
    *out = j;
    return SRM_OK;
}

namespace {
// frees every device buffer of a job (its context's pool shrinks by as much)
void release_job_buffers(srm_job *j)
{
    srm_ctx *c = j->ctx;
    DevBuf *bufs[] = {&j->d_cand, &j->d_read_words, &j->d_win_words, &j->d_read_bytes,
                      &j->d_win_bytes, &j->d_perm, &j->d_cand_tile, &j->d_tile_score,
                      &j->d_tiles, &j->d_tile_scratch,
                      &j->d_queue, &j->d_scratch,
                      &j->d_idx_rest, &j->d_idx_all, &j->d_idx_bytes, &j->d_off_rest,
                      &j->d_off_all, &j->d_off_bytes, &j->d_mat, &j->d_res, &j->d_ops,
                      &j->d_rec_len, &j->d_rec_off, &j->d_rec_info, &j->d_rec_tmp,
                      &j->d_rec_out, &j->d_rec_pos, &j->d_rec_id, &j->d_rec_names,
                      &j->d_rec_hits};
    for (DevBuf *b : bufs) c->release(*b);
}
}  // namespace

int srm_job_release(srm_job *j)
{
    srm_ctx *c = j->ctx;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (hipStream_t q : j->side) HIP_TRY(hipStreamSynchronize(q));
    destroy_graph(j);
    release_job_buffers(j);
    j->plan = HostPlan();   // and the host arrays an upload keeps for the next one
    j->all_planned = false;
    j->uploaded = false;
    return SRM_OK;
}

size_t srm_ctx_pool_used(const srm_ctx *c) { return c->pool_used; }

int srm_ctx_set_pool_cap(srm_ctx *c, size_t bytes)
{
    c->pool_cap = bytes;
    return SRM_OK;
}

void srm_debug_fail_allocs(int count) { g_fail_allocs.store(count < 0 ? 0 : count); }

void srm_job_destroy(srm_job *j)
{
    if (!j) return;
    srm_ctx *c = j->ctx;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    destroy_graph(j);
    release_job_buffers(j);
    for (hipStream_t q : j->side) {
        (void)hipStreamSynchronize(q);
        (void)hipStreamDestroy(q);
    }
    for (auto *arr : {&j->ev[0], &j->cs[0], &j->ce[0], &j->join_ev[0]}) {
        const int cnt = arr == &j->ev[0] ? 4 : arr == &j->join_ev[0] ? 8 : 17;
        for (int i = 0; i < cnt; ++i)
            if (arr[i]) (void)hipEventDestroy(arr[i]);
    }
    if (j->fork_ev) (void)hipEventDestroy(j->fork_ev);
    delete j;
}

size_t srm_job_bytes_bound(const srm_ctx *ctx, int match, int mismatch, const srm_cand *cand, int n,
                           size_t n_read_words, size_t n_win_words, size_t n_read_bytes,
                           size_t n_win_bytes, int kernel, int with_records, size_t names_bytes,
                           int name_max)
{
    // each term is one buffer as srm_job_upload, srm_job_run and the record stage
    // ask for it, rounded as ensure() rounds under a pool cap
    auto r4k = [](size_t b) { return (std::max<size_t>(b, 1) + 4095) & ~(size_t)4095; };
    if (n < 0) return 0;
    const size_t N = (size_t)n;
    // the upload plans the band for every eligible candidate whatever kernel runs;
    // the kernel decides which candidates need full matrices
    const bool band_ok = match >= 0 && mismatch >= 1;
    const bool all_plain = kernel == SRM_KERNEL_PLAIN;
    const bool unit = match == 0 && mismatch == 1;
    size_t n_cls[9] = {}, maxL_cls[9] = {};
    size_t ops = 0, rows = 0, plain_elems = 0;
    bool any_plain = false;
    for (int t = 0; t < n; ++t) {
        const srm_cand &c = cand[t];
        ops = std::max(ops, (size_t)c.ops_off + srm_ops_words(c.w, c.k));
        rows += 2 * ((size_t)c.w + c.k);
        const bool banded = band_ok && !(c.bits & SRM_CAND_BYTES) && c.k <= SRM_BAND_KMAX &&
                            band_cells(c) <= SRM_BAND_MAX && (int)c.w >= (int)c.L - (int)c.k;
        if (banded) {
            const int merge = ctx ? std::max(1, std::min(8, ctx->tuning.merge_below)) : 1;
            const int nw = std::min(8, std::max((band_cells(c) + 7) / 8, merge));
            ++n_cls[nw];
            maxL_cls[nw] = std::max<size_t>(maxL_cls[nw], c.L);
        }
        if (!banded || all_plain) {
            any_plain = true;
            plain_elems = std::max(plain_elems, (size_t)(c.w + 1) * (size_t)(c.L + 1) + 64);
        }
    }
    size_t tiles = 0, scratch = 0;
    for (int nw = 1; nw <= 8; ++nw) {
        if (!n_cls[nw]) continue;
        // every (class, L mod 8) run ends with at most one partial tile (and a
        // class split by read length has twice the runs)
        const size_t tcands = ctx && ctx->tuning.tile_cands == 32 ? 32 : 64;
        const size_t runs = ctx && ctx->tuning.split_len > 0 ? 16 : 8;
        const size_t tc = (n_cls[nw] + tcands - 1) / tcands + runs;
        tiles += tc;
        scratch += tc * (unit ? srm::bits_tile_words(nw, (int)maxL_cls[nw],
                                                     ctx && ctx->tuning.walk_layout == 1)
                              : tile_scratch_words(nw, (int)maxL_cls[nw]));
    }
    size_t b = 0;
    b += r4k(N * sizeof(srm_cand) + 64) * 2;                  // d_cand, d_cand_tile
    b += r4k(n_read_words * 4 + 64);
    if (n_win_words) b += r4k(n_win_words * 4 + 64);
    b += r4k(n_read_bytes + 64) + r4k(n_win_bytes + 64);
    b += r4k(N * 4 + 64) + r4k(std::max<size_t>(N * 4, 64));  // d_perm, d_tile_score
    b += r4k(tiles * 16 + 64) + r4k(tiles * 8 + 64);          // d_tiles, d_tile_scratch
    b += 3 * r4k(N * 4 + 64) + 3 * r4k(N * 8 + 64);           // plain index and offset lists
    b += r4k(64 * 4) + r4k(std::max<size_t>(scratch * 4, 256));
    b += r4k(std::max<size_t>(N * sizeof(srm_result), 64)) + r4k(std::max<size_t>(ops * 4, 64));
    if (any_plain)   // a chunk of matrices under a pool cap, at least one matrix
        b += r4k(std::max<size_t>(SRM_PLAIN_CHUNK, plain_elems * 2) + 256);
    if (with_records) {
        b += r4k(N * 8 + 64) + r4k(N * 4 + 64) + r4k(names_bytes + 64);
        b += 2 * r4k((N + 1) * 8) + r4k(std::max<size_t>(N * 16, 16)) + r4k(256);
        size_t tmp = 0;
        (void)srm::rec_scan(nullptr, tmp, nullptr, nullptr, n, nullptr);
        b += r4k(std::max<size_t>(tmp, 256));
        b += r4k(rows + N * ((size_t)std::max(name_max, 0) + 11 + 11 + 2 + 6 + 7));
    }
    return b;
}

int srm_job_upload(srm_job *j, int match, int mismatch, const srm_cand *cand, int n,
                   const uint32_t *read_words, size_t n_read_words, const uint32_t *win_words,
                   size_t n_win_words, const char *read_bytes, size_t n_read_bytes,
                   const char *win_bytes, size_t n_win_bytes)
{
    srm_ctx *ctx = j->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    destroy_graph(j);
    j->uploaded = false;
    j->rec_stage = false;   // positions and ids belong to the previous candidates
    if (n < 0) return fail(SRM_EINVAL, "negative candidate count");
    if (!win_words && !ctx->genome_ptr() && n_win_bytes == 0)
        return fail(SRM_EINVAL, "no window source: pass win_words or set a resident genome");
    j->n = n;
    j->ma = match;
    j->mm = mismatch;
    j->unit = (match == 0 && mismatch == 1);
    j->resident_genome = (win_words == nullptr);
    const bool band_ok = match >= 0 && mismatch >= 1;

    // ---- validate and classify, over contiguous ranges of candidates on several
    // host threads; the lists keep index order (ranges concatenated in order) and
    // the first failing candidate (lowest index) is the one reported
    const uint64_t read_codes = (uint64_t)n_read_words * 16;
    const uint64_t win_codes =
        j->resident_genome ? (uint64_t)ctx->genome_words * 16 : (uint64_t)n_win_words * 16;
    const uint64_t pad = SRM_CODE_PAD;
    enum { K_BAND, K_REST, K_BYTES, E_RANGE, E_BYTES, E_READ, E_GENOME, E_WIN };
    auto classify = [&](const srm_cand &c) -> int {
        if (c.L < 1 || c.L > SRM_MAX_L || c.w > SRM_MAX_W) return E_RANGE;
        if (c.bits & SRM_CAND_BYTES) {
            if ((uint64_t)c.read_off + c.L > n_read_bytes || (uint64_t)c.win_off + c.w > n_win_bytes)
                return E_BYTES;
            return K_BYTES;
        }
        // code arrays: pad codes readable before 0 and after the end (the caller's
        // arrays start with SRM_CODE_PAD pad codes for uploaded arrays, and offsets
        // count from the first code after them).  Written without subtraction: an
        // array shorter than the padding must fail, not wrap.
        if ((uint64_t)c.read_off + c.L + 2 * pad > read_codes) return E_READ;
        if (j->resident_genome) {
            // the genome carries SRM_CODE_PAD codes of padding at both ends
            if ((uint64_t)c.win_off < pad || (uint64_t)c.win_off + c.w + pad > win_codes)
                return E_GENOME;
        } else if ((uint64_t)c.win_off + c.w + 2 * pad > win_codes) {
            return E_WIN;
        }
        // banded: exact-band conditions, and w >= L - k (a shorter window can never
        // hold a hit; the window code stream reads up to L + 93 codes past win_off)
        if (band_ok && c.k <= SRM_BAND_KMAX && band_cells(c) <= SRM_BAND_MAX &&
            (int)c.w >= (int)c.L - (int)c.k)
            return K_BAND;
        return K_REST;
    };
    const unsigned TV = plan_threads((size_t)n);
    HostPlan &hp = j->plan;
    std::vector<UploadPart> &part = hp.part;
    part.resize(TV);
    for (UploadPart &P : part) P.reset();
    hp.mat.resize((size_t)n);
    par_ranges(TV, (size_t)n, [&](unsigned u, size_t a, size_t b) {
        UploadPart &P = part[u];
        for (size_t t = a; t < b; ++t) {
            const srm_cand &c = cand[t];
            const int r = classify(c);
            if (r >= E_RANGE) {
                P.bad = (int)t;
                P.why = r;
                return;
            }
            hp.mat[t] = plain_mat_elems(c);
            P.ops_end = std::max(P.ops_end, (size_t)c.ops_off + srm_ops_words(c.w, c.k));
            P.rows += 2 * ((size_t)c.w + c.k);
            if (r == K_BAND) P.maxL = std::max<uint32_t>(P.maxL, c.L);
            (r == K_BAND ? P.band : r == K_REST ? P.rest : P.bytes).push_back((uint32_t)t);
        }
    });
    std::vector<uint32_t> &band_idx = hp.band_idx, &rest_idx = hp.rest_idx,
                          &bytes_idx = hp.bytes_idx;
    band_idx.clear();
    rest_idx.clear();
    bytes_idx.clear();
    size_t ops_end = 0;
    j->rec_rows_bound = 0;
    for (const UploadPart &P : part) {
        if (P.bad >= 0) {
            const srm_cand &c = cand[P.bad];
            switch (P.why) {
            case E_RANGE:
                return fail(SRM_EINVAL, "candidate %d: L=%d w=%d outside 1..%d / 0..%d", P.bad,
                            c.L, c.w, SRM_MAX_L, SRM_MAX_W);
            case E_BYTES: return fail(SRM_EINVAL, "candidate %d: byte range out of bounds", P.bad);
            case E_READ: return fail(SRM_EINVAL, "candidate %d: read codes out of bounds", P.bad);
            case E_GENOME:
                return fail(SRM_EINVAL, "candidate %d: genome window out of bounds", P.bad);
            default: return fail(SRM_EINVAL, "candidate %d: window codes out of bounds", P.bad);
            }
        }
        ops_end = std::max(ops_end, P.ops_end);
        j->rec_rows_bound += P.rows;
    }
    if (TV == 1) {
        band_idx.swap(part[0].band);
        rest_idx.swap(part[0].rest);
        bytes_idx.swap(part[0].bytes);
    } else {
        for (const UploadPart &P : part) {
            band_idx.insert(band_idx.end(), P.band.begin(), P.band.end());
            rest_idx.insert(rest_idx.end(), P.rest.begin(), P.rest.end());
            bytes_idx.insert(bytes_idx.end(), P.bytes.begin(), P.bytes.end());
        }
    }
    j->n_band = band_idx.size();
    j->n_rest = rest_idx.size();
    j->n_bytes = bytes_idx.size();
    j->ops_words = ops_end;

    // ---- banded plan: sort by (class, L mod 8, L desc), 64-candidate tiles that
    // never mix L mod 8: a tile's lanes then cross the 8-column blocks of their
    // band scratch at the same traceback steps (a walk moves one column per
    // diagonal step), so the wave reloads band words together
    // cls = 2 nw + (L >= split_len): a class split by read length (tuning) runs as
    // two launches, the short reads with a smaller LDS staging size
    const int merge_below = std::max(1, std::min(8, ctx->tuning.merge_below));
    const int split_len = ctx->tuning.split_len;
    const uint32_t tile_cands = ctx->tuning.tile_cands == 32 ? 32u : 64u;
    j->block_waves = ctx->tuning.block_waves ? ctx->tuning.block_waves : 4;
    j->walk_ck = ctx->tuning.walk_layout;
    const size_t nb = band_idx.size();
    const unsigned TB = plan_threads(nb);
    // sort word of every banded candidate, in index order: its index (bits 0-31),
    // k (bits 32-35) and from bit 36 the key: class (5 bits), L mod 8 (3 bits),
    // then maxL - L in lb bits (L descending), the fields of the order above from
    // the top down
    uint32_t maxL = 0;
    for (const UploadPart &P : part) maxL = std::max(maxL, P.maxL);
    int lb = 1;
    while ((1u << lb) <= maxL) ++lb;
    std::vector<uint64_t> &kv = hp.kv;
    kv.resize(nb);
    par_ranges(TB, nb, [&](unsigned, size_t a, size_t b) {
        for (size_t i = a; i < b; ++i) {
            const srm_cand &c = cand[band_idx[i]];
            const uint32_t cl = 2u * (uint32_t)std::max((band_cells(c) + 7) / 8, merge_below) +
                                (split_len > 0 && (int)c.L >= split_len ? 1u : 0u);
            const uint32_t key = (((cl << 3) | ((uint32_t)c.L & 7u)) << lb) | (maxL - c.L);
            kv[i] = ((uint64_t)key << 36) | ((uint64_t)c.k << 32) | band_idx[i];
        }
    });
    static_assert(SRM_BAND_KMAX < 16, "k of a banded candidate fits 4 bits");
    radix_sort_kv(kv, hp.kv_tmp, 36, 8 + lb, TB);
    par_ranges(TB, nb, [&](unsigned, size_t a, size_t b) {
        for (size_t i = a; i < b; ++i) band_idx[i] = (uint32_t)kv[i];
    });
    // by position in tile order
    auto cls_at = [&](size_t i) { return (int)(kv[i] >> (36 + lb + 3)); };
    auto lmod_at = [&](size_t i) { return (int)((kv[i] >> (36 + lb)) & 7u); };
    auto k_at = [&](size_t i) { return (uint32_t)(kv[i] >> 32) & 15u; };
    auto L_at = [&](size_t i) { return maxL - ((uint32_t)(kv[i] >> 36) & ((1u << lb) - 1u)); };
    std::vector<uint32_t> &tiles = hp.tiles;   // uint4 per tile
    tiles.clear();
    j->classes.clear();
    size_t scratch_words = 0;
    std::vector<uint64_t> &tile_off = hp.tile_off;   // word offset of each tile's band
    tile_off.clear();
    // band classes run side by side on 4 streams unless the tuning says otherwise
    // (1 = one after another on the context's stream)
    const int want_streams = ctx->tuning.class_streams == 0 ? 4 : ctx->tuning.class_streams;
    const bool concurrent = want_streams > 1;
    j->class_streams = concurrent ? std::min(want_streams, 8) : 0;
    while ((int)j->side.size() < j->class_streams) {   // default priority, like every stream here
        hipStream_t q = nullptr;
        HIP_TRY(hipStreamCreateWithFlags(&q, hipStreamNonBlocking));
        j->side.push_back(q);
    }
    {
        size_t i = 0;
        while (i < band_idx.size()) {
            BandClass bc;
            const int ci = cls_at(i);
            bc.nw = ci / 2;
            bc.tile_begin = (uint32_t)(tiles.size() / 4);
            bc.maxL = 0;
            // the sort words are in key order: the class's end and each run's end by
            // binary search
            const size_t e = (size_t)(std::partition_point(
                                          kv.begin() + (std::ptrdiff_t)i, kv.end(),
                                          [&](const uint64_t &x) {
                                              return (int)(x >> (36 + lb + 3)) == ci;
                                          }) -
                                      kv.begin());
            for (size_t g = i; g < e;) {
                // one run of equal L mod 8, L descending: a tile's first read is its
                // longest (its k maximum is filled in below, over all tiles at once)
                const int lm = lmod_at(g);
                const size_t ge = (size_t)(std::partition_point(
                                               kv.begin() + (std::ptrdiff_t)g,
                                               kv.begin() + (std::ptrdiff_t)e,
                                               [&](const uint64_t &x) {
                                                   return (int)((x >> (36 + lb)) & 7u) == lm;
                                               }) -
                                           kv.begin());
                for (size_t s = g; s < ge; s += tile_cands) {
                    const uint32_t cnt = (uint32_t)std::min<size_t>(tile_cands, ge - s);
                    const uint32_t lmax = L_at(s);
                    tiles.push_back((uint32_t)s);
                    tiles.push_back(cnt);
                    tiles.push_back(lmax);
                    tiles.push_back(0u);
                    bc.maxL = std::max(bc.maxL, (int)lmax);
                }
                g = ge;
            }
            bc.ntiles = (uint32_t)(tiles.size() / 4) - bc.tile_begin;
            {
                // longest tiles first (each tile keeps one L mod 8): the persistent
                // waves then end on the short ones, and the class's last wave
                // finishes soon after the others
                std::vector<std::array<uint32_t, 4>> tv(bc.ntiles);
                std::memcpy(tv.data(), tiles.data() + 4 * (size_t)bc.tile_begin,
                            sizeof(uint32_t) * 4 * bc.ntiles);
                std::stable_sort(tv.begin(), tv.end(),
                                 [](const std::array<uint32_t, 4> &x,
                                    const std::array<uint32_t, 4> &y) { return x[2] > y[2]; });
                std::memcpy(tiles.data() + 4 * (size_t)bc.tile_begin, tv.data(),
                            sizeof(uint32_t) * 4 * bc.ntiles);
            }
            if (match == 0 && mismatch == 1) {
                // bits_band_kernel: one wave per tile, blocks of block_waves waves
                bc.blocks = (int)((bc.ntiles + j->block_waves - 1) / j->block_waves);
            } else {
                // persistent grid: as many 4-wave blocks as stay resident on every
                // CU (occupancy of this kernel instance), never more waves than tiles
                int per_cu = srm::band_occupancy(bc.nw, false);
                const int tuned = ctx->tuning.blocks_per_cu[bc.nw];
                if (tuned > 0) per_cu = std::min(per_cu, tuned);
                int max_blocks = std::max(1, ctx->num_cus * per_cu);
                bc.blocks = (int)std::min<size_t>((bc.ntiles + 3) / 4, (size_t)max_blocks);
            }
            // every tile keeps its band until trace_kernel has walked it
            for (uint32_t u = 0; u < bc.ntiles; ++u) {
                tile_off.push_back(scratch_words);
                const int tmax = (int)tiles[4 * (bc.tile_begin + u) + 2];
                scratch_words += (match == 0 && mismatch == 1)
                                     ? srm::bits_tile_words(bc.nw, tmax, j->walk_ck != 0)
                                     : tile_scratch_words(bc.nw, tmax);
            }
            j->classes.push_back(bc);
            i = e;
        }
    }
    // every tile's largest k, over the tiles on several host threads
    {
        const size_t nt = tiles.size() / 4;
        par_ranges(plan_threads(nt * tile_cands), nt, [&](unsigned, size_t a, size_t b) {
            for (size_t t = a; t < b; ++t) {
                const size_t s0 = tiles[4 * t], c0 = tiles[4 * t + 1];
                uint32_t kmax = 0;
                for (size_t u = s0; u < s0 + c0; ++u) kmax = std::max(kmax, k_at(u));
                tiles[4 * t + 3] = kmax;
            }
        });
    }

    // ---- device buffers
    int rc;
    if ((rc = upload(ctx, j->d_cand, cand, (size_t)n))) return rc;
    if ((rc = upload(ctx, j->d_read_words, read_words, n_read_words))) return rc;
    if (!j->resident_genome && (rc = upload(ctx, j->d_win_words, win_words, n_win_words))) return rc;
    if ((rc = upload(ctx, j->d_read_bytes, read_bytes, read_bytes ? n_read_bytes : 0))) return rc;
    if ((rc = upload(ctx, j->d_win_bytes, win_bytes, win_bytes ? n_win_bytes : 0))) return rc;
    if ((rc = upload(ctx, j->d_perm, band_idx.data(), band_idx.size()))) return rc;
    // the banded candidates in tile order (a wave reads its tile's descriptors as
    // one contiguous run, without the indirection through perm), gathered on the
    // device from the uploaded descriptors
    if ((rc = ctx->ensure(j->d_cand_tile, band_idx.size() * sizeof(srm_cand) + 64))) return rc;
    HIP_TRY(launch_gather_cand((const srm_cand *)j->d_cand.p, (const uint32_t *)j->d_perm.p,
                               (srm_cand *)j->d_cand_tile.p, band_idx.size(), ctx->stream));
    if ((rc = ctx->ensure(j->d_tile_score, std::max<size_t>(band_idx.size() * 4, 64)))) return rc;
    if ((rc = upload(ctx, j->d_tiles, tiles.data(), tiles.size()))) return rc;
    if ((rc = upload(ctx, j->d_tile_scratch, tile_off.data(), tile_off.size()))) return rc;
    if ((rc = upload(ctx, j->d_idx_rest, rest_idx.data(), rest_idx.size()))) return rc;
    if ((rc = upload(ctx, j->d_idx_bytes, bytes_idx.data(), bytes_idx.size()))) return rc;
    if ((rc = ctx->ensure(j->d_queue, 64 * sizeof(uint32_t)))) return rc;
    if ((rc = ctx->ensure(j->d_scratch, std::max<size_t>(scratch_words * 4, 256)))) return rc;
    if ((rc = ctx->ensure(j->d_res, std::max<size_t>((size_t)n * sizeof(srm_result), 64)))) return rc;
    if ((rc = ctx->ensure(j->d_ops, std::max<size_t>(ops_end * 4, 64)))) return rc;

    // ---- plain plans: full matrices in chunks of at most a quarter of free HBM (32
    // GB at most) -- or, under a pool cap, of SRM_PLAIN_CHUNK bytes, the room
    // srm_job_bytes_bound reserves.  A chunk holds at least one candidate.  The
    // matrix buffer itself is allocated by srm_job_run, for the path it runs: a
    // banded run whose candidates all fit the band needs none.
    // (the plan of the plain kernel over every candidate waits for a run of it:
    // plan_all)
    std::vector<uint64_t> &off_rest = hp.off_rest, &off_bytes = hp.off_bytes;
    size_t budget = SRM_PLAIN_CHUNK;
    if (!ctx->pool_cap) {
        size_t free_b = 0, total_b = 0;
        (void)hipMemGetInfo(&free_b, &total_b);
        budget = std::max<size_t>(std::min<size_t>(free_b / 4, (size_t)32 << 30), SRM_PLAIN_CHUNK);
    }
    j->mat_budget_elems = budget / 2;
    plan_plain(rest_idx, hp.mat.data(), j->mat_budget_elems, off_rest, j->chunks_rest);
    plan_plain(bytes_idx, hp.mat.data(), j->mat_budget_elems, off_bytes, j->chunks_bytes);
    j->mat_elems_band = max_chunk({&j->chunks_rest, &j->chunks_bytes});
    j->chunks_all.clear();
    j->mat_elems_plain = 0;
    j->all_planned = false;
    if ((rc = upload(ctx, j->d_off_rest, off_rest.data(), off_rest.size()))) return rc;
    if ((rc = upload(ctx, j->d_off_bytes, off_bytes.data(), off_bytes.size()))) return rc;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    j->uploaded = true;
    return SRM_OK;
}

int srm_job_run(srm_job *j, int kernel)
{
    if (!j->uploaded) return fail(SRM_ESTATE, "srm_job_run before a successful upload");
    srm_ctx *ctx = j->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    if (kernel == SRM_KERNEL_PLAIN) {   // its plan over every candidate, on its first run
        const int rc = plan_all(j);
        if (rc) return rc;
    }
    // the plain kernel's matrices for the candidates this run sends there (before
    // any capture: a graph holds the buffer's address)
    const size_t mat = kernel == SRM_KERNEL_PLAIN ? j->mat_elems_plain : j->mat_elems_band;
    if (mat) {
        const void *before = j->d_mat.p;
        int rc = ctx->ensure(j->d_mat, mat * 2 + 256);
        if (rc) return rc;
        if (j->d_mat.p != before) destroy_graph(j);
    }
    if (j->graph_enabled) {
        if (!j->graph_exec || j->graph_kernel != kernel) {
            destroy_graph(j);
            HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            int rc = run_launches(j, kernel, false);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(ctx->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return fail(SRM_EDEVICE, "graph capture failed: %s", hipGetErrorString(e));
            j->graph = g;
            HIP_TRY(hipGraphInstantiate(&j->graph_exec, g, nullptr, nullptr, 0));
            j->graph_kernel = kernel;
        }
        HIP_TRY(hipEventRecord(j->ev[0], ctx->stream));
        HIP_TRY(hipGraphLaunch(j->graph_exec, ctx->stream));
        HIP_TRY(hipEventRecord(j->ev[3], ctx->stream));
        HIP_TRY(hipEventSynchronize(j->ev[3]));
        HIP_TRY(hipEventElapsedTime(&j->last_ms, j->ev[0], j->ev[3]));
        j->band_ms = j->plain_ms = j->rec_ms = -1.f;
        return SRM_OK;
    }
    int rc = run_launches(j, kernel, true);
    if (rc) return rc;
    HIP_TRY(hipEventSynchronize(j->ev[3]));
    HIP_TRY(hipEventElapsedTime(&j->last_ms, j->ev[0], j->ev[3]));
    HIP_TRY(hipEventElapsedTime(&j->band_ms, j->ev[0], j->ev[1]));
    HIP_TRY(hipEventElapsedTime(&j->plain_ms, j->ev[1], j->ev[2]));
    HIP_TRY(hipEventElapsedTime(&j->rec_ms, j->ev[2], j->ev[3]));
    for (auto &x : j->class_ms) x = 0.f;
    if (kernel != SRM_KERNEL_PLAIN)   // the class events are recorded only with band launches
        for (size_t c = 0; c < j->classes.size(); ++c)
        {
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, j->cs[c], j->ce[c]));
            j->class_ms[j->classes[c].nw] += ms;
        }
    return SRM_OK;
}

int srm_job_download(srm_job *j, srm_result *res, uint32_t *ops, size_t n_ops_words)
{
    srm_ctx *ctx = j->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    if (res && j->n)
        HIP_TRY(hipMemcpyAsync(res, j->d_res.p, (size_t)j->n * sizeof(srm_result),
                               hipMemcpyDeviceToHost, ctx->stream));
    if (ops && n_ops_words) {
        size_t nw = std::min(n_ops_words, j->ops_words);
        HIP_TRY(hipMemcpyAsync(ops, j->d_ops.p, nw * 4, hipMemcpyDeviceToHost, ctx->stream));
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SRM_OK;
}

}  // extern "C"

namespace {
// validates and uploads what the record kernels read besides the run's results:
// window positions, read ids and the name table (or prefix); sizes the line
// buffers and the scan's temporary storage
int rec_inputs(srm_job *j, const int64_t *pos, const int32_t *read_id, const char *names,
               int name_stride, size_t n_names, const char *name_prefix)
{
    if (!j->uploaded) return fail(SRM_ESTATE, "records before a successful upload");
    if (!pos) return fail(SRM_EINVAL, "records need the window positions");
    srm_ctx *ctx = j->ctx;
    const int n = j->n;
    if (names) {
        if (name_stride < 2 || name_stride > 256 || !read_id)
            return fail(SRM_EINVAL, "a name table needs 2 <= name_stride <= 256 and read ids");
        for (int t = 0; t < n; ++t)
            if (read_id[t] < 0 || (size_t)read_id[t] >= n_names)
                return fail(SRM_EINVAL, "read id %d of candidate %d is outside the name table",
                            read_id[t], t);
    } else if (read_id) {
        for (int t = 0; t < n; ++t)
            if (read_id[t] < 0) return fail(SRM_EINVAL, "negative read id at candidate %d", t);
    }
    const char *prefix = name_prefix ? name_prefix : "";
    const size_t plen = strlen(prefix);
    if (plen > 255) return fail(SRM_EINVAL, "name prefix longer than 255 bytes");
    HIP_TRY(hipSetDevice(ctx->device));
    int rc;
    if ((rc = upload(ctx, j->d_rec_pos, pos, (size_t)n))) return rc;
    if (read_id && (rc = upload(ctx, j->d_rec_id, read_id, (size_t)n))) return rc;
    if (names) {
        if ((rc = upload(ctx, j->d_rec_names, names, n_names * (size_t)name_stride))) return rc;
    } else if ((rc = upload(ctx, j->d_rec_names, prefix, plen + 1))) {
        return rc;
    }
    if ((rc = ctx->ensure(j->d_rec_len, ((size_t)n + 1) * 8))) return rc;
    if ((rc = ctx->ensure(j->d_rec_off, ((size_t)n + 1) * 8))) return rc;
    if ((rc = ctx->ensure(j->d_rec_info, std::max<size_t>((size_t)n * 16, 16)))) return rc;
    if ((rc = ctx->ensure(j->d_rec_hits, 256))) return rc;
    size_t tmp = 0;
    HIP_TRY(srm::rec_scan(nullptr, tmp, (const uint64_t *)j->d_rec_len.p,
                          (uint64_t *)j->d_rec_off.p, n, ctx->stream));
    if ((rc = ctx->ensure(j->d_rec_tmp, std::max<size_t>(tmp, 256)))) return rc;
    j->rec_tmp = tmp;
    j->rec_names = names != nullptr;
    j->rec_ids = read_id != nullptr;
    j->rec_stride = names ? name_stride : 0;
    j->rec_prefix_len = (int)plen;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return SRM_OK;
}
}  // namespace

extern "C" {

long long srm_job_records(srm_job *j, const int64_t *pos, const int32_t *read_id,
                          const char *names, int name_stride, size_t n_names,
                          const char *name_prefix, char *out, size_t outcap, long long *n_hits)
{
    // this call's inputs replace those of a record stage set up before
    if (j->rec_stage) {
        destroy_graph(j);
        j->rec_stage = false;
    }
    int rc = rec_inputs(j, pos, read_id, names, name_stride, n_names, name_prefix);
    if (rc) return rc;
    srm_ctx *ctx = j->ctx;
    const int n = j->n;
    srm::RecArgs a = rec_args(j, true);
    HIP_TRY(hipMemsetAsync(a.hits, 0, 8, ctx->stream));
    HIP_TRY(srm::launch_rec_measure(a, ctx->stream));
    size_t tmp = j->rec_tmp;
    HIP_TRY(srm::rec_scan(j->d_rec_tmp.p, tmp, a.len, (uint64_t *)j->d_rec_off.p, n, ctx->stream));
    uint64_t total = 0;
    uint32_t hits = 0;
    HIP_TRY(hipMemcpyAsync(&total, (const uint64_t *)j->d_rec_off.p + n, 8, hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipMemcpyAsync(&hits, a.hits, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (n_hits) *n_hits = hits;
    if (total > outcap || !out) return (long long)total;   // size query
    if ((rc = ctx->ensure(j->d_rec_out, std::max<size_t>(total, 256)))) return rc;
    a.out = (char *)j->d_rec_out.p;
    HIP_TRY(srm::launch_rec_write(a, ctx->stream));
    if (total)
        HIP_TRY(hipMemcpyAsync(out, a.out, total, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return (long long)total;
}

int srm_job_set_records(srm_job *j, const int64_t *pos, const int32_t *read_id, const char *names,
                        int name_stride, size_t n_names, const char *name_prefix)
{
    destroy_graph(j);   // the launch sequence changes
    j->rec_stage = false;
    j->rec_done = false;
    int rc = rec_inputs(j, pos, read_id, names, name_stride, n_names, name_prefix);
    if (rc) return rc;
    // the output holds the longest text the lines can have: the name, two int32
    // coordinates (11 characters each), the strand, both rows of at most w + k
    // characters, the int16 score and 7 separators per candidate
    const size_t name_max = names ? (size_t)name_stride - 1 : (size_t)j->rec_prefix_len + 11;
    j->rec_cap = j->rec_rows_bound + (size_t)j->n * (name_max + 11 + 11 + 2 + 6 + 7);
    if ((rc = j->ctx->ensure(j->d_rec_out, std::max<size_t>(j->rec_cap, 256)))) return rc;
    j->rec_stage = true;
    return SRM_OK;
}

long long srm_job_records_text(srm_job *j, char *out, size_t outcap, long long *n_hits)
{
    if (!j->rec_stage || !j->rec_done)
        return fail(SRM_ESTATE, "srm_job_records_text without a run of a record stage");
    srm_ctx *ctx = j->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    uint64_t total = 0;
    uint32_t hits = 0;
    HIP_TRY(hipMemcpyAsync(&total, (const uint64_t *)j->d_rec_off.p + j->n, 8,
                           hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(&hits, j->d_rec_hits.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (total > j->rec_cap) return fail(SRM_EDEVICE, "record text %llu exceeds its bound %zu",
                                        (unsigned long long)total, j->rec_cap);
    if (n_hits) *n_hits = hits;
    if (out && total <= outcap && total) {
        HIP_TRY(hipMemcpyAsync(out, j->d_rec_out.p, total, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    return (long long)total;
}

float srm_job_last_ms(srm_job *j) { return j->last_ms; }

float srm_job_kernel_ms(srm_job *j, int kind)
{
    return kind == 0 ? j->band_ms : kind == 1 ? j->plain_ms : j->rec_ms;
}

int srm_job_use_graph(srm_job *j, int enable)
{
    j->graph_enabled = enable != 0;
    if (!enable) destroy_graph(j);
    return SRM_OK;
}

int srm_job_path_count(srm_job *j, int path)
{
    if (path == 0) return (int)j->n_band;
    if (path == 1) return (int)j->n_rest;
    return (int)j->n_bytes;
}

int srm_calibrate_copy(int device, size_t bytes, float *ms)
{
    HIP_TRY(hipSetDevice(device));
    bytes &= ~(size_t)15;
    void *a = nullptr, *b = nullptr;
    HIP_TRY(hipMalloc(&a, bytes));
    hipError_t e = hipMalloc(&b, bytes);
    if (e != hipSuccess) {
        (void)hipFree(a);
        return fail(SRM_ENOMEM, "hipMalloc(%zu) failed", bytes);
    }
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipMemset(a, 1, bytes));
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipEventRecord(e0, nullptr));
    HIP_TRY(srm::launch_copy(a, b, bytes, nullptr));
    HIP_TRY(hipEventRecord(e1, nullptr));
    HIP_TRY(hipEventSynchronize(e1));
    HIP_TRY(hipEventElapsedTime(ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return SRM_OK;
}

long long srm_pack_codes(const char *seq, size_t n, uint32_t *words, size_t code_off)
{
    // code of every byte (A0 C1 G2 T3, anything else 0) and 1 for the bytes that
    // are not ACGT: table loads, no branch per base
    static constexpr auto code = [] {
        std::array<uint8_t, 256> a{};
        a['C'] = 1;
        a['G'] = 2;
        a['T'] = 3;
        return a;
    }();
    static constexpr auto other = [] {
        std::array<uint8_t, 256> a{};
        for (auto &x : a) x = 1;
        a['A'] = a['C'] = a['G'] = a['T'] = 0;
        return a;
    }();
    long long bad = 0;
    auto put = [&](size_t i) {   // one code into its word, keeping the word's other codes
        const uint8_t ch = (uint8_t)seq[i];
        const size_t p = code_off + i;
        const uint32_t sh = (uint32_t)(p & 15) * 2;
        words[p >> 4] = (words[p >> 4] & ~(3u << sh)) | ((uint32_t)code[ch] << sh);
        bad += other[ch];
    };
    size_t i = 0;
    for (; i < n && ((code_off + i) & 15); ++i) put(i);
    for (; i + 16 <= n; i += 16) {   // whole words
        uint32_t w = 0;
        for (int j = 0; j < 16; ++j) {
            const uint8_t ch = (uint8_t)seq[i + (size_t)j];
            w |= (uint32_t)code[ch] << (2 * j);
            bad += other[ch];
        }
        words[(code_off + i) >> 4] = w;
    }
    for (; i < n; ++i) put(i);
    return bad;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// drop-in entry points
// ---------------------------------------------------------------------------
namespace {
std::mutex g_dropin_mu;
srm_ctx *g_dropin_ctx[2] = {nullptr, nullptr};

inline bool is_acgt(char c) { return c == 'A' || c == 'C' || c == 'G' || c == 'T'; }

int dropin_device()
{
    int n = srm_device_count();
    if (n <= 0) return -1;
    return use_gpu_value < n ? (int)use_gpu_value : 0;
}

srm_ctx *dropin_ctx(int which)
{
    std::lock_guard<std::mutex> lk(g_dropin_mu);
    if (!g_dropin_ctx[which]) {
        int dev = dropin_device();
        if (dev < 0) return nullptr;
        srm_ctx *c = nullptr;
        if (srm_ctx_create(dev, 0, &c) != SRM_OK) return nullptr;
        g_dropin_ctx[which] = c;
    }
    return g_dropin_ctx[which];
}

// Name of a read id in read_names (util.h:55, 62-66), as the reference finds it per
// hit (HASH_FIND_INT, cuda-functions.cu:984-985): the key's bucket in the uthash
// table, then its chain.  The bucket index is uthash's HASH_FNV (the hash SARUMAN
// is built with, src/Makefile:94, 99): FNV-1 over the key's 4 bytes, each byte
// taken as a signed char, masked by the power-of-two bucket count.  A table
// without buckets, or one built with another hash function, misses; then the
// lookup walks the insertion-order list once and answers the rest of the call
// from that map.
class NameLookup {
  public:
    const char *find(int id)
    {
        if (!read_names) return "";
        if (!walked) {
            if (const read_name_hash *e = in_bucket(id)) return e->name;
            for (read_name_hash *e = read_names; e; e = (read_name_hash *)e->hh.next)
                map.emplace(e->read_number, e->name);
            walked = true;
        }
        auto it = map.find(id);
        return it != map.end() ? it->second : "";
    }

  private:
    bool walked = false;
    std::unordered_map<int, const char *> map;

    static const read_name_hash *in_bucket(int id)
    {
        const srm_ut_hash_table *tbl = (const srm_ut_hash_table *)read_names->hh.tbl;
        if (!tbl || !tbl->buckets || tbl->num_buckets == 0 ||
            (tbl->num_buckets & (tbl->num_buckets - 1)) != 0)
            return nullptr;
        unsigned char key[sizeof(int)];
        std::memcpy(key, &id, sizeof key);
        uint32_t hv = 2166136261u;
        for (unsigned char b : key) hv = (hv * 16777619u) ^ (uint32_t)(int32_t)(signed char)b;
        const srm_ut_hash_handle *hh = tbl->buckets[hv & (tbl->num_buckets - 1)].hh_head;
        for (; hh; hh = (const srm_ut_hash_handle *)hh->hh_next) {
            if (hh->keylen == sizeof(int) && std::memcmp(hh->key, key, sizeof(int)) == 0)
                return (const read_name_hash *)((const char *)hh - tbl->hho);
        }
        return nullptr;
    }
};
}  // namespace

extern "C" {

void cuda_allocate_memory(int do_reverse)
{
    // cuda-functions.cu:541-562: do_reverse == 4 -> reverse pool, else forward
    int which = do_reverse == 4 ? 1 : 0;
    srm_ctx *c = dropin_ctx(which);
    if (!c) {
        fprintf(stderr, "ERROR: could not create GPU context: %s\n", srm_last_error());
        exit(1);
    }
    // the anchor pointer the reference exposed; here a small allocation of the context
    char **anchor = which ? &main_alloc_ptr_rev : &main_alloc_ptr_for;
    if (!*anchor) {
        (void)hipSetDevice(c->device);
        if (hipMalloc((void **)anchor, 256) != hipSuccess) *anchor = nullptr;
    }
}

void cuda_free_memory(int do_reverse)
{
    // cuda-functions.cu:564-575: any non-zero do_reverse frees the reverse pool
    int which = do_reverse ? 1 : 0;
    std::lock_guard<std::mutex> lk(g_dropin_mu);
    char **anchor = which ? &main_alloc_ptr_rev : &main_alloc_ptr_for;
    if (g_dropin_ctx[which]) {
        (void)hipSetDevice(g_dropin_ctx[which]->device);
        if (*anchor) (void)hipFree(*anchor);
        srm_ctx_destroy(g_dropin_ctx[which]);
        g_dropin_ctx[which] = nullptr;
    }
    *anchor = nullptr;
}

unsigned long int get_cuda_infos(void)
{
    // cuda-functions.cu:1051-1120
    int count = srm_device_count();
    fprintf(stderr, "\n=== CUDA hardware ===\n");
    fprintf(stderr, "Detected %d GPU(s):\n\n", count);
    size_t best_free = 0, sel_free = 0;
    int best = 0;
    for (int i = 0; i < count; ++i) {
        size_t fr = 0, tot = 0;
        hipDeviceProp_t prop;
        (void)hipSetDevice(i);
        (void)hipGetDeviceProperties(&prop, i);
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            fprintf(stderr, "Could not get memory information from GPU %d.\nShutting down.\n", i);
            exit(-1);
        }
        if (fr > best_free) {
            best_free = fr;
            best = i;
        }
        if (i == (int)use_gpu_value) sel_free = fr;
        fprintf(stderr, "GPU %d: %s\t[%*lu/%*lu MB available]\n", i, prop.name, 4,
                (unsigned long)(fr / (1024 * 1024)), 4, (unsigned long)(tot / (1024 * 1024)));
    }
    if (count <= 0 || best_free < 128000000) {
        fprintf(stderr,
                "\n-> SARUMAN was not able to find a GPU with enough VRAM available! (available: %lu MB)\n",
                (unsigned long)(best_free / (1024 * 1024)));
        exit(1);
    }
    if ((int)use_gpu_value < count) {
        fprintf(stderr, "\n-> SARUMAN was forced to run on GPU %d\n", use_gpu_value);
        return (unsigned long)sel_free;   // the selected device's memory (Q22)
    }
    fprintf(stderr, "\n-> SARUMAN will run on GPU %d\n", best);
    use_gpu_value = (unsigned short)best;
    return (unsigned long)best_free;
}

void cuda_needleman_wunsch(int match, int mismatch, char *genome_sequences_on_host,
                           char *read_sequences_on_host, int *genome_offsets, int *read_offsets,
                           int *edit_offsets, int sample_size, int *read_id_array,
                           int *genome_pos_array, char *direction_array, int error,
                           size_t total_genome_array_size, size_t total_read_sequence_array_size,
                           size_t total_edit_matrix_size, int do_reverse)
{
    (void)edit_offsets;
    (void)total_edit_matrix_size;
    fprintf(stderr, "\t\t-> starting CUDA kernel with %d samples from %s thread\n", sample_size,
            do_reverse ? "reverse" : "forward");
    if (do_reverse)
        reverse_alignment_candidates += sample_size;   // cuda-functions.cu:604-608
    else
        forward_alignment_candidates += sample_size;
    if (sample_size <= 0) return;
    const int which = do_reverse ? 1 : 0;
    srm_ctx *ctx = dropin_ctx(which);
    if (!ctx) {
        fprintf(stderr, "ERROR: no GPU context: %s\n", srm_last_error());
        return;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    const int n = sample_size, k = error;
    if (k < 0 || k > 255) {
        fprintf(stderr, "ERROR: error bound %d outside 0..255\n", k);
        return;
    }
    // the genome array holds every window plus k slot bytes per candidate
    // (hash.c:1095, cuda-functions.cu:734)
    if (total_genome_array_size < (size_t)sample_size * (size_t)error) {
        fprintf(stderr, "ERROR: genome array of %zu bytes is smaller than its %d slots of %d\n",
                total_genome_array_size, sample_size, error);
        return;
    }
    const size_t win_total = total_genome_array_size - (size_t)sample_size * (size_t)error;
    const size_t read_total = total_read_sequence_array_size;

    // candidates in the A.1 layout (hash.c:1055-1059): exclusive end offsets
    std::vector<srm_cand> cand((size_t)n);
    std::vector<int> gstart((size_t)n), rstart((size_t)n);
    bool any_bytes = false;
    uint32_t ops_acc = 0;
    const char *qenv = getenv("SARUMAN_QUIRK_FIRST_ROW0");
    const bool quirk_first = qenv && qenv[0] == '1';
    for (int t = 0; t < n; ++t) {
        int gs = t ? genome_offsets[t - 1] : 0, rs = t ? read_offsets[t - 1] : 0;
        int w = genome_offsets[t] - gs, L = read_offsets[t] - rs;
        if (w < 0 || L < 1 || (size_t)gs + (size_t)w > win_total || (size_t)rs + (size_t)L > read_total ||
            L > SRM_MAX_L || w > SRM_MAX_W) {
            fprintf(stderr, "ERROR: malformed candidate %d (L=%d w=%d)\n", t, L, w);
            return;
        }
        gstart[t] = gs;
        rstart[t] = rs;
        srm_cand &c = cand[t];
        c.L = (uint16_t)L;
        c.w = (uint16_t)w;
        c.flag = (uint8_t)direction_array[t];
        c.k = (uint8_t)k;
        c.bits = (quirk_first && t == 0) ? SRM_CAND_STOP_ROW0 : 0;
        c.pad = 0;
        c.ops_off = ops_acc;
        ops_acc += srm_ops_words((uint32_t)w, (uint32_t)k);
        bool ok = true;
        for (int i = 0; i < L && ok; ++i) ok = is_acgt(read_sequences_on_host[rs + i]);
        for (int i = 0; i < w && ok; ++i) ok = is_acgt(genome_sequences_on_host[gs + i]);
        // code index i of the kernel's view is packed byte i (the arrays below carry
        // SRM_CODE_PAD codes of padding in front)
        c.read_off = (uint32_t)rs;
        c.win_off = (uint32_t)gs;
        if (!ok) {
            c.bits |= SRM_CAND_BYTES;
            any_bytes = true;
        }
    }
    // 2-bit codes with SRM_CODE_PAD codes of padding at both ends
    std::vector<uint32_t> rw((read_total + 2 * SRM_CODE_PAD + 15) / 16 + 1, 0u);
    std::vector<uint32_t> ww((win_total + 2 * SRM_CODE_PAD + 15) / 16 + 1, 0u);
    srm_pack_codes(read_sequences_on_host, read_total, rw.data(), SRM_CODE_PAD);
    srm_pack_codes(genome_sequences_on_host, win_total, ww.data(), SRM_CODE_PAD);

    // the pool the reference sized per strand (saruman.c:82) bounds a batch; the
    // reference never compared the two (cuda-functions.cu:648-665, Q19)
    const char *kenv0 = getenv("SARUMAN_KERNEL");
    const int kernel0 = kenv0 && !strcmp(kenv0, "plain") ? SRM_KERNEL_PLAIN : SRM_KERNEL_AUTO;
    const size_t need = srm_job_bytes_bound(ctx, match, mismatch, cand.data(), n, rw.size(), ww.size(),
                                            any_bytes ? read_total : 0, any_bytes ? win_total : 0,
                                            kernel0, 0, 0, 0);
    if (cuda_memory_per_thread && need > cuda_memory_per_thread) {
        fprintf(stderr, "ERROR: batch of %d candidates needs %zu bytes, the pool holds %lu\n", n,
                need, cuda_memory_per_thread);
        return;
    }
    if (!ctx->dropin_job && srm_job_create(ctx, &ctx->dropin_job) != SRM_OK) {
        fprintf(stderr, "ERROR: %s\n", srm_last_error());
        return;
    }
    srm_job *job = ctx->dropin_job;
    const int kernel = kernel0;
    int rc = srm_job_upload(job, match, mismatch, cand.data(), n, rw.data(), rw.size(), ww.data(),
                            ww.size(), any_bytes ? read_sequences_on_host : nullptr,
                            any_bytes ? read_total : 0, any_bytes ? genome_sequences_on_host : nullptr,
                            any_bytes ? win_total : 0);
    if (rc == SRM_OK) rc = srm_job_run(job, kernel);
    std::vector<srm_result> res((size_t)n);
    std::vector<uint32_t> ops((size_t)ops_acc + 1);
    if (rc == SRM_OK) rc = srm_job_download(job, res.data(), ops.data(), ops.size());
    if (rc != SRM_OK) {
        fprintf(stderr, "ERROR: GPU alignment failed: %s\n", srm_last_error());
        return;
    }
    alignment_time[which] += srm_job_last_ms(job);

    // host post-processing and records, cuda-functions.cu:828-1006
    NameLookup names;
    std::string out;
    std::vector<char> line(4096);
    for (int t = 0; t < n; ++t) {
        if (res[t].score > k) continue;
        const srm_cand &c = cand[t];
        const char *nm = names.find(read_id_array[t]);
        int len = srm_format_record(nm, (unsigned)genome_pos_array[t], c.flag, k, c.L, c.w, &res[t],
                                    ops.data() + c.ops_off, read_sequences_on_host + rstart[t],
                                    genome_sequences_on_host + gstart[t], line.data(), (int)line.size());
        if (len >= (int)line.size()) {
            line.resize((size_t)len + 1);
            len = srm_format_record(nm, (unsigned)genome_pos_array[t], c.flag, k, c.L, c.w, &res[t],
                                    ops.data() + c.ops_off, read_sequences_on_host + rstart[t],
                                    genome_sequences_on_host + gstart[t], line.data(), (int)line.size());
        }
        if (len > 0) out.append(line.data(), (size_t)len);
        if (c.flag > 3) {                               // :997-1003
            if (reverse_output_control) reverse_output_control[read_id_array[t]] = 1;
            reverse_alignments_done++;
        } else {
            if (forward_output_control) forward_output_control[read_id_array[t]] = 1;
            forward_alignments_done++;
        }
    }
    if (!out.empty()) {
        fwrite(out.data(), 1, out.size(), stdout);
        fflush(stdout);
    }
}

}  // extern "C"
