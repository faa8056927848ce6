// static_align.hip -- the banded alignment of a candidate batch without any
// allocation: srm_job_upload + srm_job_run(SRM_KERNEL_BANDED) + the results, for
// callers that cannot allocate device memory (code that a hipGraph capture
// surrounds, tight serving loops, harnesses that forbid allocation).
//
// Everything the band kernels need besides the caller's inputs and outputs -- the
// descriptors in candidate and tile order, the tile table, the per-tile offsets,
// the scores handed from the forward pass to the walk, the class counters and the
// band scratch of every tile -- lives in one file-scope device array, `scratch`,
// of SRM_STATIC_SCRATCH_BYTES.  Each GPU holds its own copy until the process
// ends; its address is taken once per device with hipGetSymbolAddress and used
// for every kernel argument and copy.  The host plan (the order by band class, L
// mod 8 and L, and the tiles) is built in fixed host arrays, so the call
// allocates nothing on the host either.  Calls are serialised by a mutex: a call
// returns when its results are in the caller's buffers, and the scratch then
// holds that call's contents.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <mutex>

#include "nw_device.h"
#include "saruman_hip.h"

namespace {

static __device__ __align__(256) unsigned char scratch[SRM_STATIC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
constexpr size_t MAX_TILES = SRM_STATIC_MAX_CAND / 64 + 8 * 8;
constexpr int MAX_L = SRM_MAX_L;

std::mutex g_mu;
unsigned char *g_scratch[MAX_DEVICES] = {};   // the array's address on each device

// the host plan, rebuilt by each call (under g_mu)
srm_cand h_cand_tile[SRM_STATIC_MAX_CAND];
uint32_t h_perm[SRM_STATIC_MAX_CAND];
uint32_t h_tmp[SRM_STATIC_MAX_CAND];
uint8_t h_cls[SRM_STATIC_MAX_CAND];
uint32_t h_tiles[4 * MAX_TILES];
uint64_t h_tile_off[MAX_TILES];
size_t h_lcount[MAX_L + 2];

struct Class {
    int nw = 0, maxL = 0;
    uint32_t tile_begin = 0, ntiles = 0;
};

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

extern "C" int srm_align_static(int device, void *stream, int match, int mismatch,
                                const srm_cand *cand, int n, const uint32_t *d_read_words,
                                size_t n_read_words, const uint32_t *d_win_words,
                                size_t n_win_words, srm_result *d_res, uint32_t *d_ops,
                                size_t n_ops_words)
{
    using srm::set_error;
    std::lock_guard<std::mutex> lk(g_mu);
    if (device < 0 || device >= MAX_DEVICES) return set_error(SRM_EINVAL, "device %d", device);
    if (n < 0 || n > SRM_STATIC_MAX_CAND)
        return set_error(SRM_EINVAL, "srm_align_static takes at most %d candidates, got %d",
                         SRM_STATIC_MAX_CAND, n);
    if (match < 0 || mismatch < 1)
        return set_error(SRM_EINVAL, "srm_align_static needs match >= 0 and mismatch >= 1 "
                                     "(the exact band)");
    if (n > 0 && (!cand || !d_read_words || !d_win_words || !d_res || !d_ops))
        return set_error(SRM_EINVAL, "srm_align_static: null argument");
    if (n == 0) return SRM_OK;

    // Everything below up to the size check runs on the host: a batch the call
    // refuses never reaches the device.
    // ---- validate: every candidate must fit the exact band and its arrays, as
    // srm_job_upload checks them (the code arrays carry SRM_CODE_PAD codes of
    // padding at both ends; offsets count from the first code after it)
    const uint64_t pad = SRM_CODE_PAD;
    const uint64_t read_codes = (uint64_t)n_read_words * 16, win_codes = (uint64_t)n_win_words * 16;
    int maxL = 0;
    for (int t = 0; t < n; ++t) {
        const srm_cand &c = cand[t];
        if (c.L < 1 || c.L > SRM_MAX_L || c.w > SRM_MAX_W)
            return set_error(SRM_EINVAL, "candidate %d: L=%d w=%d outside 1..%d / 0..%d", t, c.L,
                             c.w, SRM_MAX_L, SRM_MAX_W);
        if (c.bits & SRM_CAND_BYTES)
            return set_error(SRM_EINVAL, "candidate %d: raw bytes need the plain kernel", t);
        if ((uint64_t)c.read_off + c.L + 2 * pad > read_codes)
            return set_error(SRM_EINVAL, "candidate %d: read codes out of bounds", t);
        if ((uint64_t)c.win_off + c.w + 2 * pad > win_codes)
            return set_error(SRM_EINVAL, "candidate %d: window codes out of bounds", t);
        if ((uint64_t)c.ops_off + srm_ops_words(c.w, c.k) > n_ops_words)
            return set_error(SRM_EINVAL, "candidate %d: transcript slot out of bounds", t);
        if (c.k > SRM_BAND_KMAX || srm::band_cells(c) > SRM_BAND_MAX || (int)c.w < (int)c.L - (int)c.k)
            return set_error(SRM_EINVAL, "candidate %d needs the plain kernel (k=%d, band %d)", t,
                             c.k, srm::band_cells(c));
        h_cls[t] = (uint8_t)((srm::band_cells(c) + 7) / 8);
        maxL = std::max(maxL, (int)c.L);
    }

    // ---- order by (class, L mod 8, L descending): a counting pass by L, then a
    // stable one by (class, L mod 8), as srm_job_upload orders them
    std::fill(h_lcount, h_lcount + maxL + 2, (size_t)0);
    for (int t = 0; t < n; ++t) ++h_lcount[(size_t)(maxL - cand[t].L) + 1];
    for (int x = 1; x <= maxL + 1; ++x) h_lcount[x] += h_lcount[x - 1];
    for (int t = 0; t < n; ++t) h_tmp[h_lcount[maxL - cand[t].L]++] = (uint32_t)t;
    size_t bk[9 * 8 + 1] = {};
    auto key = [&](uint32_t t) { return (size_t)h_cls[t] * 8 + (cand[t].L & 7); };
    for (int i = 0; i < n; ++i) ++bk[key(h_tmp[i]) + 1];
    for (size_t x = 1; x < 9 * 8 + 1; ++x) bk[x] += bk[x - 1];
    for (int i = 0; i < n; ++i) h_perm[bk[key(h_tmp[i])]++] = h_tmp[i];

    // ---- tiles of up to 64 candidates of one class and one L mod 8; the band
    // scratch of every tile, one after another
    const bool unit = match == 0 && mismatch == 1;
    Class cls[9];
    int ncls = 0;
    size_t ntiles = 0, words = 0;
    for (int i = 0; i < n;) {
        Class &bc = cls[ncls++];
        bc.nw = h_cls[h_perm[i]];
        bc.tile_begin = (uint32_t)ntiles;
        int e2 = i;
        while (e2 < n && h_cls[h_perm[e2]] == bc.nw) ++e2;
        for (int g = i; g < e2;) {
            int ge = g;
            while (ge < e2 && (cand[h_perm[ge]].L & 7) == (cand[h_perm[g]].L & 7)) ++ge;
            for (int s = g; s < ge; s += 64) {
                const int cnt = std::min(64, ge - s);
                uint32_t kmax = 0, lmax = 0;
                for (int u = s; u < s + cnt; ++u) {
                    kmax = std::max<uint32_t>(kmax, cand[h_perm[u]].k);
                    lmax = std::max<uint32_t>(lmax, cand[h_perm[u]].L);
                }
                h_tiles[4 * ntiles + 0] = (uint32_t)s;
                h_tiles[4 * ntiles + 1] = (uint32_t)cnt;
                h_tiles[4 * ntiles + 2] = lmax;
                h_tiles[4 * ntiles + 3] = kmax;
                h_tile_off[ntiles] = words;
                words += unit ? srm::bits_tile_words(bc.nw, (int)lmax, false)
                              : srm::tile_scratch_words(bc.nw, (int)lmax);
                bc.maxL = std::max(bc.maxL, (int)lmax);
                ++ntiles;
            }
            g = ge;
        }
        bc.ntiles = (uint32_t)ntiles - bc.tile_begin;
        i = e2;
    }
    for (int p = 0; p < n; ++p) h_cand_tile[p] = cand[h_perm[p]];

    // ---- the scratch layout: counters, descriptors in candidate and tile order,
    // perm, scores, tiles, tile offsets, then the band scratch
    size_t at = 0;
    const size_t o_queue = at;
    at = align256(at + 64 * sizeof(uint32_t));
    const size_t o_cand = at;
    at = align256(at + (size_t)n * sizeof(srm_cand));
    const size_t o_cand_tile = at;
    at = align256(at + (size_t)n * sizeof(srm_cand));
    const size_t o_perm = at;
    at = align256(at + (size_t)n * 4);
    const size_t o_score = at;
    at = align256(at + (size_t)n * 4);
    const size_t o_tiles = at;
    at = align256(at + ntiles * 16);
    const size_t o_toff = at;
    at = align256(at + ntiles * 8);
    const size_t o_band = at;
    at += words * 4;
    if (at > SRM_STATIC_SCRATCH_BYTES)
        return set_error(SRM_ENOMEM, "srm_align_static: the batch needs %zu bytes of scratch, the "
                                     "static scratch holds %zu; split the batch",
                         at, (size_t)SRM_STATIC_SCRATCH_BYTES);

    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return set_error(SRM_EDEVICE, "hipSetDevice(%d): %s", device, hipGetErrorString(e));
    if (!g_scratch[device]) {   // once per device, with the device current
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(SRM_EDEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch[device] = (unsigned char *)p;
    }
    unsigned char *base = g_scratch[device];
    hipStream_t st = (hipStream_t)stream;

#define SA_TRY(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return set_error(SRM_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));     \
    } while (0)

    SA_TRY(hipMemcpyAsync(base + o_cand, cand, (size_t)n * sizeof(srm_cand), hipMemcpyHostToDevice, st));
    SA_TRY(hipMemcpyAsync(base + o_cand_tile, h_cand_tile, (size_t)n * sizeof(srm_cand),
                          hipMemcpyHostToDevice, st));
    SA_TRY(hipMemcpyAsync(base + o_perm, h_perm, (size_t)n * 4, hipMemcpyHostToDevice, st));
    SA_TRY(hipMemcpyAsync(base + o_tiles, h_tiles, ntiles * 16, hipMemcpyHostToDevice, st));
    SA_TRY(hipMemcpyAsync(base + o_toff, h_tile_off, ntiles * 8, hipMemcpyHostToDevice, st));
    // the class counters of the persistent (other-cost) kernel start at zero
    SA_TRY(srm::launch_reset((uint32_t *)(base + o_queue), ncls, st));

    int num_cus = 0;
    SA_TRY(hipDeviceGetAttribute(&num_cus, hipDeviceAttributeMultiprocessorCount, device));
    for (int c = 0; c < ncls; ++c) {
        const Class &bc = cls[c];
        srm::BandArgs a{};
        a.ma = match;
        a.mm = mismatch;
        a.cand = (const srm_cand *)(base + o_cand);
        a.perm = (const uint32_t *)(base + o_perm);
        a.cand_tile = (const srm_cand *)(base + o_cand_tile);
        a.tile_score = (int32_t *)(base + o_score);
        a.tiles = (const uint4 *)(base + o_tiles) + bc.tile_begin;
        a.ntiles = bc.ntiles;
        a.queue = (uint32_t *)(base + o_queue) + c;
        a.read_words = d_read_words + SRM_CODE_PAD / 16;
        a.win_words = d_win_words + SRM_CODE_PAD / 16;
        a.scratch = (uint32_t *)(base + o_band);
        a.tile_scratch = (const uint64_t *)(base + o_toff) + bc.tile_begin;
        a.res = d_res;
        a.ops = d_ops;
        int blocks = (int)((bc.ntiles + 3) / 4);   // bits_band_kernel: a wave per tile
        if (!unit)   // persistent grid: the blocks that stay resident, at most a wave per tile
            blocks = std::min(blocks, std::max(1, num_cus * srm::band_occupancy(bc.nw, false)));
        SA_TRY(srm::launch_band(a, bc.nw, unit, blocks, bc.maxL, st));
        SA_TRY(srm::launch_trace(a, bc.nw, unit, bc.maxL, st));
    }
    // the host arrays and the scratch are reused by the next call
    SA_TRY(hipStreamSynchronize(st));
#undef SA_TRY
    return SRM_OK;
}
