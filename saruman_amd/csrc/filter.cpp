// filter.cpp -- host candidate generation with the semantics of SARUMAN's q-gram
// filter ("SEASHORE") and batch builder, redesigned for speed: a sorted 2-bit
// q-gram index instead of a string hash, reads processed in parallel, batch
// boundaries and the batch layout decided in one sequential pass afterwards.
//
// Reference semantics kept (files read as text):
//   parameters              src/saruman.c:42-75, :119
//   index of every offset   src/util.c:456-499, src/hash.c:180-240
//   position lists          src/hash.c:396-461 (0-terminated, one array per chunk)
//   run_seashore            src/hash.c:473-839
//   add_start_position      src/hash.c:40-178
//   cuda_queue_check        src/hash.c:841-851
//   prepare_cuda_alignments src/hash.c:898-1111
// including Q10 (s/e leak after a skip), Q11 (flag branches as written), Q12 (the
// first-pass scan starts at the index of P in list a and may run into the lists
// that follow, even those of the next read of the chunk), Q13, Q14, Q16.
// Generalisation: every read carries its own length and error bound (the
// mixed-length workloads); a read's q, q-gram offsets and threshold derive from
// them exactly as the reference derives the global ones.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "filter_internal.h"

using namespace srm_filter_impl;

namespace {

int nthreads(int want)
{
    int hw = (int)std::thread::hardware_concurrency();
    if (want > 0) return want;
    return std::max(1, std::min(16, hw));
}

template <class F>
void parallel_for(size_t n, int T, F f)
{
    if (T <= 1 || n < 2) {
        f(0, n, 0);
        return;
    }
    std::vector<std::thread> th;
    size_t per = (n + (size_t)T - 1) / (size_t)T;
    for (int t = 0; t < T; ++t) {
        size_t b = (size_t)t * per, e = std::min(n, b + per);
        if (b >= e) break;
        th.emplace_back(f, b, e, t);
    }
    for (auto &x : th) x.join();
}

struct Entry {
    int32_t pos;
    int8_t flag;
};

struct ReadOut {
    std::vector<Entry> ent;
    std::vector<int32_t> perfect;   // S of every perfect hit that was printed
    int queued = 0;                 // new entries with pos > 0 (hash.c:87-89, 168-170)
};

}  // namespace

namespace {

void build_one(QIndex &X, const char *g, size_t len, int q, int T)
{
    X.q = q;
    int pb = 1;
    while (((size_t)1 << pb) <= len) ++pb;
    X.pos_bits = pb;
    if (2 * q + pb > 64) X.pos_bits = 64 - 2 * q;   // guarded by the caller
    X.pre_bases = std::min(q, 11);
    const size_t npos = len >= (size_t)q ? len - (size_t)q + 1 : 0;
    std::vector<std::vector<uint64_t>> part((size_t)T);
    std::vector<std::vector<std::pair<std::string, uint32_t>>> oddp((size_t)T);
    parallel_for(npos, T, [&](size_t b, size_t e, int t) {
        auto &v = part[(size_t)t];
        v.reserve(e - b);
        // rolling code over the last q bases; `bad` counts non-ACGT bytes in it
        uint64_t code = 0;
        const uint64_t cm = (q == 32) ? ~0ULL : (((uint64_t)1 << (2 * q)) - 1);
        int bad_run = 0;   // bases since the last non-ACGT byte
        size_t start = b;
        for (size_t i = start; i < e + (size_t)q - 1 && i < len; ++i) {
            int c = code_of(g[i]);
            if (c < 0) {
                bad_run = 0;
                code = 0;
            } else {
                code = ((code << 2) | (uint64_t)c) & cm;
                ++bad_run;
            }
            if (i + 1 >= b + (size_t)q) {
                size_t p = i + 1 - (size_t)q;
                if (p < b || p >= e) continue;
                if (bad_run >= q)
                    v.push_back((code << X.pos_bits) | (uint64_t)p);
                else
                    oddp[(size_t)t].emplace_back(std::string(g + p, (size_t)q), (uint32_t)p);
            }
        }
        std::sort(v.begin(), v.end());
    });
    // merge the sorted parts
    while (part.size() > 1) {
        std::vector<std::vector<uint64_t>> nxt((part.size() + 1) / 2);
        parallel_for(nxt.size(), T, [&](size_t b, size_t e, int) {
            for (size_t i = b; i < e; ++i) {
                if (2 * i + 1 < part.size()) {
                    nxt[i].resize(part[2 * i].size() + part[2 * i + 1].size());
                    std::merge(part[2 * i].begin(), part[2 * i].end(), part[2 * i + 1].begin(),
                               part[2 * i + 1].end(), nxt[i].begin());
                } else {
                    nxt[i] = std::move(part[2 * i]);
                }
            }
        });
        part.swap(nxt);
    }
    X.keys = std::move(part[0]);
    const size_t np = (size_t)1 << (2 * X.pre_bases);
    X.pre.assign(np + 1, 0);
    const int shift = X.pos_bits + 2 * (q - X.pre_bases);
    for (uint64_t key : X.keys) X.pre[(key >> shift) + 1]++;
    for (size_t i = 0; i < np; ++i) X.pre[i + 1] += X.pre[i];
    for (auto &vv : oddp)   // ascending positions: parts are in position order
        for (auto &kv : vv) X.odd[kv.first].push_back(kv.second);
}

// the virtual data_array of one chunk of reads (hash.c:396-461), built lazily
struct Lists {
    const srm_index *ix;
    const std::vector<Params> *params;
    const std::vector<int> *pidx;
    const char *reads;
    const uint64_t *roff;
    const uint16_t *rlen;
    int do_reverse;
    size_t chunk_lo, chunk_hi;   // reads of this chunk
    // per-read cache: only the read in hand and the few it spills into
    struct R {
        size_t id = (size_t)-1;
        std::vector<uint32_t> flat;
        std::vector<uint32_t> start;
    };
    R cache[4];
    std::string oriented;

    const R &get(size_t id)
    {
        for (auto &c : cache)
            if (c.id == id) return c;
        R &c = cache[id & 3];
        c.id = id;
        c.flat.clear();
        c.start.clear();
        const Params &P = (*params)[(size_t)(*pidx)[id]];
        const char *rd = reads + roff[id];
        oriented.assign(rd, (size_t)rlen[id]);
        if (do_reverse == 4) {   // util.c:35-57
            for (auto &ch : oriented) ch = comp(ch);
            std::reverse(oriented.begin(), oriented.end());
        }
        const QIndex *X = ix->get(P.q);
        for (int x = 0; x < P.n; ++x) {
            const size_t from = c.flat.size();
            c.start.push_back((uint32_t)from);
            if (X) X->lookup(oriented.data() + P.gram[(size_t)x], c.flat);
            if (ix->base)   // a chunk's positions: base + offset (util.c:481)
                for (size_t u = from; u < c.flat.size(); ++u) c.flat[u] += ix->base;
            c.flat.push_back(0);
        }
        return c;
    }
    // data_array[offsets[read, x] + idx]
    uint32_t D(size_t id, int x, long idx)
    {
        const R *r = &get(id);
        long p = (long)r->start[(size_t)x] + idx;
        while (p >= (long)r->flat.size()) {
            p -= (long)r->flat.size();
            ++id;
            if (id >= chunk_hi) return 0;
            r = &get(id);
        }
        return r->flat[(size_t)p];
    }
};

// add_start_position, hash.c:40-178 (per-read state)
int add_start(ReadOut &ro, int32_t pos, int flag, int possible, int exact_tresh, int k)
{
    if (possible != exact_tresh) return 1;
    const int dir = flag > 3 ? 1 : 0;
    if (ro.ent.empty()) {
        ro.ent.push_back({pos, (int8_t)flag});
        if (pos > 0) ro.queued++;
        return 0;
    }
    for (auto &en : ro.ent) {
        const int tdir = en.flag > 3 ? 1 : 0;
        for (int i = -k; i <= k; ++i)
            if ((pos + i == en.pos || pos + i == -en.pos) && en.flag <= flag) return 1;
        if (en.pos == pos && tdir == dir && en.flag > flag) {
            en.flag = (int8_t)flag;
            return 2;
        } else if ((en.pos == pos && tdir == dir) || (en.pos == -pos && tdir == dir)) {
            return 1;
        }
    }
    ro.ent.push_back({pos, (int8_t)flag});
    if (pos > 0) ro.queued++;
    return 2;
}

// run_seashore for one read (hash.c:502-834).  genome/glen: the text the perfect
// check reads, whose offset 0 is genome position `base` (a -G chunk, hash.c:545-552).
// LS gives the chunk's data_array: reads, roff and D(read, list, index) as Lists has
// them.
template <class LS>
void seashore_read(LS &ls, size_t rid, const Params &P, const char *genome, size_t glen,
                   long base, int do_reverse, ReadOut &ro)
{
    const int q = P.q, half = P.half, k = P.k, L = P.L, rest = P.rest, dr = do_reverse;
    const char *rd = ls.reads + ls.roff[rid];
    std::string tmp_read(rd, (size_t)L);
    if (do_reverse == 4) {
        for (auto &ch : tmp_read) ch = comp(ch);
        std::reverse(tmp_read.begin(), tmp_read.end());
    }
    for (int a = 0; a < half; ++a) {
        const bool fqm = (a == 0);
        const long in_read = (long)a * q + 1;
        long off = 0;
        uint32_t Pa = ls.D(rid, a, off);
        while (Pa > 0) {
            const long gs = (long)Pa - in_read;
            if (gs + 1 - base > 0) {
                // perfect check (strncpy + strcmp, hash.c:550-583)
                const long S = gs + 1, Sl = S - base;
                bool perfect = (size_t)Sl + (size_t)L <= glen &&
                               std::memcmp(genome + Sl, tmp_read.data(), (size_t)L) == 0;
                if (perfect) {
                    // strcmp stops at a NUL in the genome copy: the genome has none
                    if (add_start(ro, (int32_t)(-S), 0, P.exact_tresh, P.exact_tresh, k) != 1)
                        ro.perfect.push_back((int32_t)S);
                }
                int possible = 1;
                // first pass, hash.c:606-692
                for (int b = a + 1; b < half; ++b) {
                    const long rstart = (long)b * q + 1;
                    for (int err = -k; err <= k; ++err) {
                        uint32_t Pb = ls.D(rid, b, off);
                        long nxt = 0;
                        while (Pb > 0) {
                            if ((long)Pb == gs + rstart + err) {
                                int flag;
                                long pos;
                                if (b == half - 1 && fqm) {
                                    flag = 1 + dr;
                                    pos = gs + 1;
                                } else if (a == half - 1) {
                                    if ((long)Pa + q > L) {
                                        flag = 3 + dr;
                                        pos = gs - k + 1;
                                    } else {
                                        flag = 2 + dr;
                                        pos = gs - k + 1;
                                    }
                                } else if (fqm) {
                                    flag = 1 + dr;
                                    pos = gs + 1;
                                } else {
                                    flag = 3 + dr;
                                    pos = gs - k + 1;
                                }
                                ++possible;
                                add_start(ro, (int32_t)pos, flag, possible, P.exact_tresh, k);
                            }
                            ++nxt;
                            Pb = ls.D(rid, b, nxt);
                        }
                    }
                }
                // second pass, hash.c:695-794
                for (int j = 0; j < half; ++j) {
                    const int lim = std::abs(half - 1 - j - a);
                    const long nin = (long)(half - j - 1) * q + 1 + rest;
                    for (int err = -lim; err <= lim; ++err) {
                        const int aerr = (j - a < 0) ? -err : err;
                        long so = 0;
                        uint32_t Pb = ls.D(rid, half + j, so);
                        while (Pb > 0) {
                            if ((long)Pb == gs + nin + aerr) {
                                int flag;
                                long pos;
                                if (j == 0 && fqm) {
                                    flag = 1 + dr;
                                    pos = gs + 1;
                                } else if (j == 0) {
                                    if (nin + q > L) {
                                        flag = 3 + dr;
                                        pos = gs - k + 1;
                                    } else {
                                        flag = 2 + dr;
                                        pos = gs - k + 1;
                                    }
                                } else if (fqm) {
                                    flag = 1 + dr;
                                    pos = gs + 1;
                                } else {
                                    flag = 3 + dr;
                                    pos = gs - k + 1;
                                }
                                ++possible;
                                add_start(ro, (int32_t)pos, flag, possible, P.exact_tresh, k);
                            }
                            ++so;
                            Pb = ls.D(rid, half + j, so);
                        }
                    }
                }
            }
            ++off;
            Pa = ls.D(rid, a, off);
        }
    }
}

}  // namespace

extern "C" {

int srm_filter_params(int L, int k, int *q, int *rest, int *n, int *exact_tresh)
{
    if (L < 1 || k < 0) return SRM_EINVAL;
    Params p = derive(L, k);
    *q = p.q;
    *rest = p.rest;
    *n = p.n;
    *exact_tresh = p.exact_tresh;
    return SRM_OK;
}

int srm_index_build(const char *genome, size_t len, const int *qs, int nq, int threads,
                    srm_index **out)
{
    return srm_index_build_at(genome, len, qs, nq, threads, 0, out);
}

int srm_index_build_at(const char *genome, size_t len, const int *qs, int nq, int threads,
                       uint32_t base, srm_index **out)
{
    *out = nullptr;
    int pb = 1;
    while (((size_t)1 << pb) <= len) ++pb;
    for (int i = 0; i < nq; ++i)
        if (qs[i] < 1 || qs[i] > 31 || 2 * qs[i] + pb > 64) return SRM_EINVAL;
    if (len >= ((size_t)1 << 31) || (uint64_t)base + len >= ((uint64_t)1 << 31)) return SRM_EINVAL;
    srm_index *X = new srm_index();
    X->glen = len;
    X->base = base;
    const int T = nthreads(threads);
    std::vector<int> uq(qs, qs + nq);
    std::sort(uq.begin(), uq.end());
    uq.erase(std::unique(uq.begin(), uq.end()), uq.end());
    X->idx.resize(uq.size());
    for (size_t i = 0; i < uq.size(); ++i) build_one(X->idx[i], genome, len, uq[i], T);
    *out = X;
    return SRM_OK;
}

void srm_index_destroy(srm_index *X) { delete X; }

/*
 * Candidate generation for one strand (do_reverse 0 = forward, 4 = reverse) over
 * n_reads reads (forward orientation as read from the input, read i at
 * reads + read_off[i], read_len[i] bytes, error bound read_k[i]).  chunk_value is
 * the reference's -s: reads per position-list chunk and the queue threshold that
 * flushes a batch.  blocked (may be NULL): per-genome-base mask of separator
 * bases; a window covering one is skipped like an out-of-genome window.
 */
namespace {

int filter_run(const srm_index *ix, const char *genome, size_t glen, uint64_t limit,
               const uint8_t *blocked, const char *reads, const uint64_t *read_off,
               const uint16_t *read_len, const uint8_t *read_k, size_t n_reads, int do_reverse,
               int chunk_value, int threads, srm_filter **out)
{
    *out = nullptr;
    if (chunk_value < 1 || (do_reverse != 0 && do_reverse != 4)) return SRM_EINVAL;
    // parameters per distinct (L, k)
    std::map<std::pair<int, int>, int> pmap;
    std::vector<Params> params;
    std::vector<int> pidx(n_reads);
    for (size_t i = 0; i < n_reads; ++i) {
        auto key = std::make_pair((int)read_len[i], (int)read_k[i]);
        auto it = pmap.find(key);
        if (it == pmap.end()) {
            if (key.first < 1) return SRM_EINVAL;
            params.push_back(derive(key.first, key.second));
            if (!ix->get(params.back().q)) return SRM_EINVAL;   // index lacks this q
            it = pmap.emplace(key, (int)params.size() - 1).first;
        }
        pidx[i] = it->second;
    }
    // per-read filtering in parallel
    std::vector<ReadOut> ro(n_reads);
    const int T = nthreads(threads);
    std::atomic<size_t> next{0};
    const size_t grain = 256;
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t) {
        th.emplace_back([&] {
            Lists ls;
            ls.ix = ix;
            ls.params = &params;
            ls.pidx = &pidx;
            ls.reads = reads;
            ls.roff = read_off;
            ls.rlen = read_len;
            ls.do_reverse = do_reverse;
            for (;;) {
                size_t b = next.fetch_add(grain);
                if (b >= n_reads) break;
                size_t e = std::min(n_reads, b + grain);
                for (size_t r = b; r < e; ++r) {
                    size_t run = r / (size_t)chunk_value;
                    ls.chunk_lo = run * (size_t)chunk_value;
                    ls.chunk_hi = std::min(n_reads, (run + 1) * (size_t)chunk_value);
                    seashore_read(ls, r, params[(size_t)pidx[r]], genome, glen, (long)ix->base,
                                  do_reverse, ro[r]);
                }
            }
        });
    }
    for (auto &x : th) x.join();

    struct Acc {
        const std::vector<ReadOut> &ro;
        size_t n_ent(size_t r) const { return ro[r].ent.size(); }
        int32_t pos_at(size_t r, size_t i) const { return ro[r].ent[i].pos; }
        int flag_at(size_t r, size_t i) const { return ro[r].ent[i].flag; }
        size_t n_perf(size_t r) const { return ro[r].perfect.size(); }
        int32_t perf_at(size_t r, size_t i) const { return ro[r].perfect[i]; }
        int queued(size_t r) const { return ro[r].queued; }
    };
    srm_filter *F = batch_pass(Acc{ro}, n_reads, read_len, read_k, limit, blocked, chunk_value);
    *out = F;
    return SRM_OK;
}

}  // namespace

int srm_filter_run(const srm_index *ix, const char *genome, size_t glen, const uint8_t *blocked,
                   const char *reads, const uint64_t *read_off, const uint16_t *read_len,
                   const uint8_t *read_k, size_t n_reads, int do_reverse, int chunk_value,
                   int threads, srm_filter **out)
{
    *out = nullptr;
    if (ix->base != 0) return SRM_EINVAL;   // a chunk's index: srm_filter_run_chunk
    return filter_run(ix, genome, glen, glen, blocked, reads, read_off, read_len, read_k, n_reads,
                      do_reverse, chunk_value, threads, out);
}

/*
 * One -G chunk (hash.c:545-552, 982, 998): the index holds genome positions
 * (its base plus the chunk offset); the perfect check reads chunk_text at
 * position - base; a window is queued only if it ends at or before limit, the
 * reference's total_bases_read (so never in the overhang of a full chunk).
 */
int srm_filter_run_chunk(const srm_index *ix, const char *chunk_text, size_t text_len,
                         uint64_t limit, const char *reads, const uint64_t *read_off,
                         const uint16_t *read_len, const uint8_t *read_k, size_t n_reads,
                         int do_reverse, int chunk_value, int threads, srm_filter **out)
{
    *out = nullptr;
    if (limit < ix->base) return SRM_EINVAL;
    return filter_run(ix, chunk_text, text_len, limit, nullptr, reads, read_off, read_len, read_k,
                      n_reads, do_reverse, chunk_value, threads, out);
}

int srm_filter_sizes(const srm_filter *F, uint64_t *n_cand, uint64_t *n_batches,
                     uint64_t *n_perfect)
{
    *n_cand = F->c_read.size();
    *n_batches = F->batch_end.size();
    *n_perfect = F->p_read.size();
    return SRM_OK;
}

int srm_filter_copy(const srm_filter *F, int32_t *c_read, int64_t *c_pos, uint8_t *c_flag,
                    uint16_t *c_w, uint64_t *batch_end, int32_t *p_read, int64_t *p_pos)
{
    const size_t n = F->c_read.size();
    if (c_read) std::memcpy(c_read, F->c_read.data(), n * 4);
    if (c_pos) std::memcpy(c_pos, F->c_pos.data(), n * 8);
    if (c_flag) std::memcpy(c_flag, F->c_flag.data(), n);
    if (c_w) std::memcpy(c_w, F->c_w.data(), n * 2);
    if (batch_end) std::memcpy(batch_end, F->batch_end.data(), F->batch_end.size() * 8);
    if (p_read) std::memcpy(p_read, F->p_read.data(), F->p_read.size() * 4);
    if (p_pos) std::memcpy(p_pos, F->p_pos.data(), F->p_pos.size() * 8);
    return SRM_OK;
}

void srm_filter_destroy(srm_filter *F) { delete F; }

}  // extern "C"
