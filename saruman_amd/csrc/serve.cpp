// serve.cpp -- continuous-batching request loop around the alignment path.
//
// A server keeps the packed genome resident on one GPU and runs requests (read
// batches of any size) through three overlapping stages:
//   filter   (host thread)   q-gram filter of both strands -> candidates, codes
//                            (on the host, or on the device with its own stream)
//   admit    (host thread)   a request enters as soon as one of n_slots slots is
//                            free and its device footprint fits the HBM budget
//                            (below: what the slots hold counts, not only what
//                            requests are charged)
//   slot k   (host thread)   upload, persistent banded kernel, download and record
//                            formatting on the slot's own HIP stream
// so the filtering of later requests, the device work of several requests and
// the formatting of finished ones run at the same time.  A request's records do
// not depend on what runs beside it: its candidates, kernels and output order are
// a function of the request alone.
#include <array>
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "saruman_hip.h"

namespace {

using Clock = std::chrono::steady_clock;

inline char comp(char c)
{
    // A<->T, C<->G, every other byte unchanged: one table load, no branch per base
    static constexpr auto t = [] {
        std::array<char, 256> a{};
        for (int i = 0; i < 256; ++i) a[(size_t)i] = (char)i;
        a['A'] = 'T';
        a['T'] = 'A';
        a['C'] = 'G';
        a['G'] = 'C';
        return a;
    }();
    return t[(uint8_t)c];
}

// host threads for a request's per-byte passes: one below 1 MB of work, else up
// to 8 (each slot and the filter stage run their own)
unsigned host_threads(size_t bytes)
{
    if (bytes < ((size_t)1 << 20)) return 1;
    return std::min(8u, std::max(1u, std::thread::hardware_concurrency()));
}

// fn(u, lo, hi) on T contiguous ranges of [0, n), range u on thread u
template <class Fn>
void par_ranges(unsigned T, size_t n, Fn fn)
{
    if (T <= 1 || n < T) {
        fn(0u, (size_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(T - 1);
    for (unsigned u = 0; u + 1 < T; ++u) {
        try {
            th.emplace_back(fn, u, n * u / T, n * (u + 1) / T);
        } catch (...) {   // no thread to be had: the range on this one
            fn(u, n * u / T, n * (u + 1) / T);
        }
    }
    fn(T - 1, n * (T - 1) / T, n);
    for (auto &x : th) x.join();
}

// srm_pack_codes over ranges of whole code words on several threads (a word is
// packed by one thread: the ranges start where code_off + start is a multiple of 16)
void pack_codes_par(const char *seq, size_t n, uint32_t *words, size_t code_off)
{
    const unsigned T = host_threads(n);
    if (T <= 1) {
        srm_pack_codes(seq, n, words, code_off);
        return;
    }
    auto cut = [&](size_t x) -> size_t {   // the first word boundary at or after x
        if (x == 0 || x >= n) return x >= n ? n : 0;
        const size_t p = code_off + x;
        return std::min(n, x + ((16 - (p & 15)) & 15));
    };
    par_ranges(T, n, [&](unsigned, size_t lo, size_t hi) {
        const size_t a = cut(lo), b = cut(hi);
        if (b > a) srm_pack_codes(seq + a, b - a, words, code_off + a);
    });
}

struct Request {
    uint64_t id = 0;
    size_t max_len = 0;   // longest read
    std::string reads;
    std::vector<uint64_t> off;
    std::vector<uint16_t> len;
    std::vector<uint8_t> k;
    // after the filter
    std::string reads_rev;
    std::vector<srm_cand> cand;
    std::vector<int64_t> pos;
    std::vector<int32_t> rid;
    std::vector<uint64_t> roff_bytes;
    std::vector<uint32_t> read_words;
    std::vector<int32_t> perf_read[2];
    std::vector<int64_t> perf_pos[2];
    size_t ops_total = 0;
    size_t dev_bytes = 0;
    // results
    std::string records;              // perfect hits, then (host formatting) aligned hits
    std::unique_ptr<char[]> rec_text; // aligned hits formatted on the device (after records)
    size_t rec_len = 0;
    srm_request_stats st{};
    bool done = false;
};

// a request's failure: its status and the library's message for it
void set_failed(srm_request_stats &st, int rc, const char *msg)
{
    st.status = rc;
    snprintf(st.message, sizeof st.message, "%s", msg ? msg : "");
}

}  // namespace

struct srm_server {
    srm_server_config cfg{};
    std::string genome;               // ASCII, for records
    srm_index *index = nullptr;
    srm_ctx *fctx = nullptr;          // the device filter's context (own stream)
    srm_dfilter *dfilter = nullptr;   // set when the filter runs on the device
    std::vector<srm_ctx *> ctx;       // one per slot; slot 0 owns the resident genome
    std::vector<srm_job *> job;
    std::vector<char *> stage;        // per slot: page-locked buffer for the record text
    std::vector<size_t> stage_cap;
    Clock::time_point t0;

    std::mutex mu;
    std::condition_variable cv;       // any state change
    std::deque<std::shared_ptr<Request>> incoming, ready;
    std::map<uint64_t, std::shared_ptr<Request>> all;
    std::vector<std::shared_ptr<Request>> slot_req;
    std::vector<bool> slot_busy;
    // HBM accounting.  A slot's context holds its job's buffers between requests
    // (they only grow), and slot 0 also holds the resident genome (base[0]).  An
    // idle slot counts with what it holds; a busy one with its reservation, the
    // bytes it held at admission plus the request's footprint, which its context's
    // pool cap enforces (an allocation past it fails).  Admission keeps the sum
    // within hbm_budget, releasing idle slots' buffers when that makes room.
    std::vector<size_t> base;         // bytes a slot holds with its job released
    std::vector<size_t> reserved;     // busy slots: the pool cap granted
    size_t retain = 0;                // an idle slot keeps at most base + this
    size_t peak_accounted = 0;        // largest sum seen at admission or completion
    uint64_t next_id = 1;
    bool stop = false;
    bool hold = false;                // admission paused (srm_server_hold)
    std::vector<std::thread> threads;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

    void filter_loop();
    void admit_loop();
    size_t accounted(int s, size_t s_bytes) const;   // under mu
    int pick_slot(size_t need);                       // under mu; trims idle slots
    size_t sum_base() const
    {
        size_t t = 0;
        for (size_t b : base) t += b;
        return t;
    }
    void slot_loop(int s);
    void prepare(Request &r);
    void run_on_slot(int s, Request &r);
};

void srm_server::prepare(Request &r)
{
    const size_t n = r.len.size();
    const size_t total = r.reads.size();
    r.reads_rev.resize(total);
    // reverse complements, reads split over the threads
    par_ranges(host_threads(total), n, [&](unsigned, size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i) {
            const char *src = r.reads.data() + r.off[i];
            char *dst = &r.reads_rev[r.off[i]];
            for (int j = 0; j < r.len[i]; ++j) dst[j] = comp(src[r.len[i] - 1 - j]);
        }
    });
    // both strands through the reference filter semantics (the device filter runs
    // both over one upload of the reads)
    srm_filter *both[2] = {nullptr, nullptr};
    if (dfilter) {
        int rc = srm_dfilter_run2(dfilter, nullptr, r.reads.data(), r.off.data(), r.len.data(),
                                  r.k.data(), n, cfg.chunk_value, &both[0], &both[1]);
        if (rc != SRM_OK) {
            set_failed(r.st, rc, srm_last_error());
            return;
        }
    }
    for (int s = 0; s < 2; ++s) {
        srm_filter *F = both[s];
        if (!F) {
            int rc = srm_filter_run(index, genome.data(), genome.size(), nullptr, r.reads.data(),
                                    r.off.data(), r.len.data(), r.k.data(), n, s ? 4 : 0,
                                    cfg.chunk_value, cfg.filter_threads, &F);
            if (rc != SRM_OK) {
                if (both[1]) srm_filter_destroy(both[1]);
                set_failed(r.st, rc, srm_last_error());
                return;
            }
        }
        uint64_t nc, nb, np;
        srm_filter_sizes(F, &nc, &nb, &np);
        std::vector<int32_t> cr(nc);
        std::vector<int64_t> cp(nc);
        std::vector<uint8_t> cf(nc);
        std::vector<uint16_t> cw(nc);
        r.perf_read[s].resize(np);
        r.perf_pos[s].resize(np);
        srm_filter_copy(F, cr.data(), cp.data(), cf.data(), cw.data(), nullptr,
                        r.perf_read[s].data(), r.perf_pos[s].data());
        srm_filter_destroy(F);
        for (size_t t = 0; t < nc; ++t) {
            const int32_t i = cr[t];
            srm_cand c;
            c.read_off = (uint32_t)(r.off[(size_t)i] + (cf[t] > 3 ? total : 0));
            c.win_off = (uint32_t)(cp[t] + SRM_CODE_PAD);
            c.ops_off = (uint32_t)r.ops_total;
            c.L = r.len[(size_t)i];
            c.w = cw[t];
            c.flag = cf[t];
            c.k = r.k[(size_t)i];
            c.bits = 0;
            c.pad = 0;
            r.ops_total += srm_ops_words(c.w, c.k);
            r.cand.push_back(c);
            r.pos.push_back(cp[t]);
            r.rid.push_back(i);
            r.roff_bytes.push_back(r.off[(size_t)i]);
        }
    }
    // read codes: forward reads, then their reverse complements
    r.read_words.assign((2 * total + 2 * SRM_CODE_PAD + 15) / 16 + 1, 0u);
    pack_codes_par(r.reads.data(), total, r.read_words.data(), SRM_CODE_PAD);
    pack_codes_par(r.reads_rev.data(), total, r.read_words.data(), SRM_CODE_PAD + total);
    size_t maxL = 0;
    for (auto L : r.len) maxL = std::max<size_t>(maxL, L);
    r.max_len = maxL;
    const size_t nc = r.cand.size();
    // device footprint: the library's bound for this job, record stage included
    // when the records are formatted on the device (names "r" + decimal id)
    r.dev_bytes = srm_job_bytes_bound(nullptr, 0, 1, r.cand.data(), (int)nc, r.read_words.size(), 0,
                                      0, 0, cfg.kernel, cfg.gpu_records, 2, 1 + 11);
    r.st.n_reads = (long long)n;
    r.st.n_candidates = (long long)nc;
    r.st.dev_bytes = (long long)r.dev_bytes;
}

void srm_server::run_on_slot(int s, Request &r)
{
    srm_job *j = job[(size_t)s];
    const int n = (int)r.cand.size();
    int rc = SRM_OK;
    std::vector<srm_result> res((size_t)n);
    // transcripts come to the host only when the host formats the records
    std::vector<uint32_t> ops(cfg.gpu_records ? 0 : r.ops_total + 1);
    if (n > 0) {
        rc = srm_job_upload(j, 0, 1, r.cand.data(), n, r.read_words.data(), r.read_words.size(),
                            nullptr, 0, nullptr, 0, nullptr, 0);
        if (rc == SRM_OK) rc = srm_job_run(j, cfg.kernel);
        if (rc == SRM_OK) r.st.device_ms = srm_job_last_ms(j);
        if (rc == SRM_OK)   // the transcripts stay on the device when it formats the records
            rc = srm_job_download(j, res.data(), cfg.gpu_records ? nullptr : ops.data(),
                                  cfg.gpu_records ? 0 : ops.size());
    }
    r.st.t_aligned = now();
    if (rc != SRM_OK) {
        set_failed(r.st, rc, srm_last_error());
        return;
    }
    // records: perfect hits (hash.c:580) of each strand, then aligned hits in
    // candidate order
    std::string out;
    std::vector<char> mapped(r.len.size(), 0);
    {
        // perfect hits: "r<id>\t<start>\t<stop>\t<strand>\t<genome>\t<read>\t0\n", as
        // the reference prints them (%d fields), written directly; the hits of both
        // strands in order, split over threads that each write their own part
        const size_t n0 = r.perf_read[0].size(), np = n0 + r.perf_read[1].size();
        size_t bytes = 0;
        for (int sd = 0; sd < 2; ++sd)
            for (int32_t i : r.perf_read[sd]) bytes += 2 * (size_t)r.len[(size_t)i] + 48;
        const unsigned T = host_threads(bytes);
        std::vector<std::string> part(T);
        par_ranges(T, np, [&](unsigned u, size_t lo, size_t hi) {
            std::string &o = part[u];
            size_t b = 0;
            for (size_t x = lo; x < hi; ++x) {
                const int sd = x < n0 ? 0 : 1;
                b += 2 * (size_t)r.len[(size_t)r.perf_read[sd][sd ? x - n0 : x]] + 48;
            }
            o.reserve(b);
            auto put_int = [&](long long v) {
                char d[24];
                int q = 0;
                const bool neg = v < 0;
                unsigned long long w = neg ? 0ull - (unsigned long long)v : (unsigned long long)v;
                do {
                    d[q++] = (char)('0' + w % 10);
                    w /= 10;
                } while (w);
                if (neg) o.push_back('-');
                while (q) o.push_back(d[--q]);
            };
            for (size_t x = lo; x < hi; ++x) {
                const int sd = x < n0 ? 0 : 1;
                const size_t p = sd ? x - n0 : x;
                const int32_t i = r.perf_read[sd][p];
                const int64_t S = r.perf_pos[sd][p];
                const int L = r.len[(size_t)i];
                const char *rd = (sd ? r.reads_rev.data() : r.reads.data()) + r.off[(size_t)i];
                o.push_back('r');
                put_int(i);
                o.push_back('\t');
                put_int((int)S);
                o.push_back('\t');
                put_int((int)(S + L - 1));
                o.append(sd ? "\t<<\t" : "\t>>\t");
                o.append(genome.data() + S, strnlen(genome.data() + S, (size_t)L));   // %.*s
                o.push_back('\t');
                o.append(rd, strnlen(rd, (size_t)L));
                o.append("\t0\n");
            }
        });
        size_t tot = 0;
        for (auto &x : part) tot += x.size();
        out.reserve(tot);
        for (auto &x : part) out.append(x);
        for (int sd = 0; sd < 2; ++sd)
            for (int32_t i : r.perf_read[sd]) mapped[(size_t)i] = 1;
    }
    r.st.n_perfect = (long long)(r.perf_read[0].size() + r.perf_read[1].size());
    if (n > 0 && cfg.gpu_records) {
        // the text's exact size first, then the text into a buffer of that size
        // (no zero-filled allocation of an estimate)
        long long hits = 0;
        long long total = srm_job_records(j, r.pos.data(), r.rid.data(), nullptr, 0, 0, "r",
                                          nullptr, 0, &hits);
        if (total > 0) {
            // the text comes back into the slot's page-locked buffer (a fast copy),
            // then goes to the request's own buffer on several threads
            const size_t want = (size_t)total;
            if (stage_cap[(size_t)s] < want) {
                srm_host_free(stage[(size_t)s]);
                const size_t cap = want + want / 4 + 4096;
                stage[(size_t)s] = (char *)srm_host_alloc(cap);
                stage_cap[(size_t)s] = stage[(size_t)s] ? cap : 0;
            }
            char *dst = stage[(size_t)s];
            std::unique_ptr<char[]> direct;
            if (!dst) {   // no page-locked memory: straight into the request's buffer
                direct.reset(new char[want]);
                dst = direct.get();
            }
            total = srm_job_records(j, r.pos.data(), r.rid.data(), nullptr, 0, 0, "r", dst, want,
                                    &hits);
            if (total > 0 && direct) {
                r.rec_text = std::move(direct);
            } else if (total > 0) {
                r.rec_text.reset(new char[(size_t)total]);
                char *to = r.rec_text.get();
                par_ranges(host_threads((size_t)total), (size_t)total,
                           [&](unsigned, size_t lo, size_t hi) { std::memcpy(to + lo, dst + lo, hi - lo); });
            }
        }
        if (total < 0) {
            set_failed(r.st, (int)total, srm_last_error());
            return;
        }
        r.rec_len = (size_t)total;
        r.st.n_hits = hits;
        for (int t = 0; t < n; ++t)
            if (res[(size_t)t].score <= (int)r.cand[(size_t)t].k) mapped[(size_t)r.rid[(size_t)t]] = 1;
    } else if (n > 0) {
        long long hits = 0;
        long long need = srm_format_records(n, res.data(), ops.data(), r.cand.data(), genome.data(),
                                            r.pos.data(), r.reads.data(), r.reads_rev.data(),
                                            r.roff_bytes.data(), r.rid.data(), "r", nullptr, 0,
                                            &hits, cfg.format_threads);
        if (need < 0) {
            set_failed(r.st, SRM_EINVAL, "host record formatting failed");
            return;
        }
        size_t base = out.size();
        out.resize(base + (size_t)need);
        srm_format_records(n, res.data(), ops.data(), r.cand.data(), genome.data(), r.pos.data(),
                           r.reads.data(), r.reads_rev.data(), r.roff_bytes.data(), r.rid.data(),
                           "r", &out[base], (size_t)need, &hits, cfg.format_threads);
        r.st.n_hits = hits;
        for (int t = 0; t < n; ++t)
            if (res[(size_t)t].score <= (int)r.cand[(size_t)t].k) mapped[(size_t)r.rid[(size_t)t]] = 1;
    }
    long long m = 0;
    for (char x : mapped) m += x;
    r.st.n_mapped = m;
    r.records.swap(out);
}

void srm_server::filter_loop()
{
    for (;;) {
        std::shared_ptr<Request> r;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return stop || !incoming.empty(); });
            if (stop) return;
            r = incoming.front();
            incoming.pop_front();
        }
        prepare(*r);
        r->st.t_filtered = now();
        {
            std::lock_guard<std::mutex> lk(mu);
            if (r->st.status != SRM_OK) {
                r->done = true;
                r->st.t_done = now();
            } else if (sum_base() + r->dev_bytes > cfg.hbm_budget) {
                // cannot fit even with every slot released
                set_failed(r->st, SRM_ENOMEM, "request exceeds the HBM budget");
                r->done = true;
                r->st.t_done = now();
            } else {
                ready.push_back(r);
            }
        }
        cv.notify_all();
    }
}

// the HBM the slots account for, with slot s (if >= 0) counted as s_bytes: idle
// slots with what their context holds (nothing else touches an idle slot's
// context), busy slots with their reservation
size_t srm_server::accounted(int s, size_t s_bytes) const
{
    size_t t = 0;
    for (size_t i = 0; i < ctx.size(); ++i) {
        if ((int)i == s) t += s_bytes;
        else if (slot_busy[i]) t += reserved[i];
        else t += srm_ctx_pool_used(ctx[i]);
    }
    return t;
}

// an idle slot for a request of `need` bytes, or -1.  Prefers the slot holding the
// most (its buffers are likely reused); when nothing fits, releases idle slots'
// job buffers, largest first, if that makes room.
int srm_server::pick_slot(size_t need)
{
    std::vector<int> idle;
    for (size_t i = 0; i < ctx.size(); ++i)
        if (!slot_busy[i]) idle.push_back((int)i);
    if (idle.empty()) return -1;
    std::sort(idle.begin(), idle.end(), [&](int a, int b) {
        return srm_ctx_pool_used(ctx[(size_t)a]) > srm_ctx_pool_used(ctx[(size_t)b]);
    });
    for (int s : idle)
        if (accounted(s, srm_ctx_pool_used(ctx[(size_t)s]) + need) <= cfg.hbm_budget) return s;
    // with every idle slot released, does it fit on the first?
    const int s = idle[0];
    size_t floor = 0;
    for (size_t i = 0; i < ctx.size(); ++i) floor += slot_busy[i] ? reserved[i] : base[i];
    if (floor + need > cfg.hbm_budget) return -1;
    for (int i : idle) {
        if (srm_ctx_pool_used(ctx[(size_t)i]) > base[(size_t)i]) srm_job_release(job[(size_t)i]);
        if (accounted(s, srm_ctx_pool_used(ctx[(size_t)s]) + need) <= cfg.hbm_budget) return s;
    }
    return s;   // unreachable: all idle slots are at their base now
}

void srm_server::admit_loop()
{
    for (;;) {
        std::unique_lock<std::mutex> lk(mu);
        int slot = -1;
        cv.wait(lk, [&] {
            if (stop) return true;
            if (hold || ready.empty()) return false;
            slot = pick_slot(ready.front()->dev_bytes);
            return slot >= 0;
        });
        if (stop) return;
        auto r = ready.front();
        ready.pop_front();
        const size_t s = (size_t)slot;
        reserved[s] = srm_ctx_pool_used(ctx[s]) + r->dev_bytes;
        srm_ctx_set_pool_cap(ctx[s], reserved[s]);
        slot_busy[s] = true;
        peak_accounted = std::max(peak_accounted, accounted(-1, 0));
        slot_req[s] = r;
        r->st.slot = slot;
        r->st.t_admitted = now();
        lk.unlock();
        cv.notify_all();
    }
}

void srm_server::slot_loop(int s)
{
    for (;;) {
        std::shared_ptr<Request> r;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return stop || slot_req[(size_t)s] != nullptr; });
            if (stop) return;
            r = slot_req[(size_t)s];
        }
        run_on_slot(s, *r);
        // an idle slot keeps its buffers for the next request, up to `retain`
        if (srm_ctx_pool_used(ctx[(size_t)s]) > base[(size_t)s] + retain) srm_job_release(job[(size_t)s]);
        {
            std::lock_guard<std::mutex> lk(mu);
            r->st.t_done = now();
            r->done = true;
            slot_req[(size_t)s] = nullptr;
            slot_busy[(size_t)s] = false;
            reserved[(size_t)s] = 0;
            peak_accounted = std::max(peak_accounted, accounted(-1, 0));
            // inputs are no longer needed
            std::vector<srm_cand>().swap(r->cand);
            std::vector<uint32_t>().swap(r->read_words);
        }
        cv.notify_all();
    }
}

extern "C" {

int srm_server_create(const srm_server_config *cfg, const char *genome, size_t glen,
                      const int *qs, int nq, srm_server **out)
{
    *out = nullptr;
    if (cfg->n_slots < 1 || cfg->n_slots > 16 || cfg->chunk_value < 1) return SRM_EINVAL;
    auto *S = new srm_server();
    S->cfg = *cfg;
    if (S->cfg.hbm_budget == 0) S->cfg.hbm_budget = (size_t)32 << 30;
    if (S->cfg.kernel < 0) S->cfg.kernel = SRM_KERNEL_AUTO;
    S->t0 = Clock::now();
    S->genome.assign(genome, glen);
    int rc = srm_index_build(genome, glen, qs, nq, cfg->filter_threads, &S->index);
    if (rc != SRM_OK) {
        delete S;
        return rc;
    }
    if (cfg->gpu_filter) {
        rc = srm_ctx_create(cfg->device, 0, &S->fctx);
        if (rc == SRM_OK) rc = srm_dfilter_create(S->fctx, S->index, genome, glen, &S->dfilter);
        if (rc != SRM_OK) {
            if (S->fctx) srm_ctx_destroy(S->fctx);
            srm_index_destroy(S->index);
            delete S;
            return rc;
        }
    }
    std::vector<uint32_t> gw((glen + 2 * SRM_CODE_PAD + 15) / 16 + 1, 0u);
    srm_pack_codes(genome, glen, gw.data(), SRM_CODE_PAD);
    for (int s = 0; s < cfg->n_slots; ++s) {
        srm_ctx *c = nullptr;
        rc = srm_ctx_create(cfg->device, 0, &c);
        if (rc == SRM_OK) rc = s == 0 ? srm_ctx_set_genome(c, gw.data(), gw.size())
                                      : srm_ctx_borrow_genome(c, S->ctx[0]);
        srm_job *j = nullptr;
        if (rc == SRM_OK) rc = srm_job_create(c, &j);
        if (rc != SRM_OK) {
            if (c) srm_ctx_destroy(c);
            for (size_t i = S->ctx.size(); i-- > 0;) {
                srm_job_destroy(S->job[i]);
                srm_ctx_destroy(S->ctx[i]);
            }
            srm_dfilter_destroy(S->dfilter);
            if (S->fctx) srm_ctx_destroy(S->fctx);
            srm_index_destroy(S->index);
            delete S;
            return rc;
        }
        S->ctx.push_back(c);
        S->job.push_back(j);
    }
    S->slot_req.assign((size_t)cfg->n_slots, nullptr);
    S->stage.assign((size_t)cfg->n_slots, nullptr);
    S->stage_cap.assign((size_t)cfg->n_slots, 0);
    S->slot_busy.assign((size_t)cfg->n_slots, false);
    S->reserved.assign((size_t)cfg->n_slots, 0);
    for (srm_ctx *c : S->ctx) S->base.push_back(srm_ctx_pool_used(c));
    S->retain = S->cfg.hbm_budget / (2 * (size_t)cfg->n_slots);
    if (S->sum_base() > S->cfg.hbm_budget) {
        srm_server_destroy(S);
        return SRM_ENOMEM;   // the resident genome alone exceeds the budget
    }
    S->threads.emplace_back([S] { S->filter_loop(); });
    S->threads.emplace_back([S] { S->admit_loop(); });
    for (int s = 0; s < cfg->n_slots; ++s) S->threads.emplace_back([S, s] { S->slot_loop(s); });
    *out = S;
    return SRM_OK;
}

long long srm_server_submit(srm_server *S, const char *reads, const uint64_t *read_off,
                            const uint16_t *read_len, const uint8_t *read_k, size_t n_reads)
{
    auto r = std::make_shared<Request>();
    uint64_t total = 0;
    for (size_t i = 0; i < n_reads; ++i) {
        if (read_len[i] < 1 || read_len[i] > SRM_MAX_L) return SRM_EINVAL;
        total = std::max<uint64_t>(total, read_off[i] + read_len[i]);
    }
    // repack the reads contiguously
    r->off.resize(n_reads);
    r->len.assign(read_len, read_len + n_reads);
    r->k.assign(read_k, read_k + n_reads);
    size_t acc = 0;
    for (size_t i = 0; i < n_reads; ++i) {
        r->off[i] = acc;
        acc += read_len[i];
    }
    r->reads.resize(acc);
    for (size_t i = 0; i < n_reads; ++i)
        std::memcpy(&r->reads[r->off[i]], reads + read_off[i], read_len[i]);
    std::lock_guard<std::mutex> lk(S->mu);
    r->id = S->next_id++;
    r->st.t_submit = S->now();
    r->st.slot = -1;
    S->all[r->id] = r;
    S->incoming.push_back(r);
    S->cv.notify_all();
    return (long long)r->id;
}

/* waits up to timeout_ms (< 0: forever); 1 = done, 0 = not yet, < 0 = unknown id */
int srm_server_wait(srm_server *S, long long id, int timeout_ms)
{
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->all.find((uint64_t)id);
    if (it == S->all.end()) return SRM_EINVAL;
    auto r = it->second;
    auto pred = [&] { return r->done; };
    if (timeout_ms < 0)
        S->cv.wait(lk, pred);
    else
        S->cv.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred);
    return r->done ? 1 : 0;
}

/* copies the request's records (if out is large enough) and stats; returns the
 * records' size in bytes.  release != 0 forgets the request afterwards. */
long long srm_server_result(srm_server *S, long long id, char *out, size_t outcap,
                            srm_request_stats *st, int release)
{
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find((uint64_t)id);
    if (it == S->all.end() || !it->second->done) return SRM_EINVAL;
    auto r = it->second;
    if (st) *st = r->st;
    long long need = (long long)(r->records.size() + r->rec_len);
    if (out && (size_t)need <= outcap) {
        std::memcpy(out, r->records.data(), r->records.size());
        if (r->rec_len) std::memcpy(out + r->records.size(), r->rec_text.get(), r->rec_len);
    }
    if (release && out && (size_t)need <= outcap) S->all.erase(it);
    return need;
}

/* HBM the alignment slots account for now (held by idle slots, reserved by busy
 * ones) and the most they have accounted for at once; both stay within hbm_budget */
void srm_server_memory(srm_server *S, size_t *now_bytes, size_t *peak_bytes)
{
    std::lock_guard<std::mutex> lk(S->mu);
    if (now_bytes) *now_bytes = S->accounted(-1, 0);
    if (peak_bytes) *peak_bytes = S->peak_accounted;
}

/* on != 0 pauses admission: filtered requests wait in the ready queue and
 * requests already on a slot finish; on == 0 resumes it. */
void srm_server_hold(srm_server *S, int on)
{
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->hold = on != 0;
    }
    S->cv.notify_all();
}

/* queue depths: submitted and not yet filtered, filtered and waiting for
 * admission, on a slot */
void srm_server_queue(srm_server *S, uint64_t *incoming, uint64_t *ready, uint64_t *running)
{
    std::lock_guard<std::mutex> lk(S->mu);
    if (incoming) *incoming = S->incoming.size();
    if (ready) *ready = S->ready.size();
    if (running) {
        uint64_t n = 0;
        for (auto &q : S->slot_req) n += q != nullptr;
        *running = n;
    }
}

void srm_server_destroy(srm_server *S)
{
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv.notify_all();
    for (auto &t : S->threads) t.join();
    for (size_t i = S->ctx.size(); i-- > 0;) {
        srm_job_destroy(S->job[i]);
        srm_ctx_destroy(S->ctx[i]);
    }
    for (char *b : S->stage) srm_host_free(b);
    srm_dfilter_destroy(S->dfilter);
    if (S->fctx) srm_ctx_destroy(S->fctx);
    srm_index_destroy(S->index);
    delete S;
}

}  // extern "C"
