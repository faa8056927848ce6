// Runs the unit-cost band kernels of saruman_amd/csrc/band_bits.hip on the CPU
// (through tests/emu/hip/hip_runtime.h: every lane of every wave in turn) over
// random candidates, and compares each score, transcript and end cell with the
// oracle (oracle/nw_oracle.c: oracle_fill, oracle_traceback).  TEST INFRASTRUCTURE
// ONLY.
//
// usage: band_bits_emu N [seed [plant]] -> one summary line per scratch layout (stored D
// bits, checkpoints), exit status 1 on a mismatch.
// "plant" reads the oracle's gap rows the wrong way round (top for left), a mistake
// planted in the checker, which the comparison must report.
#include <algorithm>
#include <stdio.h>
#include <stdlib.h>
#include <vector>

#include "band_bits.hip"

thread_local EmuIdx threadIdx, blockIdx, blockDim;
namespace srm {
uint32_t band_lds[256 * 64];    // one 256-thread block's LDS (the kernels' staging)
uint32_t trace_lds[256 * 64];
uint32_t walk_lds[256 * 64];
}

extern "C" int oracle_fill(int ma, int mm, int k, const char *win, int w, const char *read, int L,
                           int flag, int16_t *M);
extern "C" int oracle_traceback(int ma, int mm, int k, const char *win, int w, const char *read,
                                int L, int flag, int stop_at_row0, const int16_t *M, char *R,
                                char *G, int cap, int *end_row, int *end_col);

namespace {

uint64_t rng = 88172645463325252ull;
bool plant = false;
uint32_t rnd()
{
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return (uint32_t)rng;
}
int imn(int a, int b) { return a < b ? a : b; }
int imx(int a, int b) { return a > b ? a : b; }
int fs(int f, int k) { return (f <= 7 && (f & 2)) ? k : 0; }
int fe(int f, int k) { return (f <= 7 && (f & 1)) ? k : 0; }

struct Case {
    std::vector<uint8_t> r, g;
    int L, w, k, flag, stop;
};
struct Tally {
    long hits = 0, bad = 0, ops[3] = {0, 0, 0};
};

// the oracle's traceback as 2-bit ops in walk order: its rows come out reversed
// (first alignment column first); '_' in the read row is a top step, '_' in the
// window row a left step
int oracle_ops(const Case &c, const std::vector<char> &rch, const std::vector<char> &wch,
               const std::vector<int16_t> &M, std::vector<int> &ops, int &row, int &col)
{
    const int cap = c.w + c.k;
    std::vector<char> R(cap + 1, 0), G(cap + 1, 0);
    const int n = oracle_traceback(0, 1, c.k, wch.data(), c.w, rch.data(), c.L, c.flag, c.stop,
                                   M.data(), R.data(), G.data(), cap, &row, &col);
    for (int u = n - 1; u >= 0; --u) {
        int op = R[u] == '_' ? 1 : G[u] == '_' ? 2 : 0;
        if (plant && op) op = 3 - op;
        ops.push_back(op);
    }
    return n;
}

template <int NW>
void run_class(std::vector<Case> &cs, Tally &tally, bool ck)
{
    const int n = (int)cs.size();
    if (!n) return;
    const int PAD = SRM_CODE_PAD;
    // 2-bit code arrays with random codes in the padding
    std::vector<uint8_t> rc, wc;
    for (int i = 0; i < PAD; ++i) {
        rc.push_back(rnd() & 3);
        wc.push_back(rnd() & 3);
    }
    std::vector<srm_cand> cand(n);
    std::vector<uint32_t> ops_off(n + 1, 0);
    for (int t = 0; t < n; ++t) {
        srm_cand &cd = cand[t];
        cd.read_off = (uint32_t)(rc.size() - PAD);
        cd.win_off = (uint32_t)(wc.size() - PAD);
        rc.insert(rc.end(), cs[t].r.begin(), cs[t].r.end());
        wc.insert(wc.end(), cs[t].g.begin(), cs[t].g.end());
        cd.L = (uint16_t)cs[t].L;
        cd.w = (uint16_t)cs[t].w;
        cd.k = (uint8_t)cs[t].k;
        cd.flag = (uint8_t)cs[t].flag;
        cd.bits = cs[t].stop ? SRM_CAND_STOP_ROW0 : 0;
        cd.pad = 0;
        cd.ops_off = ops_off[t];
        ops_off[t + 1] = ops_off[t] + (cs[t].w + cs[t].k + 15) / 16 + 1;
    }
    for (int i = 0; i < PAD + 64; ++i) {
        rc.push_back(rnd() & 3);
        wc.push_back(rnd() & 3);
    }
    auto pack = [](const std::vector<uint8_t> &c) {
        std::vector<uint32_t> w(c.size() / 16 + 2, 0u);
        for (size_t i = 0; i < c.size(); ++i) w[i / 16] |= (uint32_t)c[i] << (2 * (i % 16));
        return w;
    };
    const std::vector<uint32_t> rw = pack(rc), ww = pack(wc);
    // tiles of 64 by L descending, as the runtime cuts them
    std::vector<uint32_t> perm(n);
    for (int t = 0; t < n; ++t) perm[t] = t;
    std::stable_sort(perm.begin(), perm.end(),
                     [&](uint32_t a, uint32_t b) { return cs[a].L > cs[b].L; });
    std::vector<uint4> tiles;
    std::vector<uint64_t> toff;
    size_t words = 0;
    int maxL = 0;
    for (int b = 0; b < n; b += 64) {
        const int cnt = imn(64, n - b);
        int mL = 0, mk = 0;
        for (int u = 0; u < cnt; ++u) {
            mL = imx(mL, cs[perm[b + u]].L);
            mk = imx(mk, cs[perm[b + u]].k);
        }
        maxL = imx(maxL, mL);
        tiles.push_back(uint4{(uint32_t)b, (uint32_t)cnt, (uint32_t)mL, (uint32_t)mk});
        toff.push_back(words);
        words += srm::bits_tile_words(NW, mL, ck);
    }
    std::vector<srm_cand> ctile(n);
    for (int p = 0; p < n; ++p) ctile[p] = cand[perm[p]];
    std::vector<uint32_t> scratch(words + 64, 0xDEADBEEFu);
    std::vector<int32_t> tscore(n, 0x5A5A5A5A);
    std::vector<srm_result> res(n);
    std::vector<uint32_t> ops(ops_off[n] + 16, 0);
    srm::BandArgs a{};
    a.ma = 0;
    a.mm = 1;
    a.cand = cand.data();
    a.perm = perm.data();
    a.cand_tile = ctile.data();
    a.tile_score = tscore.data();
    a.tiles = tiles.data();
    a.ntiles = (uint32_t)tiles.size();
    a.read_words = rw.data() + PAD / 16;
    a.win_words = ww.data() + PAD / 16;
    a.scratch = scratch.data();
    a.tile_scratch = toff.data();
    a.res = res.data();
    a.ops = ops.data();
    a.walk_ck = ck;
    const uint32_t blocks = (a.ntiles + 3) / 4;
    // forward pass: the LDS staging rule of launch_bits_band
    a.code_words = maxL <= 250 ? ((maxL + 15) >> 4) + 3 : 0;
    for (uint32_t blk = 0; blk < blocks; ++blk)
        for (uint32_t th = 0; th < 256; ++th) {
            threadIdx.x = th;
            blockIdx.x = blk;
            blockDim.x = 256;
            if (ck)
                srm::bits_band_kernel<NW, true>(a);
            else
                srm::bits_band_kernel<NW, false>(a);
        }
    // walk (codes from the tile's rows; transcript words staged as the launch does)
    a.code_words = 0;
    a.stage_words = srm::bits_trace_stage_words(maxL);
    for (uint32_t blk = 0; blk < blocks; ++blk)
        for (uint32_t th = 0; th < 256; ++th) {
            threadIdx.x = th;
            blockIdx.x = blk;
            blockDim.x = 256;
            if (ck)
                srm::bits_trace_kernel<NW, true>(a);
            else
                srm::bits_trace_kernel<NW, false>(a);
        }
    std::vector<int16_t> M;
    for (int t = 0; t < n; ++t) {
        const Case &c = cs[t];
        std::vector<char> rch(c.L + 1), wch(c.w + 1);
        for (int i = 0; i < c.L; ++i) rch[i] = "ACGT"[c.r[i]];
        for (int i = 0; i < c.w; ++i) wch[i] = "ACGT"[c.g[i]];
        M.assign((size_t)(c.w + 1) * (c.L + 1), 0);
        const int sc = oracle_fill(0, 1, c.k, wch.data(), c.w, rch.data(), c.L, c.flag, M.data());
        bool ok = res[t].score == imn(sc, c.k + 1);
        if (ok && sc <= c.k) {
            ++tally.hits;
            std::vector<int> o2;
            int row, col;
            const int nn = oracle_ops(c, rch, wch, M, o2, row, col);
            ok = res[t].n_ops == nn && res[t].end_row == row && res[t].end_col == col;
            for (int u = 0; ok && u < nn; ++u)
                ok = ((ops[cand[t].ops_off + u / 16] >> (2 * (u % 16))) & 3u) == (uint32_t)o2[u];
            for (int u = 0; u < nn; ++u) tally.ops[o2[u]]++;
        }
        if (!ok) {
            if (tally.bad < 5)
                fprintf(stderr, "mismatch NW%d%s L=%d w=%d k=%d flag=%d stop=%d: score %d (oracle %d)\n",
                        NW, ck ? " (checkpoint layout)" : "", c.L, c.w, c.k, c.flag, c.stop,
                        res[t].score, sc);
            ++tally.bad;
        }
    }
}

} // namespace

int main(int argc, char **argv)
{
    const int N = argc > 1 ? atoi(argv[1]) : 20000;
    if (argc > 2) rng = strtoull(argv[2], nullptr, 0) | 1;
    plant = argc > 3 && argv[3][0] == 'p';
    // by class, reads up to 250 and longer ones apart (LDS and global code paths)
    std::vector<Case> by[9][2];
    for (int t = 0; t < N; ++t) {
        Case c;
        const int k = rnd() % 15, flag = rnd() % 9;
        const int L = 1 + rnd() % (t % 10 == 0 ? 900 : 160);
        const int s = fs(flag, k);
        const int w = imx(0, L + (int)(rnd() % (2 * k + 1)) - k + (rnd() % 3 == 0 ? 2 * s : 0));
        const int nw = (k + imx(2 * s + k, w - L) + 1 + 7) / 8;
        if (nw > 8 || w < L - k) continue;
        c.g.resize(w);
        c.r.resize(L);
        for (int i = 0; i < w; ++i) c.g[i] = rnd() & 3;
        // random reads, substituted copies, copies with indels, shifted copies
        const int mode = rnd() % 4;
        for (int i = 0; i < L; ++i) {
            c.r[i] = (mode && i < w) ? c.g[i] : (rnd() & 3);
            if (mode && rnd() % 10 == 0) c.r[i] = rnd() & 3;
        }
        if (mode == 2) {
            int p = (int)(rnd() % (2 * s + 1)), i = 0;
            while (i < L) {
                const int r = rnd() % 24;
                if (r == 0) {
                    c.r[i++] = rnd() & 3;
                } else if (r == 1) {
                    ++p;
                } else {
                    c.r[i++] = p < w ? c.g[p] : (rnd() & 3);
                    ++p;
                }
            }
        }
        if (mode == 3) {
            const int off = w > L ? (int)(rnd() % (w - L + 1)) : 0;
            for (int i = 0; i < L; ++i)
                if (i + off < w) c.r[i] = c.g[i + off];
        }
        c.L = L;
        c.w = w;
        c.k = k;
        c.flag = flag;
        c.stop = rnd() % 5 == 0;
        by[nw][L > 250].push_back(c);
    }
    // both scratch layouts: stored D bits, and checkpoints the walk recomputes from
    Tally tally[2];
    for (int ck = 0; ck < 2; ++ck)
        for (int g = 0; g < 2; ++g) {
            run_class<1>(by[1][g], tally[ck], ck);
            run_class<2>(by[2][g], tally[ck], ck);
            run_class<3>(by[3][g], tally[ck], ck);
            run_class<4>(by[4][g], tally[ck], ck);
            run_class<5>(by[5][g], tally[ck], ck);
            run_class<6>(by[6][g], tally[ck], ck);
            run_class<7>(by[7][g], tally[ck], ck);
            run_class<8>(by[8][g], tally[ck], ck);
        }
    for (int ck = 0; ck < 2; ++ck)
        printf("%s: cases %d hits %ld bad %ld diag %ld top %ld left %ld\n",
               ck ? "checkpoint" : "stored", N, tally[ck].hits, tally[ck].bad, tally[ck].ops[0],
               tally[ck].ops[1], tally[ck].ops[2]);
    return tally[0].bad != 0 || tally[1].bad != 0;
}
