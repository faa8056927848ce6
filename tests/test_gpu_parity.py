"""GPU parity: both HIP kernels against the C oracle, record by record.

Every hit's TSV record (position, strand, error count, both alignment rows) must
equal the oracle's byte for byte; scores must agree (the banded kernel saturates
scores above k at k + 1).  Shapes follow the workloads: 36 bp / k 2, 100 bp / k 5,
150 bp / k 8, and mixed 50-250 bp with k = ceil(5% L).
"""
import math
import sys
import zlib

import numpy as np
import pytest

import saruman_amd as sa
from saruman_amd.testing import random_batch

pytestmark = pytest.mark.gpu


def banner(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def check(batch, got, want, kernel):
    ks = batch.kvec()
    sc_got = got.scores
    sc_want = np.asarray(want.scores, np.int64)
    if kernel == "plain":
        assert (sc_got == sc_want).all()
    else:
        assert (np.minimum(sc_got, ks + 1) == np.minimum(sc_want, ks + 1)).all()
    bad = [t for t in range(batch.n) if got.records[t] != want.records[t]]
    assert not bad, (len(bad), got.records[bad[0]], want.records[bad[0]])


SHAPES = {
    "36bp_k2": dict(L=36, k=2),
    "100bp_k5": dict(L=100, k=5),
    "150bp_k8": dict(L=150, k=8),
    "mixed_50_250": dict(L=(50, 250), k=13, per_cand_k=lambda L: math.ceil(0.05 * L)),
}


@pytest.mark.parametrize("shape", list(SHAPES))
@pytest.mark.parametrize("kernel", ["banded", "plain"])
def test_shapes(orc, shape, kernel):
    banner(f"[gpu] shapes {shape} {kernel}")
    rng = np.random.default_rng(zlib.crc32(f"{shape}/{kernel}".encode()))
    n = 3000 if kernel == "banded" else 1500
    batch = random_batch(rng, n, flags=(1, 3, 5, 7), **SHAPES[shape])
    got = sa.align_batch(batch, kernel=kernel)
    want = orc.align_batch(batch)
    if kernel == "banded":
        assert got.paths[0] == batch.n, got.paths
    check(batch, got, want, kernel)
    assert sum(1 for r in want.records if r) > n // 3


@pytest.mark.parametrize("ma,mm", [(0, 2), (1, 2), (2, 1), (0, 3)])
def test_costs(orc, ma, mm):
    banner(f"[gpu] costs {ma} {mm}")
    rng = np.random.default_rng(7 + ma * 10 + mm)
    batch = random_batch(rng, 1500, L=(30, 80), k=4, flags=(0, 1, 2, 3, 4, 5, 6, 7), match=ma,
                         mismatch=mm, per_cand_k=lambda L: 2 + L // 30)
    want = orc.align_batch(batch)
    for kernel in ("banded", "plain"):
        got = sa.align_batch(batch, kernel=kernel)
        check(batch, got, want, kernel)


def test_negative_match_goes_plain(orc):
    banner("[gpu] negative match")
    rng = np.random.default_rng(9)
    batch = random_batch(rng, 500, L=30, k=3, match=-1, mismatch=1)
    got = sa.align_batch(batch, kernel="banded")
    assert got.paths[0] == 0 and got.paths[1] == batch.n
    check(batch, got, orc.align_batch(batch), "plain")


def test_raw_bytes_path(orc):
    banner("[gpu] raw bytes")
    rng = np.random.default_rng(10)
    batch = random_batch(rng, 800, L=40, k=3)
    for t in range(0, batch.n, 3):   # N in window or read (raw byte compare, N == N)
        wl = bytearray(batch.windows[t])
        wl[int(rng.integers(0, len(wl)))] = ord("N")
        batch.windows[t] = bytes(wl)
        if t % 2 == 0:
            rl = bytearray(batch.reads[t])
            rl[int(rng.integers(0, len(rl)))] = ord("N")
            batch.reads[t] = bytes(rl)
    got = sa.align_batch(batch, kernel="banded")
    assert got.paths[2] > 0
    check(batch, got, orc.align_batch(batch), "banded")


@pytest.mark.parametrize("kernel", ["banded", "plain"])
def test_first_candidate_quirk(orc, kernel):
    banner(f"[gpu] q1 {kernel}")
    rng = np.random.default_rng(12)
    batch = random_batch(rng, 200, L=36, k=2)
    batch.windows[0] = b"ACGAA" + b"C" * 33
    batch.reads[0] = b"TACG" + b"C" * 32     # leading insertion on row 0
    batch.flags[0] = 1
    batch.first_row0 = True
    check(batch, sa.align_batch(batch, kernel=kernel), orc.align_batch(batch), kernel)


def test_edges(orc):
    banner("[gpu] edges")
    rng = np.random.default_rng(13)
    # 600 bp: the longest reads the walk stages in LDS; 700 and 880 bp: the walk
    # reads their codes from read_words
    for n, L, k in ((1, 36, 2), (63, 1, 0), (65, 5, 0), (130, 20, 14), (64, 300, 3),
                    (64, 600, 5), (70, 700, 5), (40, 880, 4)):
        batch = random_batch(rng, n, L=L, k=k, flags=(0, 1, 3, 5, 7))
        want = orc.align_batch(batch)
        for kernel in ("banded", "plain"):
            got = sa.align_batch(batch, kernel=kernel)
            if kernel == "banded" and L >= 300:
                assert got.paths[0] == batch.n, (L, got.paths)
            check(batch, got, want, kernel)


# (match, mismatch, k, window, read, flag, pos, record) worked by hand; the same
# cases pin the oracle in tests/test_oracle.py
HAND = [
    (0, 1, 2, b"ACGTAA", b"ACGT", 1, 1000, "r0\t1000\t1003\t>>\tACGT\tACGT\t0\n"),
    (0, 1, 2, b"ACGTAA", b"ACTT", 5, 7, "r1\t7\t10\t<<\tACTT\tACGT\t1\n"),
    (0, 1, 1, b"ACGTA", b"CGT", 3, 50, "r2\t51\t53\t>>\tCGT\tCGT\t0\n"),        # Q4
    (0, 1, 1, b"ACGAA", b"TACG", 1, 10, "r3\t10\t12\t>>\tTACG\t_ACG\t1\n"),     # row 0
    (0, 1, 2, b"CCCCCGATACAC", b"GATTACAC", 3, 200,
     "r4\t200\t206\t>>\tGATTACAC\tGA_TACAC\t2\n"),                          # Q5
]


@pytest.mark.parametrize("kernel", ["banded", "plain"])
def test_hand_worked_records(kernel):
    """Both kernels reproduce the hand-worked records (one launch, each case
    behind a filler candidate so none is the first of the launch)."""
    banner(f"[gpu] hand-worked records {kernel}")
    from saruman_amd.batch import Batch
    win, rd, fl, ps, nm, ks = [], [], [], [], [], []
    for i, (ma, mm, k, w, r, f, p, _) in enumerate(HAND):
        win += [b"ACGTACGTAC", w]
        rd += [b"ACGTACGT", r]
        fl += [1, f]
        ps += [1, p]
        nm += [f"f{i}", f"r{i}"]
        ks += [1, k]
    batch = Batch(0, 1, 2, win, rd, np.array(fl, np.uint8), np.array(ps, np.int64), nm,
                  ks=np.array(ks, np.int64))
    got = sa.align_batch(batch, kernel=kernel)
    assert [got.records[2 * i + 1] for i in range(len(HAND))] == [h[7] for h in HAND]


def test_graph_replay(orc):
    """Replays of the captured launch sequence, with eager work and syncs between
    them, each checked against the eager result and the oracle."""
    banner("[gpu] graph")
    from saruman_amd.batch import Job, _layout, context, format_records, pack_codes
    rng = np.random.default_rng(14)
    batch = random_batch(rng, 4000, L=(50, 250), k=13, per_cand_k=lambda L: math.ceil(0.05 * L))
    want = orc.align_batch(batch)
    eager = sa.align_batch(batch, kernel="banded")
    check(batch, eager, want, "banded")
    lay = _layout(batch)
    job = Job(context(0))
    try:
        job.upload(batch.match, batch.mismatch, lay["cand"], pack_codes(lay["rcat"]),
                   pack_codes(lay["wcat"]), ops_total=lay["ops_total"])
        job.use_graph(True)
        for rep in range(4):
            job.run("banded")
            res, ops = job.download()
            assert (res == eager.res).all(), rep
            assert format_records(batch, res, ops, lay) == want.records
            sa.align_batch(random_batch(rng, 100, L=36, k=2), kernel="banded")   # eager work
    finally:
        job.close()


def test_tuned_launch_shapes_keep_results(orc):
    """Every launch shape the autotuner may pick gives the same results: grids,
    merged classes, classes run concurrently on several streams, tiles of 32
    candidates, blocks of 1 or 2 waves, classes split by read length and the
    checkpoint scratch layout, eagerly and replayed from a hipGraph."""
    banner("[gpu] tuning shapes")
    from saruman_amd import tune
    from saruman_amd.batch import Context, Job, _layout, pack_codes
    rng = np.random.default_rng(15)
    batch = random_batch(rng, 3000, L=(36, 250), k=13, per_cand_k=lambda L: math.ceil(0.05 * L))
    lay = _layout(batch)
    ctx = Context(0)
    try:
        outs = []
        for table, graph in ((None, False),
                             ({"blocks_per_cu": [0] * 9, "merge_below": 0, "class_streams": 1},
                              False),
                             ({"blocks_per_cu": [0] + [1] * 8, "merge_below": 3,
                               "class_streams": 1}, False),
                             ({"blocks_per_cu": [0, 2, 1, 3, 1, 2, 1, 1, 1], "merge_below": 2},
                              False),
                             ({"blocks_per_cu": [0] * 9, "merge_below": 0, "class_streams": 4},
                              False),
                             ({"blocks_per_cu": [0] * 9, "merge_below": 2, "class_streams": 2},
                              True),
                             ({"blocks_per_cu": [0] * 9, "merge_below": 0, "class_streams": 8},
                              True),
                             ({"blocks_per_cu": [0] * 9, "class_streams": 4, "tile_cands": 32,
                               "block_waves": 2}, False),
                             ({"blocks_per_cu": [0] * 9, "class_streams": 2, "block_waves": 1,
                               "split_len": 120}, True),
                             ({"blocks_per_cu": [0] * 9, "merge_below": 2, "class_streams": 1,
                               "tile_cands": 32, "split_len": 100}, False),
                             ({"blocks_per_cu": [0] * 9, "walk_layout": 1}, False),
                             ({"blocks_per_cu": [0] * 9, "walk_layout": 1, "class_streams": 2,
                               "block_waves": 1, "merge_below": 2}, True),
                             ({"blocks_per_cu": [0] * 9, "walk_layout": 1, "tile_cands": 32,
                               "split_len": 120, "class_streams": 8}, False)):
            tune.apply(ctx, table)
            job = Job(ctx)
            job.upload(0, 1, lay["cand"], pack_codes(lay["rcat"]), pack_codes(lay["wcat"]),
                       ops_total=lay["ops_total"])
            job.use_graph(graph)
            for rep in range(3 if graph else 1):     # graph: replays with syncs between
                job.run("banded")
                outs.append(job.download())
                sa.align_batch(random_batch(rng, 50, L=36, k=2), kernel="banded")
            job.close()
        want = orc.align_batch(batch)
        assert (np.minimum(outs[0][0]["score"].astype(np.int64), batch.kvec() + 1) ==
                np.minimum(np.asarray(want.scores, np.int64), batch.kvec() + 1)).all()
        # transcript words each walk wrote (the rest of a slot is never written)
        off = lay["cand"]["ops_off"].astype(np.int64)
        used = np.zeros(len(outs[0][1]), bool)
        for t in range(batch.n):
            nw = (int(outs[0][0]["n_ops"][t]) + 15) // 16
            used[off[t]:off[t] + nw] = True
        for res, ops in outs[1:]:
            assert (res == outs[0][0]).all()
            assert (ops[used] == outs[0][1][used]).all()
    finally:
        ctx.close()


def test_tile_shapes_other_costs(orc):
    """Tiles of 32 candidates and classes split by read length on the nibble
    kernels (costs other than match 0 / mismatch 1): the default launch's results."""
    banner("[gpu] tuning shapes, other costs")
    from saruman_amd import tune
    from saruman_amd.batch import Context, Job, _layout, pack_codes
    rng = np.random.default_rng(16)
    batch = random_batch(rng, 2000, L=(30, 160), k=4, flags=(0, 1, 3, 5, 7), match=1,
                         mismatch=3, per_cand_k=lambda L: 2 + L // 30)
    lay = _layout(batch)
    ctx = Context(0)
    try:
        outs = []
        for table in (None, {"blocks_per_cu": [0] * 9, "tile_cands": 32, "split_len": 80,
                             "class_streams": 2}):
            tune.apply(ctx, table)
            job = Job(ctx)
            job.upload(batch.match, batch.mismatch, lay["cand"], pack_codes(lay["rcat"]),
                       pack_codes(lay["wcat"]), ops_total=lay["ops_total"])
            job.run("banded")
            outs.append(job.download())
            job.close()
        want = orc.align_batch(batch)
        k1 = batch.kvec() + 1
        assert (np.minimum(outs[0][0]["score"].astype(np.int64), k1) ==
                np.minimum(np.asarray(want.scores, np.int64), k1)).all()
        assert (outs[1][0] == outs[0][0]).all()
    finally:
        ctx.close()
