"""The benchmark configurations through the whole pipeline (host q-gram filter,
resident 2-bit genome, banded kernels with the default launch shapes):

- the 36 bp configuration at its full size (1 Mbp genome, 100,000 reads, k = 2):
  every candidate's score and every hit's record against the oracle;
- the 100 bp configuration on a 100,000-read slice (the oracle's share of the
  full size): the same checks;
- at full size, where the oracle would take minutes, a property that holds at
  any size: the banded kernels and the plain kernel (two independent
  implementations of the recurrence) give the same scores, end cells and
  transcripts -- the 100 bp configuration (1,000,000 reads), the mixed-length
  configuration (2,000,000 reads) and the first 1,000,000-read batch of the
  150 bp configuration.
"""
import numpy as np
import pytest

import saruman_amd as sa
from saruman_amd import workload as wl
from saruman_amd.batch import Batch

pytestmark = pytest.mark.gpu


def pipeline(name, n_reads):
    cfg, genome, reads = wl.make_workload(name, n_reads=n_reads)
    index = wl.Index(genome, wl.needed_qs(reads))
    strands = [wl.run_filter(index, genome, reads, dr, cfg["chunk"]) for dr in (0, 4)]
    index.close()
    db = wl.device_batch(reads, strands)
    ctx = sa.Context(0)
    ctx.set_genome(wl.genome_words(genome))
    job = sa.Job(ctx)
    job.upload(0, 1, db.cand, db.read_words, None, ops_total=db.ops_total)
    return genome, reads, db, job


def banded_text(job, db):
    job.run("banded")
    res, ops = job.download()
    text, hits = job.records(db.pos, read_id=np.ascontiguousarray(db.read_id, np.int32),
                             prefix="r")
    return res, ops, text, hits


def oracle_of(genome, reads, db):
    from oracle import oracle as orc
    orc.load()
    g = genome.seq
    fwd = reads.seq
    rev = wl.revcomp_all(reads).tobytes()
    wins, rds = [], []
    for t in range(db.n):
        p, w = int(db.pos[t]), int(db.w[t])
        o, L = int(reads.off[db.read_id[t]]), int(reads.length[db.read_id[t]])
        wins.append(g[p:p + w])
        rds.append((rev if db.flag[t] > 3 else fwd)[o:o + L])
    b = Batch(0, 1, int(db.k.max()), wins, rds, db.flag, db.pos,
              [f"r{int(x)}" for x in db.read_id], ks=db.k)
    return orc.align_batch(b)


@pytest.mark.parametrize("name,n_reads", [("small", None), ("single", 100_000)],
                         ids=["36bp_full_size", "100bp_100k_reads"])
def test_workload_against_oracle(name, n_reads):
    print(f"[gpu] workload {name} against the oracle", flush=True)
    genome, reads, db, job = pipeline(name, n_reads)
    try:
        assert job.paths()[0] >= 0.9 * db.n
        res, ops, text, hits = banded_text(job, db)
    finally:
        job.close()
    want = oracle_of(genome, reads, db)
    k1 = db.k.astype(np.int64) + 1
    got_s = np.minimum(res["score"].astype(np.int64), k1)
    want_s = np.minimum(want.scores.astype(np.int64), k1)
    assert (got_s == want_s).all(), int((got_s != want_s).sum())
    assert hits == int((want_s < k1).sum()) and hits > db.n // 3
    assert text.decode() == "".join(want.records)


@pytest.mark.parametrize("name,n_reads", [("single", None), ("mixed", None),
                                          ("scale", 1_000_000)],
                         ids=["100bp_1M_reads", "mixed_2M_reads", "150bp_1M_read_batch"])
def test_full_size_banded_equals_plain(name, n_reads):
    print(f"[gpu] workload {name} at full size: banded against plain", flush=True)
    genome, reads, db, job = pipeline(name, n_reads)
    try:
        assert job.paths()[0] >= 0.9 * db.n
        res, ops, text, hits = banded_text(job, db)
        job.run("plain")
        pres, pops = job.download()
    finally:
        job.close()
    k1 = db.k.astype(np.int64) + 1
    assert (np.minimum(pres["score"], k1) == np.minimum(res["score"], k1)).all()
    for f in ("n_ops", "end_row", "end_col"):
        assert (pres[f] == res[f]).all(), f
    assert (pops == ops).all()
    assert hits == int((np.minimum(res["score"], k1) < k1).sum()) and hits > db.n // 3


def test_scale_workload_ten_batches():
    """The 150 bp configuration at its full size: a 12 Mbp genome in two contigs
    (8 + 4 Mbp, 45% GC), 10,000,000 reads, k = 8, the filter on the device, the
    candidates cut into ten batches of 1,000,000 reads run over the resident
    genome.  Every batch: the banded kernels equal the plain kernel (scores, end
    cells, transcripts) and the record stage's text equals the standalone record
    writer's after the plain run.  The first 100,000 reads' candidates: every score
    and record against the oracle."""
    print("[gpu] scale workload: 10M reads in ten 1M-read batches", flush=True)
    cfg, genome, reads = wl.make_workload("scale")
    assert reads.n == 10_000_000 and len(genome.contigs) == 2
    index = wl.Index(genome, wl.needed_qs(reads))
    ctx = sa.Context(0)
    df = wl.DeviceFilter(ctx, index, genome)
    try:
        strands = df.run_both(reads, cfg["chunk"])
    finally:
        df.close()
        index.close()
    batches = wl.batches_by_reads(reads, strands, 1_000_000)
    assert len(batches) == 10
    ctx.set_genome(wl.genome_words(genome))
    job = sa.Job(ctx)
    total_hits = 0
    try:
        for b, (lo, hi, db) in enumerate(batches):
            print(f"[gpu] scale batch {b}: reads {lo}-{hi}, {db.n} candidates", flush=True)
            job.upload(0, 1, db.cand, db.read_words, None, ops_total=db.ops_total)
            assert job.paths()[0] == db.n
            rid = np.ascontiguousarray(db.read_id + lo, np.int32)
            job.set_records(db.pos, read_id=rid, prefix="r")
            job.use_graph(True)
            job.run("banded")
            res, ops = job.download()
            text, hits = job.records_text()
            job.use_graph(False)
            job.run("plain")
            pres, pops = job.download()
            ptext, _ = job.records_text()
            k1 = db.k.astype(np.int64) + 1
            assert (np.minimum(pres["score"], k1) == np.minimum(res["score"], k1)).all()
            for f in ("n_ops", "end_row", "end_col"):
                assert (pres[f] == res[f]).all(), (b, f)
            assert (pops == ops).all(), b
            assert ptext == text, b
            assert hits == int((np.minimum(res["score"], k1) < k1).sum()) and hits > db.n // 3
            total_hits += hits
            if b == 0:
                keep = db.read_id < 100_000
                sub = wl.subset_reads(reads, 0, 100_000)
                part = wl.DeviceBatch(db.cand[keep], db.read_words, db.ops_total,
                                      db.read_id[keep], db.pos[keep], db.flag[keep],
                                      db.k[keep], db.L[keep], db.w[keep])
                want = oracle_of(genome, sub, part)
                got_s = np.minimum(res["score"][keep].astype(np.int64), k1[keep])
                want_s = np.minimum(want.scores.astype(np.int64), k1[keep])
                assert (got_s == want_s).all()
                lines = text.decode().splitlines(keepends=True)
                hit = np.minimum(res["score"].astype(np.int64), k1) < k1
                by_cand = iter(lines)
                mine = [next(by_cand) if h else "" for h in hit]
                assert [mine[t] for t in np.flatnonzero(keep)] == want.records
    finally:
        job.close()
    assert total_hits > 5_000_000
