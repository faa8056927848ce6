"""The batch pipeline (saruman_amd/pipeline.py): candidate batches from page-locked
memory through two lanes (contexts sharing the resident genome), each batch
uploaded, run as one graph replay of the band phase and the record stage, and its
record text copied back.  Every batch's text equals the host formatter's lines
(srm_format_records) for a single job's results on the same batch, whichever lane
ran it and in whichever order the batches came: each lane reuses its job, its
graph buffers and its upload's host arrays from one batch to the next, with sizes
going up and down.
"""
import ctypes as C

import numpy as np
import pytest

import saruman_amd as sa
from saruman_amd import tune
from saruman_amd import workload as wl
from saruman_amd._lib import ptr
from saruman_amd.pipeline import BatchPipeline, PinnedBatch

pytestmark = pytest.mark.gpu


def host_text(lib, db, reads, lo, genome, res, ops):
    """srm_format_records over one batch's results: its lines in candidate order."""
    hi = lo + int(db.read_id.max()) + 1 if db.n else lo
    sub = wl.subset_reads(reads, lo, hi)
    fwd = np.frombuffer(sub.seq, np.uint8)
    rev = wl.revcomp_all(sub)
    roff = np.ascontiguousarray(sub.off[db.read_id], np.uint64)
    rid = np.ascontiguousarray(db.read_id + lo, np.int32)
    g = np.frombuffer(genome.seq, np.uint8)
    pos = np.ascontiguousarray(db.pos, np.int64)
    nh = C.c_longlong()
    args = (int(db.n), ptr(res), ptr(ops), ptr(db.cand), ptr(g), ptr(pos), ptr(fwd), ptr(rev),
            ptr(roff), ptr(rid), b"r")
    need = lib.srm_format_records(*args, None, 0, C.byref(nh), 4)
    out = np.empty(max(int(need), 1), np.uint8)
    got = lib.srm_format_records(*args, ptr(out), out.size, C.byref(nh), 4)
    assert got == need
    return out[:got].tobytes()


def test_batches_through_two_lanes():
    print("[gpu] batch pipeline: 3 batches, 6 passes over 2 lanes", flush=True)
    cfg, genome, reads = wl.make_workload("mixed", n_reads=60_000)
    index = wl.Index(genome, wl.needed_qs(reads))
    strands = [wl.run_filter(index, genome, reads, dr, cfg["chunk"]) for dr in (0, 4)]
    index.close()
    batches = wl.batches_by_reads(reads, strands, 24_000)   # 24k, 24k and 12k reads
    assert len(batches) == 3 and batches[2][2].n < batches[0][2].n
    lib = sa.load()
    ctx = sa.Context(0)
    ctx.set_genome(wl.genome_words(genome))
    job = sa.Job(ctx)
    want = []
    try:
        for lo, hi, db in batches:
            job.upload(0, 1, db.cand, db.read_words, None, ops_total=db.ops_total)
            job.run("banded")
            res, ops = job.download()
            want.append(host_text(lib, db, reads, lo, genome, res, ops))
    finally:
        job.close()
        ctx.close()
    pinned = [PinnedBatch.of(db, lo) for lo, hi, db in batches]
    seq = (0, 2, 1, 1, 2, 0)    # lane 0 takes 24k, 24k, 12k reads; lane 1 12k, 24k, 24k
    try:
        with BatchPipeline(wl.genome_words(genome), lanes=2, tuning=tune.load_table()) as pipe:
            wall, res, texts = pipe.run([pinned[i] for i in seq], keep_text=True)
    finally:
        for p in pinned:
            p.free()
    assert sorted({r.lane for r in res}) == [0, 1]
    for bi, src in enumerate(seq):
        assert texts[bi] == want[src], (bi, src)
        assert res[bi].text_bytes == len(want[src])
        assert res[bi].hits == want[src].count(b"\n")
