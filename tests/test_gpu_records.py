"""Records formatted on the device (srm_job_records, saruman_amd/csrc/records_gpu.hip):
the text equals, byte for byte, the host formatter's records of the same run
(srm_format_record, the post-processing of src/cuda-functions.cu:828-994) in
candidate order, and the oracle's records."""
import math

import numpy as np
import pytest

import saruman_amd as sa
from saruman_amd.testing import random_batch

pytestmark = pytest.mark.gpu


def check_text(batch, got, want=None):
    host = "".join(r for r in got.records if r)
    assert got.extra["device_text"].decode() == host
    assert got.extra["device_hits"] == sum(1 for r in got.records if r)
    if want is not None:
        assert got.records == want.records


@pytest.mark.parametrize("shape", ["36bp_k2", "150bp_k8", "mixed_50_250"])
def test_device_records_match_host_and_oracle(orc, shape):
    print(f"[gpu] device records {shape}", flush=True)
    rng = np.random.default_rng(21)
    if shape == "36bp_k2":
        batch = random_batch(rng, 3000, L=36, k=2)
    elif shape == "150bp_k8":
        batch = random_batch(rng, 2000, L=150, k=8)
    else:
        batch = random_batch(rng, 3000, L=(50, 250), k=13,
                             per_cand_k=lambda L: math.ceil(0.05 * L))
    want = orc.align_batch(batch)
    for kernel in ("banded", "plain"):
        got = sa.align_batch(batch, kernel=kernel, device_records=True)
        check_text(batch, got, want)


def test_device_records_costs_bytes_and_negative_match(orc):
    """General costs, the raw-byte path (N in reads and windows) and a negative
    match, whose walks may stop at the slot cap (all on the plain kernel)."""
    print("[gpu] device records: costs, bytes, negative match", flush=True)
    rng = np.random.default_rng(22)
    batch = random_batch(rng, 1200, L=(30, 80), k=4, flags=(0, 1, 2, 3, 4, 5, 6, 7), match=1,
                         mismatch=3, per_cand_k=lambda L: 2 + L // 30)
    check_text(batch, sa.align_batch(batch, device_records=True), orc.align_batch(batch))
    batch = random_batch(rng, 800, L=40, k=3)
    for t in range(0, batch.n, 3):
        wl = bytearray(batch.windows[t])
        wl[int(rng.integers(0, len(wl)))] = ord("N")
        batch.windows[t] = bytes(wl)
        if t % 2 == 0:
            rl = bytearray(batch.reads[t])
            rl[int(rng.integers(0, len(rl)))] = ord("N")
            batch.reads[t] = bytes(rl)
    got = sa.align_batch(batch, device_records=True)
    assert got.paths[2] > 0
    check_text(batch, got, orc.align_batch(batch))
    batch = random_batch(rng, 500, L=30, k=3, match=-1, mismatch=1)
    check_text(batch, sa.align_batch(batch, device_records=True), orc.align_batch(batch))


def test_device_records_prefix_names_and_no_hits():
    print("[gpu] device records: prefix names, no hits", flush=True)
    from saruman_amd.batch import Job, _layout, context, format_records, pack_codes
    rng = np.random.default_rng(23)
    batch = random_batch(rng, 700, L=(36, 120), k=6, per_cand_k=lambda L: math.ceil(0.05 * L))
    batch.names = [f"r{t}" for t in range(batch.n)]
    lay = _layout(batch)
    job = Job(context(0))
    try:
        job.upload(0, 1, lay["cand"], pack_codes(lay["rcat"]), pack_codes(lay["wcat"]),
                   ops_total=lay["ops_total"])
        job.run("banded")
        res, ops = job.download()
        text, hits = job.records(batch.pos, prefix="r")          # ids = candidate index
        host = "".join(r for r in format_records(batch, res, ops, lay) if r)
        assert text.decode() == host and hits > 0
        # a batch without hits: every read is random against its window
        none = random_batch(rng, 200, L=60, k=1, p_random=1.0, edge_cases=False)
        lay2 = _layout(none)
        job.upload(0, 1, lay2["cand"], pack_codes(lay2["rcat"]), pack_codes(lay2["wcat"]),
                   ops_total=lay2["ops_total"])
        job.run("banded")
        res2, _ = job.download()
        text2, hits2 = job.records(none.pos, prefix="x")
        n_hits = int((res2["score"].astype(np.int64) <= 1).sum())
        assert hits2 == n_hits and (n_hits > 0 or text2 == b"")
    finally:
        job.close()


@pytest.mark.parametrize("graph", [False, True])
def test_record_stage_in_the_launch_sequence(orc, graph):
    """srm_job_set_records makes the record writer a stage of every run: the text
    of each run (eager, or replayed from a hipGraph several times with eager work
    and synchronisations between the replays) equals the standalone writer's and
    the oracle's records."""
    print(f"[gpu] record stage graph={graph}", flush=True)
    from saruman_amd.batch import Job, _layout, context, name_table, pack_codes
    rng = np.random.default_rng(24)
    batch = random_batch(rng, 2500, L=(50, 250), k=13, per_cand_k=lambda L: math.ceil(0.05 * L))
    lay = _layout(batch)
    want = orc.align_batch(batch)
    want_text = "".join(r for r in want.records if r).encode()
    names = name_table(batch.names)
    rid = np.arange(batch.n, dtype=np.int32)
    job = Job(context(0))
    other = Job(context(0))
    try:
        job.upload(0, 1, lay["cand"], pack_codes(lay["rcat"]), pack_codes(lay["wcat"]),
                   ops_total=lay["ops_total"])
        job.set_records(batch.pos, read_id=rid, names=names)
        job.use_graph(graph)
        for rep in range(3):
            job.run("banded")
            text, hits = job.records_text()
            assert text == want_text, rep
            assert hits == sum(1 for r in want.records if r)
            # eager work on the same device between replays
            small = random_batch(rng, 300, L=36, k=2)
            sl = _layout(small)
            other.upload(0, 1, sl["cand"], pack_codes(sl["rcat"]), pack_codes(sl["wcat"]),
                         ops_total=sl["ops_total"])
            other.run("auto")
            other.download()
        # the standalone writer on the same run gives the same text and removes the stage
        text2, _ = job.records(batch.pos, read_id=rid, names=names)
        assert text2 == want_text
    finally:
        job.close()
        other.close()


def test_device_records_short_and_long_lines_side_by_side(orc):
    """Lines of 36-860 bp reads in one batch: the short ones are written in runs by
    the block kernel, the long ones (over 512 printed positions or 1 KiB of text)
    by the wave kernel, into neighbouring bytes of the same output; plus names
    from a table and a prefix, and raw-byte candidates in between."""
    print("[gpu] device records: short and long lines", flush=True)
    rng = np.random.default_rng(25)
    batch = random_batch(rng, 1500, L=(36, 860), k=14, flags=(1, 3, 5, 7),
                         per_cand_k=lambda L: min(14, math.ceil(0.03 * L)))
    for t in range(0, batch.n, 7):
        wl = bytearray(batch.windows[t])
        wl[int(rng.integers(0, len(wl)))] = ord("N")
        batch.windows[t] = bytes(wl)
    want = orc.align_batch(batch)
    got = sa.align_batch(batch, device_records=True)
    assert got.paths[2] > 0
    check_text(batch, got, want)
