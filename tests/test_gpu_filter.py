"""The q-gram filter on the GPU (saruman_amd/csrc/filter_gpu.hip): the same
batches, candidates (read, position, flag, window length) in the same order and
the same perfect hits as the host filter (filter.cpp), and as the Python
transcription of the reference filter (oracle/filter_oracle.py, src/hash.c) on
small inputs."""
import numpy as np
import pytest

from oracle import filter_oracle as fo
from saruman_amd import workload as wl
from saruman_amd.batch import Context

pytestmark = pytest.mark.gpu

FIELDS = ("read", "pos", "flag", "w", "batch_end", "perfect_read", "perfect_pos")


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def same(h, d, tag):
    for f in FIELDS:
        a, b = getattr(h, f), getattr(d, f)
        assert a.shape == b.shape and np.array_equal(a, b), (tag, f, a.shape, b.shape)


def host_vs_device(ctx, genome, reads, chunks):
    idx = wl.Index(genome, wl.needed_qs(reads))
    df = wl.DeviceFilter(ctx, idx, genome)
    out = []
    try:
        for chunk in chunks:
            for dr in (0, 4):
                h = wl.run_filter(idx, genome, reads, dr, chunk)
                d = df.run(reads, dr, chunk)
                same(h, d, (chunk, dr))
                out.append((h.n, len(h.perfect_read), df.stats()))
    finally:
        df.close()
        idx.close()
    return out


def test_against_reference_transcription(ctx):
    print("[gpu] device filter vs transcription", flush=True)
    L, k, chunk = 36, 2, 60
    genome = wl.synth_genome(101, [30_000], 0.5)
    reads = wl.synth_reads(102, genome, 400, L, L, 0, max_err=k, p_sub=0.8, p_ins=0.1, k=k)
    q = wl.filter_params(L, k)[0]
    idx = wl.Index(genome, [q])
    df = wl.DeviceFilter(ctx, idx, genome)
    a = np.frombuffer(reads.seq, np.uint8)
    rl = [a[int(o):int(o) + int(n)].tobytes() for o, n in zip(reads.off, reads.length)]
    oidx = fo.build_index(genome.seq, q)
    total = 0
    for dr in (0, 4):
        want_batches, want_perfect = fo.run_strand(genome.seq, rl, [f"r{i}" for i in range(reads.n)],
                                                   L, k, dr, chunk, oidx)
        got = df.run(reads, dr, chunk)
        ends = [0] + got.batch_end.tolist()
        assert len(ends) - 1 == len(want_batches)
        for bi, wb in enumerate(want_batches):
            lo, hi = ends[bi], ends[bi + 1]
            g = list(zip(got.read[lo:hi].tolist(), got.pos[lo:hi].tolist(),
                         got.flag[lo:hi].tolist(), got.w[lo:hi].tolist()))
            assert g == [(c.read_id, c.pos, c.flag, c.w) for c in wb], (dr, bi)
            total += len(wb)
        gp = sorted(zip(got.perfect_read.tolist(), got.perfect_pos.tolist()))
        wp = sorted((int(x.split("\t")[0][1:]), int(x.split("\t")[1])) for x in want_perfect)
        assert gp == wp
    df.close()
    idx.close()
    assert total > 200


def test_repeats_overflow_and_spill(ctx):
    """Reads from 40-copy repeat families hold more result-hash entries and
    perfect hits than the LDS slots: they run again with HBM slots.  A tiny
    chunk makes list walks run into the next reads' lists and stop at chunk ends."""
    print("[gpu] device filter repeats", flush=True)
    genome = wl.synth_genome(202, [40_000], 0.5, (3, 300, 600, 40, 0.5, 0.001))
    reads = wl.synth_reads(203, genome, 3000, 40, 40, 0, max_err=3, p_sub=0.8, p_ins=0.1, k=3)
    res = host_vs_device(ctx, genome, reads, chunks=(7, 25, 100_000))
    assert sum(r[0] for r in res) > 1000
    assert max(r[2]["overflow_reads"] for r in res) > 0


def test_non_acgt_bytes(ctx):
    """N runs and other letters in the genome and the reads: q-grams holding them
    are looked up by their text (FNV-1a hash, then a byte compare)."""
    print("[gpu] device filter non-ACGT", flush=True)
    rng = np.random.default_rng(7)
    g = wl.synth_genome(303, [60_000], 0.5)
    arr = np.frombuffer(g.seq, np.uint8).copy()
    arr[rng.choice(len(arr), 120, replace=False)] = ord("N")
    for s in rng.choice(len(arr) - 40, 6, replace=False):
        arr[s:s + 25] = ord("N")
    arr[rng.choice(len(arr), 40, replace=False)] = ord("R")
    genome = wl.Genome(arr.tobytes(), [(0, len(arr))], None)
    reads = wl.synth_reads(304, genome, 3000, 36, 120, 1, p_sub=0.01, p_ins=0.001, indel_max=2,
                           k_rate=0.05)
    ra = np.frombuffer(reads.seq, np.uint8).copy()
    ra[rng.choice(len(ra), 300, replace=False)] = ord("N")
    reads = wl.Reads(ra.tobytes(), reads.off, reads.length, reads.true_pos, reads.strand, reads.k)
    res = host_vs_device(ctx, genome, reads, chunks=(50, 100_000))
    assert sum(r[0] for r in res) > 500


def test_mixed_lengths_two_contigs(ctx):
    """Lengths 50-250 with their own q and k, two contigs with a blocked separator."""
    print("[gpu] device filter mixed", flush=True)
    genome = wl.synth_genome(404, [150_000, 80_000], 0.45, (4, 500, 2000, 30, 0.02, 0.005))
    reads = wl.synth_reads(405, genome, 20_000, 50, 250, 1, p_sub=0.01, p_ins=0.001,
                           indel_max=3, p_random=0.02, k_rate=0.05)
    res = host_vs_device(ctx, genome, reads, chunks=(5000, 200_000))
    assert sum(r[0] for r in res) > 10_000


def test_no_reads(ctx):
    genome = wl.synth_genome(505, [10_000], 0.5)
    reads = wl.Reads(b"", np.zeros(0, np.uint64), np.zeros(0, np.uint16), np.zeros(0, np.int64),
                     np.zeros(0, np.uint8), np.zeros(0, np.uint8))
    idx = wl.Index(genome, [11])
    df = wl.DeviceFilter(ctx, idx, genome)
    d = df.run(reads, 0, 100)
    assert d.n == 0 and len(d.batch_end) == 0 and len(d.perfect_read) == 0
    df.close()
    idx.close()
