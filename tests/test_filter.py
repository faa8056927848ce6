"""The product's C++ candidate generation (saruman_amd/csrc/filter.cpp) against
the line-by-line Python transcription of the reference filter
(oracle/filter_oracle.py, src/hash.c:40-1111): same batches, same candidates in
the same order (read, position, flag, window length), same perfect hits."""
import numpy as np
import pytest

from oracle import filter_oracle as fo
from saruman_amd import workload as wl


def strand_reads(reads: wl.Reads):
    a = np.frombuffer(reads.seq, np.uint8)
    return [a[int(o):int(o) + int(n)].tobytes() for o, n in zip(reads.off, reads.length)]


def make(seed, glen, n, L, k, mode=0, repeats=None, tail_reads=0):
    genome = wl.synth_genome(seed, [glen], 0.5, repeats)
    reads = wl.synth_reads(seed + 1, genome, n, L, L, mode, max_err=k, p_sub=0.8, p_ins=0.1,
                           k=k)
    if tail_reads:
        # reads copied from the last bases of the genome: their flag-1/3 windows run
        # past the end and are skipped (hash.c:982-984)
        g = genome.seq
        extra = []
        for j in range(tail_reads):
            t = bytearray(g[glen - L - j:glen - j])
            t[L // 2] = ord("A") if t[L // 2] != ord("A") else ord("C")   # not a perfect hit
            extra.append(bytes(t))
        seq = reads.seq + b"".join(extra)
        off = np.concatenate([reads.off, len(reads.seq) + L * np.arange(tail_reads)]).astype(np.uint64)
        reads = wl.Reads(seq, off, np.full(n + tail_reads, L, np.uint16),
                         np.concatenate([reads.true_pos, np.full(tail_reads, -1)]),
                         np.concatenate([reads.strand, np.zeros(tail_reads, np.uint8)]),
                         np.full(n + tail_reads, k, np.uint8))
    return genome, reads


def compare(genome, reads, L, k, chunk):
    q = wl.filter_params(L, k)[0]
    idx = wl.Index(genome, [q])
    rl = strand_reads(reads)
    names = [f"r{i}" for i in range(reads.n)]
    oidx = fo.build_index(genome.seq, q)
    total = 0
    for dr in (0, 4):
        want_batches, want_perfect = fo.run_strand(genome.seq, rl, names, L, k, dr, chunk, oidx)
        got = wl.run_filter(idx, genome, reads, dr, chunk)
        ends = [0] + got.batch_end.tolist()
        assert len(ends) - 1 == len(want_batches)
        for bi, wb in enumerate(want_batches):
            lo, hi = ends[bi], ends[bi + 1]
            g = list(zip(got.read[lo:hi].tolist(), got.pos[lo:hi].tolist(),
                         got.flag[lo:hi].tolist(), got.w[lo:hi].tolist()))
            w = [(c.read_id, c.pos, c.flag, c.w) for c in wb]
            assert g == w, (dr, bi)
            total += len(w)
        gp = sorted(zip(got.perfect_read.tolist(), got.perfect_pos.tolist()))
        wp = sorted((int(line.split("\t")[0][1:]), int(line.split("\t")[1]))
                    for line in want_perfect)
        assert gp == wp
    idx.close()
    return total


def test_filter_small_uniform():
    genome, reads = make(101, 30_000, 400, 36, 2, tail_reads=6)
    skips = fo.STATS["skip"]
    assert compare(genome, reads, 36, 2, chunk=60) > 200
    assert fo.STATS["skip"] > skips          # out-of-genome windows were skipped


def test_filter_repeats_spill():
    # heavy repeats: q-grams with many positions make the first pass read past
    # the end of list b (Q12), into the next lists and the next read's lists
    genome, reads = make(202, 20_000, 300, 40, 3, repeats=(3, 300, 600, 40, 0.5, 0.01))
    spills = fo.STATS["spill"]
    assert compare(genome, reads, 40, 3, chunk=25) > 150
    assert fo.STATS["spill"] > spills        # list-b reads landed past list b


@pytest.mark.parametrize("L,k", [(50, 3), (100, 5)])
def test_filter_lengths(L, k):
    genome, reads = make(303 + L, 25_000, 250, L, k, mode=0)
    assert compare(genome, reads, L, k, chunk=40) > 100


def test_params_match_reference_derivation():
    for L in range(36, 251):
        k = int(wl.k_for(L))
        P = fo.derive_params(L, k)
        assert wl.filter_params(L, k) == (P.q, P.rest, P.n, P.exact_tresh)


def test_filter_leak_after_skip():
    """A flag-3 candidate whose window runs past the genome end is skipped without
    the reset of hash.c:1079-1080, so the next candidate (flag 1) inherits
    start_offset = k and gets a window of L + 2k bytes (Q10)."""
    L, k = 36, 2
    genome = wl.synth_genome(404, [20_000], 0.5)
    g = genome.seq
    glen = len(g)

    def mutated(s, i):
        t = bytearray(s)
        t[i] = ord("A") if t[i] != ord("A") else ord("C")
        return bytes(t)

    rd = []
    rd.append(mutated(g[5000:5000 + L], 20))
    rd.append(mutated(g[glen - L - 1:glen - 1], 3))    # error in q-gram 0 -> flag 3, skipped
    rd.append(mutated(g[9000:9000 + L], 30))           # flag 1: inherits s = k
    rd.append(mutated(g[12000:12000 + L], 5))
    seq = b"".join(rd)
    n = len(rd)
    reads = wl.Reads(seq, (L * np.arange(n)).astype(np.uint64), np.full(n, L, np.uint16),
                     np.full(n, -1), np.zeros(n, np.uint8), np.full(n, k, np.uint8))
    leaks = fo.STATS["leak"]
    compare(genome, reads, L, k, chunk=100)
    assert fo.STATS["leak"] > leaks
    got = wl.run_filter(wl.Index(genome, [wl.filter_params(L, k)[0]]), genome, reads, 0, 100)
    i = got.read.tolist().index(2)
    assert got.flag[i] == 1 and got.w[i] == L + 2 * k


def _comp(b: bytes) -> bytes:
    return b.translate(bytes.maketrans(b"ACGT", b"TGCA"))[::-1]


def test_exact_reads_are_perfect_hits():
    """Independent of the transcription: a read that occurs verbatim at genome
    offset T > 0 has its q-gram 0 listed at T, so the perfect check at S = T
    fires (hash.c:550-583) -- on strand 0 for a forward copy, on strand 4 for a
    reverse-complement copy.  Both the transcription and the product filter must
    report every one of them (offset 0 is the list terminator, Q14, so T >= 1)."""
    L, k = 36, 2
    genome = wl.synth_genome(606, [40_000], 0.5)
    rng = np.random.default_rng(606)
    T = rng.integers(1, len(genome.seq) - L, 300)
    strand = rng.integers(0, 2, 300)
    rd = [genome.seq[t:t + L] if s == 0 else _comp(genome.seq[t:t + L]) for t, s in zip(T, strand)]
    n = len(rd)
    reads = wl.Reads(b"".join(rd), (L * np.arange(n)).astype(np.uint64), np.full(n, L, np.uint16),
                     T.astype(np.int64), strand.astype(np.uint8), np.full(n, k, np.uint8))
    q = wl.filter_params(L, k)[0]
    idx = wl.Index(genome, [q])
    oidx = fo.build_index(genome.seq, q)
    for dr, want_strand in ((0, 0), (4, 1)):
        got = wl.run_filter(idx, genome, reads, dr, 100)
        gp = set(zip(got.perfect_read.tolist(), got.perfect_pos.tolist()))
        _, wperf = fo.run_strand(genome.seq, rd, [f"r{i}" for i in range(n)], L, k, dr, 100, oidx)
        wp = {(int(x.split("\t")[0][1:]), int(x.split("\t")[1])) for x in wperf}
        need = {(i, int(T[i])) for i in range(n) if strand[i] == want_strand}
        assert need <= gp and need <= wp
    idx.close()


def test_single_substitution_reads_are_found():
    """q-gram lemma: a 36 bp read with one substitution keeps floor(L/q) - 1 = 2
    of its 3 non-overlapping q-grams exact (q = 11, k = 2), enough for
    exact_tresh = 2, so its true window is among the candidates.  The first-pass
    scan of Q12 may skip index 0 of a list, so on a random genome (lists of one
    entry) nearly all -- not provably all -- are found."""
    L, k = 36, 2
    genome = wl.synth_genome(707, [60_000], 0.5)
    rng = np.random.default_rng(707)
    n = 400
    T = rng.integers(1, len(genome.seq) - L - 2 * k, n)
    rd = []
    for t in T:
        r = bytearray(genome.seq[t:t + L])
        j = int(rng.integers(0, L))
        r[j] = b"ACGT"[(b"ACGT".index(r[j]) + int(rng.integers(1, 4))) % 4]
        rd.append(bytes(r))
    reads = wl.Reads(b"".join(rd), (L * np.arange(n)).astype(np.uint64), np.full(n, L, np.uint16),
                     T.astype(np.int64), np.zeros(n, np.uint8), np.full(n, k, np.uint8))
    q = wl.filter_params(L, k)[0]
    idx = wl.Index(genome, [q])
    got = wl.run_filter(idx, genome, reads, 0, 1000)
    idx.close()
    found = set()
    for r, p, w in zip(got.read.tolist(), got.pos.tolist(), got.w.tolist()):
        if p <= T[r] and T[r] + L <= p + w:
            found.add(r)
    assert len(found) >= 0.99 * n
