"""-G mode end to end on the GPU (saruman_amd/chunked.py): a FASTA genome read
in passes (src/saruman.c:139-415), each pass indexed and filtered on the host and
its batches aligned by the device kernels, against the oracle's restatement of
the same passes (oracle/genome_chunks.py, oracle/filter_oracle.py with
base/limit, the scalar C oracle for the alignments, each filter batch one call
whose first candidate stops at row 0).  Every record must be identical: perfect
hits as a set per pass, aligned hits in order.  Reads are placed across the
pass starts and limits, so that windows past a limit are dropped and start-free
windows begin before their pass."""
import numpy as np
import pytest

from oracle import filter_oracle as fo
from oracle import genome_chunks as ogc
from saruman_amd import chunked, ingest
from saruman_amd import workload as wl
from saruman_amd.batch import Batch

pytestmark = pytest.mark.gpu


def fasta(bases: bytes, width: int) -> bytes:
    return b">chrT\n" + b"\n".join(bases[i:i + width] for i in range(0, len(bases), width)) + b"\n"


def boundary_reads(genome: bytes, n, L, k, starts):
    g = wl.Genome(genome, [(0, len(genome))], None)
    reads = wl.synth_reads(0x5A2C, g, n, L, L, 0, max_err=k, p_sub=0.8, p_ins=0.1, k=k)
    extra = []
    for p in starts:
        for j in (L // 2, 2):
            t = bytearray(genome[p:p + L])
            t[j] = ord("A") if t[j] != ord("A") else ord("C")
            extra.append(bytes(t))
    m = len(extra)
    return wl.Reads(reads.seq + b"".join(extra),
                    np.concatenate([reads.off, len(reads.seq) + L * np.arange(m)]).astype(np.uint64),
                    np.full(n + m, L, np.uint16), np.full(n + m, -1, np.int64),
                    np.zeros(n + m, np.uint8), np.full(n + m, k, np.uint8))


def oracle_records(data, G, L, k, reads, chunk_value):
    from oracle import oracle as orc
    orc.load()
    passes = ogc.read_chunks(data, G, L)
    whole = ogc.base_stream(passes)
    q = wl.filter_params(L, k)[0]
    rl = [reads.seq[int(o):int(o) + L] for o in reads.off]
    names = [f"r{i}" for i in range(reads.n)]
    per_pass = []
    before = 0
    for text, base, limit in passes:
        idx = fo.build_index(text, q, base)
        perfect, aligned = [], []
        for dr in (0, 4):
            batches, perf = fo.run_strand(text, rl, names, L, k, dr, chunk_value, idx, base=base,
                                          limit=limit, window_genome=whole)
            perfect += perf
            for cands in batches:
                before += sum(1 for c in cands if c.pos < base)
                b = Batch(0, 1, k, [c.window for c in cands], [c.read for c in cands],
                          np.array([c.flag for c in cands], np.uint8),
                          np.array([c.pos for c in cands], np.int64),
                          [names[c.read_id] for c in cands], first_row0=True)
                aligned += [x for x in orc.align_batch(b).records if x]
        per_pass.append((perfect, aligned))
    return per_pass, before


def test_chunked_mapping_against_oracle(tmp_path):
    rng = np.random.default_rng(0x5A2C)
    bases = bytes(rng.choice(np.frombuffer(b"ACGT", np.uint8), 40_000))
    G, L, k, chunk_value = 12_000, 50, 2, 300
    data = fasta(bases, 60)
    path = tmp_path / "genome.fa"
    path.write_bytes(data)
    starts = [B + d for B in (G, 2 * G, 3 * G) for d in range(-L - 3, 6)]
    starts += [B - L - j for B in (G, 2 * G, 3 * G) for j in range(-2, 8)]
    reads = boundary_reads(bases, 2000, L, k, starts)
    print(f"[gpu] -G mapping: {len(bases)} bp in passes of {G}, {reads.n} reads of {L} bp",
          flush=True)
    chunks = ingest.load_genome_chunks(path, G, L)
    assert [(c.text, c.base, c.limit) for c in chunks] == ogc.read_chunks(data, G, L)
    stats = {}
    got = chunked.map_genome_chunks(chunks, reads, chunk_value, stats=stats)
    want, before = oracle_records(data, G, L, k, reads, chunk_value)
    assert before > 0 and fo.STATS["skip"] > 0
    assert len(stats["passes"]) == len(want) == 4
    i = 0
    for st, (perfect, aligned) in zip(stats["passes"], want):
        n_perf = st["perfect"]
        assert n_perf == len(perfect)
        assert sorted(got[i:i + n_perf]) == sorted(perfect)
        i += n_perf
        assert got[i:i + len(aligned)] == aligned
        assert st["hits"] == len(aligned) and st["candidates"] >= len(aligned)
        i += len(aligned)
    assert i == len(got)
    assert sum(st["hits"] for st in stats["passes"]) > 1000
