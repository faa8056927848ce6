"""-G mode: mapping reads against a genome read in chunks, as the reference's
pass loop does it (src/saruman.c:139-415).

Each pass (ingest.load_genome_chunks) gets its own q-gram index whose positions
start at the pass's base (util.c:481).  For each strand, the host filter runs
over that pass (srm_filter_run_chunk, hash.c:545-552, 982, 998).  Every filter
batch is one alignment call on the device, as each prepare_cuda_alignments call
is one cuda_needleman_wunsch call (hash.c:1094-1097): its first candidate stops
its traceback at row 0 (Q1).  Records carry genome positions.

Windows come from the genome in genome coordinates (ingest.chunk_genome).  A
window ends at or before its pass's limit, so it never reaches the overhang and
holds the same bytes as the pass text.  The one exception is a start-free window
of a later pass that begins up to k bases before the pass.  There the reference
indexes its pass buffer with a negative offset, which is undefined; here such a
window holds the genome's own bases.

Output order per pass: the forward strand's perfect hits, the reverse strand's
perfect hits (hash.c:580), then the aligned hits of every forward batch, then of
every reverse batch (cuda-functions.cu:987).  The reference runs the two strands
on two threads, so its interleaving of them is not fixed.
"""
from __future__ import annotations

import numpy as np

from . import batch as _batch
from . import workload as wl
from .ingest import GenomeChunk, chunk_genome, load_genome_chunks

_COMP = bytes.maketrans(b"ACGT", b"TGCA")


def _revcomp(s: bytes) -> bytes:
    # util.c:35-57: A<->T, C<->G, other bytes unchanged, then reversed
    return s.translate(_COMP)[::-1]


def _strlen(b: bytes) -> int:
    z = b.find(b"\0")
    return len(b) if z < 0 else z


def map_genome_chunks(chunks: list, reads: wl.Reads, chunk_value: int, names=None, match=0,
                      mismatch=1, threads=0, device=0, stats=None) -> list:
    """TSV records of every pass, in the order above (a list of lines).

    chunks: ingest.GenomeChunk list (one per pass).  names: read names (default
    r<id>).  stats: optional dict that receives per-pass counts.
    """
    if names is None:
        names = [f"r{i}" for i in range(reads.n)]
    whole = chunk_genome(chunks)
    qs = wl.needed_qs(reads)
    seq = reads.seq if isinstance(reads.seq, bytes) else bytes(reads.seq)
    fwd = [seq[int(o):int(o) + int(n)] for o, n in zip(reads.off, reads.length)]
    rev = [_revcomp(r) for r in fwd]
    # construct_qgram_index ends the reference's run at a pass of no more than q
    # bases (util.c:466-469), after the records of the passes before it; here the
    # call fails before mapping anything
    for ch in chunks:
        n = _strlen(ch.text)
        if n <= max(qs):
            raise ValueError(f"genome pass at {ch.base} has {n} bases, not more than q = "
                             f"{max(qs)}")
    out = []
    for ch in chunks:
        n = _strlen(ch.text)
        text = ch.text[:n]
        index = wl.ChunkIndex(text, qs, ch.base, threads)
        try:
            strands = [wl.run_filter_chunk(index, text, ch.limit, reads, dr, chunk_value, threads)
                       for dr in (0, 4)]
        finally:
            index.close()
        for s in strands:
            arrow = "<<" if s.do_reverse else ">>"
            for r, S in zip(s.perfect_read.tolist(), s.perfect_pos.tolist()):
                rd = (rev if s.do_reverse else fwd)[r]
                out.append(f"{names[r]}\t{S}\t{S + len(rd) - 1}\t{arrow}\t{rd.decode()}\t"
                           f"{rd.decode()}\t0\n")
        n_cand = n_hits = 0
        for s in strands:
            lo = 0
            for hi in s.batch_end.tolist():
                rid = s.read[lo:hi]
                pos = s.pos[lo:hi]
                flag = s.flag[lo:hi]
                w = s.w[lo:hi]
                ks = reads.k[rid].astype(np.int64)
                wins = [whole[int(p):int(p) + int(x)] for p, x in zip(pos, w)]
                rds = [(rev if f > 3 else fwd)[int(r)] for r, f in zip(rid, flag)]
                b = _batch.Batch(match, mismatch, int(ks.max()), wins, rds, flag, pos,
                                 [names[int(r)] for r in rid], ks=ks, first_row0=True)
                res = _batch.align_batch(b, device=device)
                out.extend(x for x in res.records if x)
                n_cand += b.n
                n_hits += sum(1 for x in res.records if x)
                lo = hi
        if stats is not None:
            stats.setdefault("passes", []).append(
                {"base": ch.base, "limit": ch.limit, "bases": n, "candidates": n_cand,
                 "hits": n_hits, "perfect": sum(len(s.perfect_read) for s in strands)})
    return out


def map_file(genome_path, reads: wl.Reads, chunk: int, read_len: int, chunk_value: int,
             **kw) -> list:
    """The reference's -G run: genome_path read in passes of `chunk` bases with an
    overhang of read_len (its -l), every pass mapped (map_genome_chunks)."""
    return map_genome_chunks(load_genome_chunks(genome_path, chunk, read_len), reads,
                             chunk_value, **kw)


__all__ = ["GenomeChunk", "map_genome_chunks", "map_file"]
