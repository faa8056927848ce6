"""From files to records: a two-contig genome in a FASTA file and reads in a FASTQ
file go through the file readers (saruman_amd/csrc/ingest.cpp), the host q-gram
filter of both strands and the same filter on the device, the banded kernels and
the device record writer, with the read names of the file.

The reference's readers (src/util.c:293-454 for reads, src/saruman.c:139-271 for
the genome) take a single genome record; the second contig here is joined after a
separator of N bases that no candidate window may cross.  The FASTA file has
60-column lines, a description on each header and a lower-case stretch; the FASTQ
reads are 36-150 bp with k = ceil(5% of L).  Every score and every record is
checked against the oracle on the same concatenated genome.
"""
import numpy as np
import pytest

import saruman_amd as sa
from saruman_amd import ingest, workload as wl
from saruman_amd.batch import Batch, name_table

pytestmark = pytest.mark.gpu


def write_fasta(path, genome, names):
    with open(path, "w") as f:
        for (lo, hi), name in zip(genome.contigs, names):
            s = genome.seq[lo:hi].decode()
            if name.startswith("chr2"):
                s = s[:1000] + s[1000:4000].lower() + s[4000:]
            f.write(f">{name}\n")
            for i in range(0, len(s), 60):
                f.write(s[i:i + 60] + "\n")


def write_fastq(path, reads, names):
    a = np.frombuffer(reads.seq, np.uint8)
    with open(path, "w") as f:
        for i in range(reads.n):
            o, L = int(reads.off[i]), int(reads.length[i])
            f.write(f"@{names[i]}\n{a[o:o + L].tobytes().decode()}\n+\n{'I' * L}\n")


def test_fasta_fastq_two_contigs_to_records(tmp_path, orc):
    print("[gpu] files: two-contig FASTA + FASTQ through filter, kernels and records",
          flush=True)
    genome = wl.synth_genome(0x5A2_00F1, [300_000, 200_000], 0.45)
    pool = wl.synth_reads(0x5A2_00F2, genome, 20_000, 36, 150, 1, p_sub=0.01, p_ins=0.002,
                          indel_max=3, k_rate=0.05)
    contig_names = ["chr1 synthetic 300 kbp", "chr2 synthetic 200 kbp"]
    read_names = [f"read{i} len={int(pool.length[i])}" for i in range(pool.n)]
    gpath, rpath = tmp_path / "genome.fa", tmp_path / "reads.fq"
    write_fasta(gpath, genome, contig_names)
    write_fastq(rpath, pool, read_names)

    g, gnames = ingest.load_genome(gpath, sep=300)
    assert gnames == contig_names and g.seq == genome.seq and g.contigs == genome.contigs
    reads, names, skipped = ingest.load_reads(rpath, error_rate=0.05, k_rule="ceil")
    assert names == read_names and skipped == (0, 0)
    assert reads.seq == pool.seq and (reads.length == pool.length).all()
    assert (reads.k == pool.k).all()

    index = wl.Index(g, wl.needed_qs(reads))
    ctx = sa.Context(0)
    try:
        host = [wl.run_filter(index, g, reads, dr, 5000) for dr in (0, 4)]
        df = wl.DeviceFilter(ctx, index, g)
        try:
            dev = df.run_both(reads, 5000)
        finally:
            df.close()
        for h, d in zip(host, dev):
            for f in ("read", "pos", "flag", "w"):
                assert np.array_equal(getattr(h, f), getattr(d, f)), f
            assert np.array_equal(np.sort(h.perfect_read), np.sort(d.perfect_read))
        db = wl.device_batch(reads, host)
        assert db.n > reads.n // 2
        # no window reaches into the separator
        sep = np.concatenate([[0], np.cumsum(g.blocked.astype(np.int64))])
        assert (sep[db.pos + db.w] - sep[db.pos] == 0).all()
        assert ((db.pos >= g.contigs[1][0]).sum() > 0) and ((db.pos < g.contigs[0][1]).sum() > 0)

        ctx.set_genome(wl.genome_words(g))
        job = sa.Job(ctx)
        try:
            job.upload(0, 1, db.cand, db.read_words, None, ops_total=db.ops_total)
            job.set_records(db.pos, read_id=np.ascontiguousarray(db.read_id, np.int32),
                            names=name_table(names))
            job.run("banded")
            res, _ = job.download()
            text, hits = job.records_text()
        finally:
            job.close()
    finally:
        index.close()
        ctx.close()

    seq = g.seq
    fwd = reads.seq
    rev = wl.revcomp_all(reads).tobytes()
    wins, rds = [], []
    for t in range(db.n):
        p, w = int(db.pos[t]), int(db.w[t])
        o, L = int(reads.off[db.read_id[t]]), int(reads.length[db.read_id[t]])
        wins.append(seq[p:p + w])
        rds.append((rev if db.flag[t] > 3 else fwd)[o:o + L])
    b = Batch(0, 1, int(db.k.max()), wins, rds, db.flag, db.pos,
              [names[int(x)] for x in db.read_id], ks=db.k)
    want = orc.align_batch(b)
    k1 = db.k.astype(np.int64) + 1
    got_s = np.minimum(res["score"].astype(np.int64), k1)
    want_s = np.minimum(want.scores.astype(np.int64), k1)
    assert (got_s == want_s).all(), int((got_s != want_s).sum())
    assert hits == int((want_s < k1).sum()) and hits > db.n // 3
    assert text.decode() == "".join(want.records)
