"""FASTA/FASTQ input (saruman_amd/csrc/ingest.cpp) with the reference reader's
filters (src/util.c:293-454): upper case, names cut to 59 characters, reads
shorter than the fixed length skipped and longer ones cut, reads with more than
k = floor(e * L) N dropped; multi-contig genomes joined by an N separator."""
import numpy as np

from saruman_amd import ingest


def test_reads_fasta_reference_filters(tmp_path):
    p = tmp_path / "r.fa"
    long_name = "x" * 80
    p.write_text(
        ">r0 first\nacgtACGTAC\n"        # lower case -> upper
        ">r1\nACGTACGTACGTAA\n"          # longer than 10 -> cut
        ">r2\nACGTA\n"                   # shorter -> skipped
        f">{long_name}\nNACGTACGTA\n"    # one N, k = floor(0.1 * 10) = 1 -> kept
        ">r4\nNNACGTACGT\n"              # two N > k -> dropped
    )
    reads, names, (short, nn) = ingest.load_reads(p, fixed_len=10, error_rate=0.1, k_rule="floor")
    assert names == ["r0 first", "r1", "x" * 59]
    a = np.frombuffer(reads.seq, np.uint8)
    got = [bytes(a[int(o):int(o) + int(n)]) for o, n in zip(reads.off, reads.length)]
    assert got == [b"ACGTACGTAC", b"ACGTACGTAC", b"NACGTACGTA"]
    assert (short, nn) == (1, 1)
    assert list(reads.k) == [1, 1, 1]


def test_reads_fastq_variable_length(tmp_path):
    p = tmp_path / "r.fq"
    p.write_text("@a\nACGTACGTACGTACGTACGTACGT\n+\nIIIIIIIIIIIIIIIIIIIIIIII\n"
                 "@b\nACGTAC\n+\nIIIIII\n")
    reads, names, _ = ingest.load_reads(p, error_rate=0.05)
    assert names == ["a", "b"]
    assert list(reads.length) == [24, 6]
    assert list(reads.k) == [2, 1]          # ceil(5% L)


def test_genome_multi_contig(tmp_path):
    p = tmp_path / "g.fa"
    p.write_text(">chr1 desc\nACGT\nacgt\n>chr2\nTTTT\n")
    g, names = ingest.load_genome(p, sep=5)
    assert names == ["chr1 desc", "chr2"]
    assert g.seq == b"ACGTACGT" + b"N" * 5 + b"TTTT"
    assert g.contigs == [(0, 8), (13, 17)]
    assert g.blocked is not None and int(g.blocked.sum()) == 5
