"""FASTA / FASTQ input through libsaruman_hip.so (saruman_amd/csrc/ingest.cpp).

load_genome: every FASTA record is a contig; contigs are joined by `sep` N bases
and the separator mask keeps candidate windows from spanning two contigs.
load_reads: reads with the reference's filters (upper case, at most k N, names
of at most 59 characters; with fixed_len, shorter reads skipped, longer cut).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, ptr
from .workload import Genome, Reads, k_for


def _load(path, kind, fixed_len=0, n_rate=-1.0, n_mode=1):
    lib = _lib.load()
    h = C.c_void_p()
    check(lib.srm_fastx_load(str(path).encode(), kind, fixed_len, n_rate, n_mode, C.byref(h)),
          f"srm_fastx_load({path})")
    try:
        n, b, ss, sn = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint64()
        lib.srm_fastx_sizes(h, C.byref(n), C.byref(b), C.byref(ss), C.byref(sn))
        seq = np.zeros(max(b.value, 1), np.uint8)
        off = np.zeros(n.value, np.uint64)
        ln = np.zeros(n.value, np.uint32)
        names = np.zeros(max(n.value, 1) * 60, np.uint8)
        lib.srm_fastx_copy(h, ptr(seq), ptr(off), ptr(ln), ptr(names))
    finally:
        lib.srm_fastx_destroy(h)
    nm = [bytes(names[60 * i:60 * i + 60]).split(b"\0")[0].decode(errors="replace")
          for i in range(n.value)]
    return seq[:b.value].tobytes(), off, ln, nm, (ss.value, sn.value)


def load_genome(path, sep: int = 300) -> tuple:
    """Returns (Genome, contig names)."""
    seq, off, ln, names, _ = _load(path, 1)
    parts, contigs, pos = [], [], 0
    for i in range(len(off)):
        if i:
            parts.append(b"N" * sep)
            pos += sep
        parts.append(seq[int(off[i]):int(off[i]) + int(ln[i])])
        contigs.append((pos, pos + int(ln[i])))
        pos += int(ln[i])
    g = b"".join(parts)
    blocked = None
    if len(off) > 1:
        blocked = (np.frombuffer(g, np.uint8) == ord("N")).astype(np.uint8)
    return Genome(g, contigs, blocked), names


def load_reads(path, fixed_len: int = 0, error_rate: float = 0.05, k_rule: str = "ceil",
               kind: int = 0):
    """Reads and names.  k per read: ceil(rate * L) ("ceil", the mixed-length
    workloads) or floor(rate * L) ("floor", the reference's k, saruman.c:42)."""
    n_mode = 1 if k_rule == "ceil" else 0
    seq, off, ln, names, skipped = _load(path, kind, fixed_len, error_rate, n_mode)
    L = ln.astype(np.int64)
    k = (np.ceil(L * error_rate - 1e-9) if n_mode else np.floor(L * error_rate)).astype(np.uint8)
    r = Reads(seq, off.astype(np.uint64), ln.astype(np.uint16), np.full(len(off), -1, np.int64),
              np.zeros(len(off), np.uint8), k)
    return r, names, skipped


@dataclass
class GenomeChunk:
    """One pass of the reference's -G reader: `text` (its G part, then the
    overhang), `base` (genome position of text[0]) and `limit` (the reference's
    total_bases_read after this pass: windows end at or before it)."""
    text: bytes
    base: int
    limit: int


def load_genome_chunks(path, chunk: int, read_len: int) -> list:
    """The genome file cut into chunks of `chunk` bases as the reference's -G
    reader does it, overhang of read_len bases included (srm_genome_chunks_load;
    src/saruman.c:139-271, 415)."""
    lib = _lib.load()
    h = C.c_void_p()
    check(lib.srm_genome_chunks_load(str(path).encode(), int(chunk), int(read_len), C.byref(h)),
          f"srm_genome_chunks_load({path})")
    try:
        n, b = C.c_uint64(), C.c_uint64()
        lib.srm_genome_chunks_sizes(h, C.byref(n), C.byref(b))
        text = np.zeros(max(b.value, 1), np.uint8)
        off = np.zeros(n.value + 1, np.uint64)
        lim = np.zeros(max(n.value, 1), np.uint64)
        lib.srm_genome_chunks_copy(h, ptr(text), ptr(off), ptr(lim))
    finally:
        lib.srm_genome_chunks_destroy(h)
    raw = text[:b.value].tobytes()
    return [GenomeChunk(raw[int(off[i]):int(off[i + 1])], i * int(chunk), int(lim[i]))
            for i in range(n.value)]


def chunk_genome(chunks) -> bytes:
    """Genome positions 0..limit of the last chunk: each chunk's bases up to its
    limit (no window of any chunk reads past that)."""
    return b"".join(c.text[:c.limit - c.base] for c in chunks)
