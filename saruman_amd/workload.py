"""Synthetic workloads, the host filter, and candidate batches for the device.

Pipeline (all host steps are C++ in libsaruman_hip.so):
  genome (+ repeats / contigs)  ->  reads  ->  q-gram index  ->  per-strand filter
  (candidates in the reference's batch order + perfect hits)  ->  device batches.

Workloads (seeds 0x5A2U0001..5), one per benchmark configuration:
  small   1 Mbp, 100,000 x 36 bp, 0-2 errors per read (80% sub, 10% ins, 10% del), k 2
  single  5 Mbp 50% GC, 1,000,000 x 100 bp, sub 1%, indel 0.1% (1-3 bp), 5% random, k 5
  mixed   4 Mbp with 200 diverged repeat copies covering 2%, 2,000,000 reads of
          50-250 bp, k = ceil(5% L), sub 1%, indel 0.1%
  scale   12 Mbp in two contigs (8 + 4 Mbp, 45% GC), 10,000,000 x 150 bp,
          sub 1%, indel 0.2%, k 8
  serve   5 Mbp resident genome, request batches of 1,000-200,000 reads of 36-250 bp
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import CAND_DTYPE, SRM_CODE_PAD, check, ptr
from .batch import pack_codes

_vp = C.c_void_p


def lib():
    return _lib.load()


def filter_params(L: int, k: int):
    q, rest, n, et = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    check(lib().srm_filter_params(L, k, C.byref(q), C.byref(rest), C.byref(n), C.byref(et)))
    return q.value, rest.value, n.value, et.value


def k_for(L, rate=0.05):
    """k = ceil(5% of L) for the mixed-length workloads."""
    return np.ceil(np.asarray(L, np.float64) * rate - 1e-9).astype(np.int64)


@dataclass
class Genome:
    seq: bytes
    contigs: list                     # [(lo, hi)] in concatenated coordinates
    blocked: np.ndarray | None = None  # separator mask (uint8) or None

    @property
    def n(self):
        return len(self.seq)


def synth_genome(seed: int, contig_lengths, gc=0.5, repeats=None, sep=300) -> Genome:
    """Contigs joined by `sep` N bases (only when there are several)."""
    L = lib()
    parts, contigs, pos = [], [], 0
    for i, cl in enumerate(contig_lengths):
        buf = C.create_string_buffer(int(cl))
        check(L.srm_synth_genome(seed + 7919 * i, int(cl), float(gc), buf))
        if repeats:
            L.srm_synth_repeats(seed + 31 * i, buf, int(cl), *repeats)
        if i:
            parts.append(b"N" * sep)
            pos += sep
        parts.append(buf.raw)
        contigs.append((pos, pos + int(cl)))
        pos += int(cl)
    seq = b"".join(parts)
    blocked = None
    if len(contig_lengths) > 1:
        blocked = (np.frombuffer(seq, np.uint8) == ord("N")).astype(np.uint8)
    return Genome(seq, contigs, blocked)


@dataclass
class Reads:
    seq: bytes
    off: np.ndarray
    length: np.ndarray
    true_pos: np.ndarray
    strand: np.ndarray
    k: np.ndarray

    @property
    def n(self):
        return len(self.off)

    def name(self, i: int) -> str:
        return f"r{i}"


def synth_reads(seed, genome: Genome, n, Lmin, Lmax, mode, max_err=0, p_sub=0.0, p_ins=0.0,
                indel_max=1, p_random=0.0, k=None, k_rate=None) -> Reads:
    L = lib()
    lo = np.array([c[0] for c in genome.contigs], np.uint64)
    hi = np.array([c[1] for c in genome.contigs], np.uint64)
    off = np.zeros(n, np.uint64)
    ln = np.zeros(n, np.uint16)
    tp = np.zeros(n, np.int64)
    st = np.zeros(n, np.uint8)
    total = n * Lmax
    out = np.zeros(total, np.uint8)
    gbuf = np.frombuffer(genome.seq, np.uint8)
    check(L.srm_synth_reads(seed, ptr(gbuf), ptr(lo), ptr(hi), len(lo), n, Lmin, Lmax, mode,
                            max_err, p_sub, p_ins, indel_max, p_random, ptr(out), ptr(off),
                            ptr(ln), ptr(tp), ptr(st)))
    used = int(off[-1] + ln[-1]) if n else 0
    kk = np.full(n, k, np.uint8) if k is not None else k_for(ln, k_rate).astype(np.uint8)
    return Reads(out[:used].tobytes(), off, ln, tp, st, kk)


class Index:
    def __init__(self, genome: Genome, qs, threads=0):
        self.lib = lib()
        self.qs = sorted(set(int(q) for q in qs))
        h = _vp()
        qa = np.array(self.qs, np.int32)
        self._g = np.frombuffer(genome.seq, np.uint8)
        check(self.lib.srm_index_build(ptr(self._g), len(self._g), ptr(qa), len(qa), threads,
                                       C.byref(h)), "srm_index_build")
        self.h = h

    def close(self):
        if self.h:
            self.lib.srm_index_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ChunkIndex(Index):
    """The q-gram index of one -G chunk: positions base + offset over the chunk
    text (srm_index_build_at)."""

    def __init__(self, text: bytes, qs, base: int, threads=0):
        self.lib = lib()
        self.qs = sorted(set(int(q) for q in qs))
        self.base = int(base)
        h = _vp()
        qa = np.array(self.qs, np.int32)
        self._g = np.frombuffer(text, np.uint8)
        check(self.lib.srm_index_build_at(ptr(self._g), len(self._g), ptr(qa), len(qa), threads,
                                          self.base, C.byref(h)), "srm_index_build_at")
        self.h = h


@dataclass
class StrandCands:
    do_reverse: int
    read: np.ndarray
    pos: np.ndarray
    flag: np.ndarray
    w: np.ndarray
    batch_end: np.ndarray
    perfect_read: np.ndarray
    perfect_pos: np.ndarray
    seconds: float = 0.0

    @property
    def n(self):
        return len(self.read)


def needed_qs(reads: Reads):
    pairs = set(zip(reads.length.tolist(), reads.k.tolist()))
    return sorted({filter_params(L, k)[0] for L, k in pairs})


class Pinned:
    """Page-locked host buffer (srm_host_alloc) viewed as a uint8 array."""

    def __init__(self, nbytes: int):
        self.lib = lib()
        p = self.lib.srm_host_alloc(max(int(nbytes), 1))
        if not p:
            raise MemoryError(f"srm_host_alloc({nbytes}) failed")
        self.p = p
        self.array = np.ctypeslib.as_array((C.c_uint8 * max(int(nbytes), 1)).from_address(p))

    def free(self):
        if self.p:
            self.array = None
            self.lib.srm_host_free(self.p)
            self.p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def pinned_reads(reads: Reads):
    """The reads with their bases in page-locked host memory (the Pinned buffer
    must outlive the returned Reads)."""
    buf = Pinned(len(reads.seq))
    buf.array[:len(reads.seq)] = np.frombuffer(reads.seq, np.uint8)
    view = buf.array[:len(reads.seq)]
    return Reads(view, reads.off, reads.length, reads.true_pos, reads.strand, reads.k), buf


def _reads_args(reads: Reads):
    rs = reads.seq if isinstance(reads.seq, np.ndarray) else np.frombuffer(reads.seq, np.uint8)
    off = np.ascontiguousarray(reads.off, np.uint64)
    ln = np.ascontiguousarray(reads.length, np.uint16)
    kk = np.ascontiguousarray(reads.k, np.uint8)
    return rs, off, ln, kk


def _collect(L, h, do_reverse: int, secs: float) -> StrandCands:
    nc, nb, npf = C.c_uint64(), C.c_uint64(), C.c_uint64()
    L.srm_filter_sizes(h, C.byref(nc), C.byref(nb), C.byref(npf))
    n, b, p = nc.value, nb.value, npf.value
    out = StrandCands(do_reverse, np.zeros(n, np.int32), np.zeros(n, np.int64),
                      np.zeros(n, np.uint8), np.zeros(n, np.uint16), np.zeros(b, np.uint64),
                      np.zeros(p, np.int32), np.zeros(p, np.int64), secs)
    L.srm_filter_copy(h, ptr(out.read), ptr(out.pos), ptr(out.flag), ptr(out.w),
                      ptr(out.batch_end), ptr(out.perfect_read), ptr(out.perfect_pos))
    L.srm_filter_destroy(h)
    return out


def run_filter(index: Index, genome: Genome, reads: Reads, do_reverse: int, chunk_value: int,
               threads: int = 0) -> StrandCands:
    """Host filter (srm_filter_run) of one strand."""
    import time
    L = index.lib
    h = _vp()
    g = np.frombuffer(genome.seq, np.uint8)
    rs, off, ln, kk = _reads_args(reads)
    t0 = time.perf_counter()
    check(L.srm_filter_run(index.h, ptr(g), len(g), ptr(genome.blocked), ptr(rs), ptr(off),
                           ptr(ln), ptr(kk), reads.n, do_reverse, chunk_value, threads,
                           C.byref(h)), "srm_filter_run")
    return _collect(L, h, do_reverse, time.perf_counter() - t0)


def run_filter_chunk(index: ChunkIndex, text: bytes, limit: int, reads: Reads, do_reverse: int,
                     chunk_value: int, threads: int = 0) -> StrandCands:
    """Host filter of one strand over one -G chunk (srm_filter_run_chunk):
    positions are genome positions; windows end at or before `limit`."""
    import time
    L = index.lib
    h = _vp()
    g = np.frombuffer(text, np.uint8)
    rs, off, ln, kk = _reads_args(reads)
    t0 = time.perf_counter()
    check(L.srm_filter_run_chunk(index.h, ptr(g), len(g), int(limit), ptr(rs), ptr(off), ptr(ln),
                                 ptr(kk), reads.n, do_reverse, chunk_value, threads, C.byref(h)),
          "srm_filter_run_chunk")
    return _collect(L, h, do_reverse, time.perf_counter() - t0)


class DeviceFilter:
    """The same filter with its per-read work on the GPU (srm_dfilter_*): the
    index and the genome are copied to ctx's device once."""

    def __init__(self, ctx, index: Index, genome: Genome):
        self.lib = index.lib
        self.index, self.genome = index, genome     # kept alive: the object borrows them
        g = np.frombuffer(genome.seq, np.uint8)
        h = _vp()
        check(self.lib.srm_dfilter_create(ctx.h, index.h, ptr(g), len(g), C.byref(h)),
              "srm_dfilter_create")
        self.h = h

    def run(self, reads: Reads, do_reverse: int, chunk_value: int) -> StrandCands:
        import time
        h = _vp()
        rs, off, ln, kk = _reads_args(reads)
        t0 = time.perf_counter()
        check(self.lib.srm_dfilter_run(self.h, ptr(self.genome.blocked), ptr(rs), ptr(off),
                                       ptr(ln), ptr(kk), reads.n, do_reverse, chunk_value,
                                       C.byref(h)), "srm_dfilter_run")
        return _collect(self.lib, h, do_reverse, time.perf_counter() - t0)

    def run_both(self, reads: Reads, chunk_value: int):
        """Both strands over one upload of the reads: (forward, reverse)."""
        import time
        hf, hr = _vp(), _vp()
        rs, off, ln, kk = _reads_args(reads)
        t0 = time.perf_counter()
        check(self.lib.srm_dfilter_run2(self.h, ptr(self.genome.blocked), ptr(rs), ptr(off),
                                        ptr(ln), ptr(kk), reads.n, chunk_value, C.byref(hf),
                                        C.byref(hr)), "srm_dfilter_run2")
        secs = time.perf_counter() - t0
        return _collect(self.lib, hf, 0, secs), _collect(self.lib, hr, 4, secs)

    def stats(self) -> dict:
        ms = C.c_float()
        a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self.lib.srm_dfilter_stats(self.h, C.byref(ms), C.byref(a), C.byref(b), C.byref(c))
        return {"device_ms": ms.value, "lists": a.value, "list_entries": b.value,
                "overflow_reads": c.value}

    def close(self):
        if getattr(self, "h", None):
            self.lib.srm_dfilter_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_COMP = np.zeros(256, np.uint8)
for _a, _b in zip(b"ACGT", b"TGCA"):
    _COMP[_a] = _b


def revcomp_all(reads: Reads) -> np.ndarray:
    """Reverse complement of every read, at the same offsets."""
    a = np.frombuffer(reads.seq, np.uint8)
    ln = reads.length.astype(np.int64)
    off = reads.off.astype(np.int64)
    rid = np.repeat(np.arange(reads.n), ln)
    pos_in = np.arange(len(a)) - np.repeat(off, ln)
    src = np.repeat(off + ln - 1, ln) - pos_in
    out = a[src]
    comp = _COMP[out]
    keep = comp == 0           # bytes other than ACGT stay unchanged
    comp[keep] = out[keep]
    del rid
    return comp


@dataclass
class DeviceBatch:
    """Candidates of one or more filter batches as srm_cand rows over the
    resident genome and a per-batch read code array (forward reads then their
    reverse complements)."""
    cand: np.ndarray
    read_words: np.ndarray
    ops_total: int
    read_id: np.ndarray
    pos: np.ndarray
    flag: np.ndarray
    k: np.ndarray
    L: np.ndarray
    w: np.ndarray
    extra: dict = field(default_factory=dict)

    @property
    def n(self):
        return len(self.cand)


def genome_words(genome: Genome) -> np.ndarray:
    """Resident genome codes: SRM_CODE_PAD codes of padding before base 0."""
    return pack_codes(genome.seq)


def device_batch(reads: Reads, cands_list, read_words=None) -> DeviceBatch:
    """srm_cand rows for the concatenation of strand candidate lists."""
    rid = np.concatenate([c.read for c in cands_list]).astype(np.int64)
    pos = np.concatenate([c.pos for c in cands_list]).astype(np.int64)
    flag = np.concatenate([c.flag for c in cands_list])
    w = np.concatenate([c.w for c in cands_list]).astype(np.int64)
    L = reads.length[rid].astype(np.int64)
    k = reads.k[rid].astype(np.int64)
    if read_words is None:
        fwd = np.frombuffer(reads.seq, np.uint8)
        both = np.concatenate([fwd, revcomp_all(reads)])
        read_words = pack_codes(both)
    total = len(reads.seq)
    roff = reads.off[rid].astype(np.int64) + np.where(flag > 3, total, 0)
    ow = (w + k + 15) // 16
    ooff = np.concatenate([[0], np.cumsum(ow)[:-1]]) if len(ow) else np.zeros(0, np.int64)
    cand = np.zeros(len(rid), CAND_DTYPE)
    cand["read_off"] = roff
    cand["win_off"] = pos + SRM_CODE_PAD
    cand["ops_off"] = ooff
    cand["L"] = L
    cand["w"] = w
    cand["flag"] = flag
    cand["k"] = k
    ops_total = int(ooff[-1] + ow[-1]) if len(ow) else 0
    return DeviceBatch(cand, read_words, ops_total, rid, pos, flag, k, L, w)


def subset_reads(reads: Reads, lo: int, hi: int) -> Reads:
    """Reads lo..hi-1 with their bases copied out and offsets from 0."""
    a = np.frombuffer(reads.seq, np.uint8) if not isinstance(reads.seq, np.ndarray) else reads.seq
    b0 = int(reads.off[lo]) if hi > lo else 0
    b1 = int(reads.off[hi - 1] + reads.length[hi - 1]) if hi > lo else 0
    return Reads(a[b0:b1].tobytes(), (reads.off[lo:hi] - b0).astype(np.uint64),
                 reads.length[lo:hi], reads.true_pos[lo:hi], reads.strand[lo:hi],
                 reads.k[lo:hi])


def batches_by_reads(reads: Reads, strands, batch_reads: int):
    """The candidates of both strands cut by read id into batches of batch_reads
    reads.  Returns [(lo, hi, DeviceBatch)]: the batch's read ids are local (read
    lo is 0), its read codes hold reads lo..hi-1 and their reverse complements."""
    out = []
    for lo in range(0, reads.n, batch_reads):
        hi = min(reads.n, lo + batch_reads)
        sub = subset_reads(reads, lo, hi)
        parts = []
        for s in strands:
            m = (s.read >= lo) & (s.read < hi)
            pm = (s.perfect_read >= lo) & (s.perfect_read < hi)
            parts.append(StrandCands(s.do_reverse, (s.read[m] - lo).astype(np.int32), s.pos[m],
                                     s.flag[m], s.w[m], np.zeros(0, np.uint64),
                                     (s.perfect_read[pm] - lo).astype(np.int32),
                                     s.perfect_pos[pm]))
        db = device_batch(sub, parts)
        db.extra["perfect_read"] = np.unique(np.concatenate([p.perfect_read for p in parts]))
        out.append((lo, hi, db))
    return out


# ---------------------------------------------------------------------------
# the benchmark configurations
# ---------------------------------------------------------------------------
CONFIGS = {
    "small": dict(seed=0x5A2_0001, contigs=[1_000_000], gc=0.5, n=100_000, L=(36, 36), mode=0,
                  max_err=2, p_sub=0.8, p_ins=0.1, k=2, chunk=100_000),
    "single": dict(seed=0x5A2_0002, contigs=[5_000_000], gc=0.5, n=1_000_000, L=(100, 100),
                   mode=1, p_sub=0.01, p_ins=0.001, indel_max=3, p_random=0.05, k=5,
                   chunk=1_000_000),
    "mixed": dict(seed=0x5A2_0003, contigs=[4_000_000], gc=0.5,
                  repeats=(20, 1000, 5000, 200, 0.02, 0.005), n=2_000_000, L=(50, 250), mode=1,
                  p_sub=0.01, p_ins=0.001, indel_max=3, k_rate=0.05, chunk=1_000_000),
    "scale": dict(seed=0x5A2_0004, contigs=[8_000_000, 4_000_000], gc=0.45, n=10_000_000,
                  L=(150, 150), mode=1, p_sub=0.01, p_ins=0.002, indel_max=3, k=8,
                  chunk=1_000_000),
    "serve": dict(seed=0x5A2_0005, contigs=[5_000_000], gc=0.5, n=0, L=(36, 250), mode=1,
                  p_sub=0.01, p_ins=0.001, indel_max=3, k_rate=0.05, chunk=200_000),
}


def make_workload(name: str, n_reads: int | None = None, scale_genome: float = 1.0):
    cfg = dict(CONFIGS[name])
    contigs = [max(20_000, int(c * scale_genome)) for c in cfg["contigs"]]
    genome = synth_genome(cfg["seed"], contigs, cfg["gc"], cfg.get("repeats"))
    n = cfg["n"] if n_reads is None else n_reads
    reads = synth_reads(cfg["seed"] + 1, genome, n, cfg["L"][0], cfg["L"][1], cfg["mode"],
                        max_err=cfg.get("max_err", 0), p_sub=cfg.get("p_sub", 0.0),
                        p_ins=cfg.get("p_ins", 0.0), indel_max=cfg.get("indel_max", 1),
                        p_random=cfg.get("p_random", 0.0), k=cfg.get("k"),
                        k_rate=cfg.get("k_rate"))
    return cfg, genome, reads
