"""Candidate batches and their alignment through the C ABI.

A `Batch` holds what one call of the reference entry point receives
(src/hash.c:1094-1097 -> src/cuda-functions.cu:584): per candidate the genome
window, the read in the orientation it is aligned in, the flag
(direction_array), the genome position and the read name, plus the costs and
the error bound.  The extended API also allows a per-candidate error bound.

`align_batch` packs the sequences into 2-bit codes, uploads one job, runs the
banded or plain kernel and formats the reference's TSV record for every hit.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import CAND_DTYPE, RES_DTYPE, SRM_CODE_PAD, check, ptr

_LUT = np.zeros(256, np.uint32)
_LUT[ord("C")] = 1
_LUT[ord("G")] = 2
_LUT[ord("T")] = 3
_ACGT = np.zeros(256, bool)
for _c in b"ACGT":
    _ACGT[_c] = True


def pack_codes(seq: bytes | np.ndarray, pad: int = SRM_CODE_PAD) -> np.ndarray:
    """2-bit codes (A0 C1 G2 T3, 16 per uint32, code j in bits 2j..2j+1) with
    `pad` zero codes before code 0 and at least `pad` after the last one."""
    a = np.frombuffer(seq, np.uint8) if isinstance(seq, (bytes, bytearray)) else seq
    total = pad + len(a) + pad
    total = (total + 15) // 16 * 16 + 16
    buf = np.zeros(total, np.uint32)
    buf[pad:pad + len(a)] = _LUT[a]
    buf = buf.reshape(-1, 16) << (2 * np.arange(16, dtype=np.uint32))
    return np.bitwise_or.reduce(buf, axis=1).astype(np.uint32)


def ops_words(w, k):
    return (np.asarray(w, np.int64) + np.asarray(k, np.int64) + 15) // 16


@dataclass
class Batch:
    match: int
    mismatch: int
    k: int
    windows: list
    reads: list
    flags: np.ndarray
    pos: np.ndarray
    names: list
    ks: np.ndarray | None = None     # per-candidate error bound (extended API)
    first_row0: bool = False         # Q1: candidate 0 stops its traceback at row 0

    @property
    def n(self) -> int:
        return len(self.windows)

    def kvec(self) -> np.ndarray:
        if self.ks is not None:
            return np.asarray(self.ks, np.int64)
        return np.full(self.n, self.k, np.int64)


@dataclass
class AlignResult:
    res: np.ndarray                   # RES_DTYPE per candidate
    ops: np.ndarray                   # transcript words
    records: list                     # TSV line per candidate ('' for non-hits)
    device_ms: float = 0.0
    band_ms: float = 0.0
    plain_ms: float = 0.0
    paths: tuple = ()
    extra: dict = field(default_factory=dict)

    @property
    def scores(self) -> np.ndarray:
        return self.res["score"].astype(np.int64)


def _layout(batch: Batch):
    """Concatenated sequences, code arrays and srm_cand rows for a batch."""
    n = batch.n
    wl = np.fromiter((len(x) for x in batch.windows), np.int64, n)
    rl = np.fromiter((len(x) for x in batch.reads), np.int64, n)
    wcat = b"".join(batch.windows)
    rcat = b"".join(batch.reads)
    woff = np.concatenate([[0], np.cumsum(wl)[:-1]]) if n else np.zeros(0, np.int64)
    roff = np.concatenate([[0], np.cumsum(rl)[:-1]]) if n else np.zeros(0, np.int64)
    ks = batch.kvec()
    ow = ops_words(wl, ks)
    ooff = np.concatenate([[0], np.cumsum(ow)[:-1]]) if n else np.zeros(0, np.int64)
    wa = np.frombuffer(wcat, np.uint8)
    ra = np.frombuffer(rcat, np.uint8)
    # a candidate with any non-ACGT byte goes to the raw-byte path
    wbad = np.concatenate([[0], np.cumsum(~_ACGT[wa])])
    rbad = np.concatenate([[0], np.cumsum(~_ACGT[ra])])
    isbytes = ((wbad[woff + wl] - wbad[woff]) > 0) | ((rbad[roff + rl] - rbad[roff]) > 0)
    cand = np.zeros(n, CAND_DTYPE)
    cand["read_off"] = roff
    cand["win_off"] = woff
    cand["ops_off"] = ooff
    cand["L"] = rl
    cand["w"] = wl
    cand["flag"] = np.asarray(batch.flags, np.uint8)
    cand["k"] = ks
    bits = np.where(isbytes, _lib.SRM_CAND_BYTES, 0).astype(np.uint8)
    if batch.first_row0 and n:
        bits[0] |= _lib.SRM_CAND_STOP_ROW0
    cand["bits"] = bits
    return dict(cand=cand, wcat=wcat, rcat=rcat, woff=woff, roff=roff, wl=wl, rl=rl,
                ops_total=int(ooff[-1] + ow[-1]) if n else 0, any_bytes=bool(isbytes.any()))


class Context:
    """One HIP stream + device pool (srm_ctx) on one GPU."""

    def __init__(self, device: int = 0, pool_cap: int = 0):
        self.lib = _lib.load()
        h = C.c_void_p()
        check(self.lib.srm_ctx_create(device, pool_cap, C.byref(h)), "srm_ctx_create")
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            self.lib.srm_ctx_destroy(self.h)
            self.h = None

    def set_genome(self, words: np.ndarray):
        words = np.ascontiguousarray(words, np.uint32)
        check(self.lib.srm_ctx_set_genome(self.h, ptr(words), words.size), "srm_ctx_set_genome")


class Job:
    """One uploaded candidate batch (srm_job); runnable many times."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.lib = ctx.lib
        h = C.c_void_p()
        check(self.lib.srm_job_create(ctx.h, C.byref(h)), "srm_job_create")
        self.h = h
        self.n = 0
        self.ops_total = 0
        self._keep = ()

    def upload(self, match, mismatch, cand, read_words, win_words=None, read_bytes=None,
               win_bytes=None, ops_total=None):
        cand = np.ascontiguousarray(cand, CAND_DTYPE)
        read_words = np.ascontiguousarray(read_words, np.uint32)
        if win_words is not None:
            win_words = np.ascontiguousarray(win_words, np.uint32)
        rb = np.frombuffer(read_bytes, np.uint8) if read_bytes is not None else None
        wb = np.frombuffer(win_bytes, np.uint8) if win_bytes is not None else None
        self._keep = (cand, read_words, win_words, rb, wb)
        check(self.lib.srm_job_upload(
            self.h, match, mismatch, ptr(cand), len(cand), ptr(read_words), read_words.size,
            ptr(win_words), 0 if win_words is None else win_words.size,
            ptr(rb), 0 if rb is None else rb.size, ptr(wb), 0 if wb is None else wb.size),
            "srm_job_upload")
        self.n = len(cand)
        if ops_total is None:
            ops_total = int((cand["ops_off"].astype(np.int64) +
                             ops_words(cand["w"], cand["k"])).max()) if len(cand) else 0
        self.ops_total = ops_total

    def use_graph(self, on: bool = True):
        check(self.lib.srm_job_use_graph(self.h, int(on)), "srm_job_use_graph")

    def run(self, kernel: str = "auto") -> float:
        check(self.lib.srm_job_run(self.h, _lib.SRM_KERNEL[kernel]), "srm_job_run")
        return float(self.lib.srm_job_last_ms(self.h))

    def kernel_ms(self):
        return (float(self.lib.srm_job_kernel_ms(self.h, 0)),
                float(self.lib.srm_job_kernel_ms(self.h, 1)))

    def paths(self):
        return tuple(int(self.lib.srm_job_path_count(self.h, i)) for i in range(3))

    def download(self):
        res = np.zeros(self.n, RES_DTYPE)
        ops = np.zeros(max(self.ops_total, 1), np.uint32)
        check(self.lib.srm_job_download(self.h, ptr(res), ptr(ops), ops.size), "srm_job_download")
        return res, ops

    def records(self, pos, read_id=None, names=None, name_stride=60, prefix="", out=None,
                copy=True):
        """TSV records of every hit of the last run, formatted on the device
        (srm_job_records).  names: None (prefix + decimal read id) or a bytes
        table of name_stride-byte NUL-terminated names indexed by read_id.
        out: optional preallocated uint8 buffer (page-locked memory makes the
        copy-out fast).  Returns (text bytes, hits), or (byte count, hits) with
        copy=False (the text is then in out)."""
        pos = np.ascontiguousarray(pos, np.int64)
        rid = None if read_id is None else np.ascontiguousarray(read_id, np.int32)
        nb = None if names is None else np.frombuffer(names, np.uint8)
        n_names = 0 if nb is None else nb.size // name_stride
        pfx = prefix.encode()
        hits = C.c_longlong()
        if out is None:
            out = np.empty(1, np.uint8)
        got = self.lib.srm_job_records(self.h, ptr(pos), ptr(rid), ptr(nb), name_stride, n_names,
                                       pfx, ptr(out), out.size, C.byref(hits))
        if got < 0:
            check(int(got), "srm_job_records")
        if got > out.size:          # the first call was a size query
            out = np.empty(int(got), np.uint8)
            got = self.lib.srm_job_records(self.h, ptr(pos), ptr(rid), ptr(nb), name_stride,
                                           n_names, pfx, ptr(out), out.size, C.byref(hits))
            if got < 0:
                check(int(got), "srm_job_records")
        return (out[:got].tobytes() if copy else int(got)), int(hits.value)

    def set_records(self, pos, read_id=None, names=None, name_stride=60, prefix=""):
        """Makes record formatting on the device a stage of every later run
        (srm_job_set_records); arguments as for records()."""
        pos = np.ascontiguousarray(pos, np.int64)
        rid = None if read_id is None else np.ascontiguousarray(read_id, np.int32)
        nb = None if names is None else np.frombuffer(names, np.uint8)
        n_names = 0 if nb is None else nb.size // name_stride
        self._rec_keep = (pos, rid, nb)
        check(self.lib.srm_job_set_records(self.h, ptr(pos), ptr(rid), ptr(nb), name_stride,
                                           n_names, prefix.encode()), "srm_job_set_records")

    def records_text(self, out=None, copy=True):
        """The last run's record text (srm_job_records_text): (bytes, hits), or
        (byte count, hits) with copy=False and the text in out."""
        hits = C.c_longlong()
        total = self.lib.srm_job_records_text(self.h, None, 0, C.byref(hits))
        if total < 0:
            check(int(total), "srm_job_records_text")
        if out is None or out.size < total:
            out = np.empty(max(int(total), 1), np.uint8)
        got = self.lib.srm_job_records_text(self.h, ptr(out), out.size, C.byref(hits))
        if got < 0:
            check(int(got), "srm_job_records_text")
        return (out[:got].tobytes() if copy else int(got)), int(hits.value)

    def close(self):
        if self.h:
            self.lib.srm_job_destroy(self.h)
            self.h = None


_CTX: dict = {}


def context(device: int = 0) -> Context:
    if device not in _CTX:
        _CTX[device] = Context(device)
    return _CTX[device]


def format_records(batch: Batch, res: np.ndarray, ops: np.ndarray, lay=None) -> list:
    """The reference's TSV record for every hit (score <= k), '' otherwise."""
    lib = _lib.load()
    if lay is None:
        lay = _layout(batch)
    ks = batch.kvec()
    out = []
    buf = C.create_string_buffer(8192)
    wa = np.frombuffer(lay["wcat"], np.uint8)
    ra = np.frombuffer(lay["rcat"], np.uint8)
    cand = lay["cand"]
    for t in range(batch.n):
        if int(res["score"][t]) > int(ks[t]):
            out.append("")
            continue
        r = res[t:t + 1]
        o = ops[int(cand["ops_off"][t]):]
        n = lib.srm_format_record(
            batch.names[t].encode(), int(batch.pos[t]) & 0xFFFFFFFF, int(batch.flags[t]),
            int(ks[t]), int(lay["rl"][t]), int(lay["wl"][t]), r.ctypes.data, o.ctypes.data,
            ra.ctypes.data + int(lay["roff"][t]), wa.ctypes.data + int(lay["woff"][t]), buf,
            len(buf))
        if n < 0:
            raise RuntimeError(f"srm_format_record failed for candidate {t}")
        out.append(buf.value.decode())
    return out


def name_table(names, stride: int = 60) -> bytes:
    """Fixed-stride NUL-terminated names (at most stride - 1 bytes each)."""
    return b"".join(n.encode()[:stride - 1].ljust(stride, b"\0") for n in names)


def align_batch(batch: Batch, kernel: str = "auto", device: int = 0, graph: bool = False,
                repeats: int = 1, device_records: bool = False) -> AlignResult:
    """Aligns a batch; records are formatted on the host.  device_records also
    formats them on the device (srm_job_records) into extra["device_text"]."""
    lay = _layout(batch)
    rw = pack_codes(lay["rcat"])
    ww = pack_codes(lay["wcat"])
    ctx = context(device)
    job = Job(ctx)
    try:
        job.upload(batch.match, batch.mismatch, lay["cand"], rw, ww,
                   lay["rcat"] if lay["any_bytes"] else None,
                   lay["wcat"] if lay["any_bytes"] else None, ops_total=lay["ops_total"])
        if graph:
            job.use_graph(True)
        ms = 0.0
        for _ in range(max(1, repeats)):
            ms = job.run(kernel)
        bms, pms = job.kernel_ms()
        res, ops = job.download()
        paths = job.paths()
        extra = {}
        if device_records:
            text, hits = job.records(batch.pos, read_id=np.arange(batch.n, dtype=np.int32),
                                     names=name_table(batch.names), name_stride=60)
            extra = {"device_text": text, "device_hits": hits}
    finally:
        job.close()
    recs = format_records(batch, res, ops, lay)
    return AlignResult(res, ops, recs, ms, bms, pms, paths, extra)
