"""Candidate batches through the device with transfers overlapped.

The reference aligns each flushed batch synchronously on its strand's pool
(src/hash.c:841-851 -> src/cuda-functions.cu:584-1049): copy in, kernel, copy
out, records, one after another.  Here a batch is uploaded from page-locked host
memory, aligned and formatted by one launch sequence (a hipGraph replay of the
band phase and the record stage), and its record text copied back to page-locked
memory.  Two or more lanes, each a context (own HIP stream and buffers) sharing
the resident genome, take batches in turn: while one lane's step runs on the
device, another lane uploads its next batch or copies its records out.
"""
from __future__ import annotations

import ctypes as C
import threading
import time
from dataclasses import dataclass

import numpy as np

from . import _lib
from .batch import Context, Job
from .workload import DeviceBatch, Pinned


@dataclass
class PinnedBatch:
    """A DeviceBatch's upload arrays and record inputs (window positions and read
    ids) copied into page-locked memory."""
    n: int
    cand: np.ndarray
    read_words: np.ndarray
    ops_total: int
    pos: np.ndarray
    read_id: np.ndarray
    n_reads: int
    _bufs: tuple = ()

    @classmethod
    def of(cls, db: DeviceBatch, read_id_base: int = 0, n_reads: int = 0):
        cb = Pinned(db.cand.nbytes)
        cand = cb.array[:db.cand.nbytes].view(db.cand.dtype)
        cand[:] = db.cand
        wb = Pinned(db.read_words.nbytes)
        rw = wb.array[:db.read_words.nbytes].view(np.uint32)
        rw[:] = db.read_words
        pbuf = Pinned(max(db.n, 1) * 8)
        pos = pbuf.array[:db.n * 8].view(np.int64)
        pos[:] = db.pos
        ibuf = Pinned(max(db.n, 1) * 4)
        rid = ibuf.array[:db.n * 4].view(np.int32)
        rid[:] = np.asarray(db.read_id, np.int64) + read_id_base
        return cls(db.n, cand, rw, db.ops_total, pos, rid, n_reads, (cb, wb, pbuf, ibuf))

    def free(self):
        for b in self._bufs:
            b.free()
        self._bufs = ()


@dataclass
class LaneResult:
    batch: int
    lane: int
    upload_s: float
    step_ms: float
    copy_s: float
    text_bytes: int
    hits: int


class BatchPipeline:
    def __init__(self, genome_words: np.ndarray, lanes: int = 2, device: int = 0,
                 graph: bool = True, prefix: str = "r", tuning: dict | None = None):
        self.lib = _lib.load()
        self.ctxs = [Context(device) for _ in range(max(1, lanes))]
        if tuning is not None:          # the launch tuning of every lane (tune.apply)
            from . import tune
            for c in self.ctxs:
                tune.apply(c, tuning)
        self.ctxs[0].set_genome(genome_words)
        for c in self.ctxs[1:]:
            _lib.check(self.lib.srm_ctx_borrow_genome(c.h, self.ctxs[0].h),
                       "srm_ctx_borrow_genome")
        self.jobs = [Job(c) for c in self.ctxs]
        self.graph = graph
        self.prefix = prefix
        self.out = [None] * len(self.jobs)

    def _lane(self, li: int, batches, order, results, errors, keep_text):
        job = self.jobs[li]
        try:
            for bi in order:
                b = batches[bi]
                t0 = time.perf_counter()
                job.upload(0, 1, b.cand, b.read_words, None, ops_total=b.ops_total)
                job.set_records(b.pos, read_id=b.read_id, prefix=self.prefix)
                job.use_graph(self.graph)
                t1 = time.perf_counter()
                ms = job.run("banded")
                t2 = time.perf_counter()
                total = self.lib.srm_job_records_text(job.h, None, 0, None)
                if total < 0:
                    _lib.check(int(total), "srm_job_records_text")
                if self.out[li] is None or self.out[li].array.size < total:
                    if self.out[li] is not None:
                        self.out[li].free()
                    self.out[li] = Pinned(int(total * 1.25) + 4096)
                hits = C.c_longlong()
                got = self.lib.srm_job_records_text(job.h, self.out[li].p,
                                                    self.out[li].array.size, C.byref(hits))
                if got < 0:
                    _lib.check(int(got), "srm_job_records_text")
                t3 = time.perf_counter()
                results[bi] = LaneResult(bi, li, t1 - t0, ms, t3 - t2, int(got), int(hits.value))
                if keep_text is not None:
                    keep_text[bi] = bytes(self.out[li].array[:got])
        except Exception as e:      # reported by run()
            errors.append(e)

    def run(self, batches, keep_text: bool = False):
        """Runs every batch once; returns (wall seconds, [LaneResult] in batch
        order, {batch: text} when keep_text)."""
        L = len(self.jobs)
        results = [None] * len(batches)
        errors = []
        texts = {} if keep_text else None
        orders = [list(range(li, len(batches), L)) for li in range(L)]
        th = [threading.Thread(target=self._lane, args=(li, batches, orders[li], results, errors,
                                                        texts)) for li in range(L)]
        t0 = time.perf_counter()
        for t in th:
            t.start()
        for t in th:
            t.join()
        wall = time.perf_counter() - t0
        if errors:
            raise errors[0]
        return wall, results, texts

    def close(self):
        for j in self.jobs:
            j.close()
        for c in reversed(self.ctxs):
            c.close()
        for o in self.out:
            if o is not None:
                o.free()
        self.jobs, self.ctxs, self.out = [], [], []

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
