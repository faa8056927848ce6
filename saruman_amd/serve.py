"""Serving: continuous batching over a resident genome (srm_server_* in
saruman_amd/csrc/serve.cpp), plus a Poisson load generator that reports p50/p99
request latency (arrival -> all records of the request on the host) and reads/s.
"""
from __future__ import annotations

import ctypes as C
import threading
import time
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, ptr
from .workload import Genome, Reads, filter_params, k_for


class ServerConfig(C.Structure):
    _fields_ = [("device", C.c_int), ("n_slots", C.c_int), ("hbm_budget", C.c_size_t),
                ("filter_threads", C.c_int), ("format_threads", C.c_int),
                ("chunk_value", C.c_int), ("kernel", C.c_int), ("gpu_filter", C.c_int),
                ("gpu_records", C.c_int)]


class RequestStats(C.Structure):
    _fields_ = [("t_submit", C.c_double), ("t_filtered", C.c_double),
                ("t_admitted", C.c_double), ("t_aligned", C.c_double), ("t_done", C.c_double),
                ("device_ms", C.c_double), ("n_reads", C.c_longlong),
                ("n_candidates", C.c_longlong), ("n_hits", C.c_longlong),
                ("n_perfect", C.c_longlong), ("n_mapped", C.c_longlong),
                ("dev_bytes", C.c_longlong), ("slot", C.c_int), ("status", C.c_int),
                ("message", C.c_char * 160)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["message"] = self.message.decode(errors="replace")
        return d


def qs_for_lengths(Lmin: int, Lmax: int, k=None, k_rate=0.05):
    Ls = np.arange(Lmin, Lmax + 1)
    ks = np.full(len(Ls), k) if k is not None else k_for(Ls, k_rate)
    return sorted({filter_params(int(L), int(kk))[0] for L, kk in zip(Ls, ks)})


class Server:
    def __init__(self, genome: Genome, qs, n_slots=4, hbm_budget=0, filter_threads=0,
                 format_threads=0, chunk_value=200_000, kernel="auto", device=0,
                 gpu_filter=False, gpu_records=False):
        self.lib = _lib.load()
        cfg = ServerConfig(device, n_slots, hbm_budget, filter_threads, format_threads,
                           chunk_value, _lib.SRM_KERNEL[kernel], int(gpu_filter),
                           int(gpu_records))
        q = np.array(sorted(set(qs)), np.int32)
        self._g = np.frombuffer(genome.seq, np.uint8)
        h = C.c_void_p()
        check(self.lib.srm_server_create(C.byref(cfg), ptr(self._g), len(self._g), ptr(q), len(q),
                                         C.byref(h)), "srm_server_create")
        self.h = h

    def submit(self, reads: Reads, lo: int = 0, hi: int | None = None) -> int:
        hi = reads.n if hi is None else hi
        a = np.frombuffer(reads.seq, np.uint8)
        off = np.ascontiguousarray(reads.off[lo:hi], np.uint64)
        ln = np.ascontiguousarray(reads.length[lo:hi], np.uint16)
        kk = np.ascontiguousarray(reads.k[lo:hi], np.uint8)
        rid = self.lib.srm_server_submit(self.h, ptr(a), ptr(off), ptr(ln), ptr(kk), hi - lo)
        if rid <= 0:
            raise RuntimeError(f"srm_server_submit failed ({rid})")
        return int(rid)

    def wait(self, rid: int, timeout_ms: int = -1) -> bool:
        return self.lib.srm_server_wait(self.h, rid, timeout_ms) == 1

    def result(self, rid: int, release: bool = True):
        st = RequestStats()
        need = self.lib.srm_server_result(self.h, rid, None, 0, C.byref(st), 0)
        if need < 0:
            raise RuntimeError(f"request {rid} not done")
        buf = C.create_string_buffer(max(int(need), 1))
        self.lib.srm_server_result(self.h, rid, buf, len(buf), C.byref(st), int(release))
        if st.status != 0:
            raise RuntimeError(f"request {rid} failed with status {st.status}: "
                               f"{st.message.decode(errors='replace')}")
        return buf.raw[:need].decode(), st.as_dict()

    def hold(self, on: bool = True) -> None:
        """Pauses (True) or resumes (False) admission; running requests finish."""
        self.lib.srm_server_hold(self.h, int(on))

    def memory(self):
        """(HBM the alignment slots account for now, the most at once), bytes"""
        a, b = C.c_size_t(), C.c_size_t()
        self.lib.srm_server_memory(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def queue(self):
        """(not yet filtered, waiting for admission, on a slot)"""
        a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self.lib.srm_server_queue(self.h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def close(self):
        if self.h:
            self.lib.srm_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


@dataclass
class LoadResult:
    latencies: np.ndarray      # seconds, per request
    stats: list
    reads: int
    mapped: int
    wall: float

    @property
    def p50(self):
        return float(np.percentile(self.latencies, 50))

    @property
    def p99(self):
        return float(np.percentile(self.latencies, 99))


def request_sizes(rng, n_requests, lo=1_000, hi=200_000):
    """Request sizes log-uniform in [lo, hi]."""
    return np.exp(rng.uniform(np.log(lo), np.log(hi), n_requests)).astype(np.int64)


def poisson_load(server: Server, pool: Reads, rate_hz: float, sizes, seed=0,
                 keep_records=False) -> LoadResult:
    """Submits requests of the given sizes (consecutive slices of the read pool,
    wrapping around) at Poisson arrival times; waits for all of them."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_hz, len(sizes))
    ids, records = [], {}
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]]) % max(pool.n - int(max(sizes)), 1)
    t0 = time.perf_counter()
    when = t0 + np.cumsum(gaps)

    def submitter():
        for i, sz in enumerate(sizes):
            d = when[i] - time.perf_counter()
            if d > 0:
                time.sleep(d)
            lo = int(starts[i])
            ids.append(server.submit(pool, lo, lo + int(sz)))

    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    stats = []
    for rid in ids:
        server.wait(rid)
        rec, st = server.result(rid)
        stats.append(st)
        if keep_records:
            records[rid] = rec
    wall = time.perf_counter() - t0
    lat = np.array([s["t_done"] - s["t_submit"] for s in stats])
    res = LoadResult(lat, stats, int(sum(s["n_reads"] for s in stats)),
                     int(sum(s["n_mapped"] for s in stats)), wall)
    res.records = records
    return res
