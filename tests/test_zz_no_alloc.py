"""srm_align_static, the entry point that allocates nothing, against the existing
one it is a variant of (srm_job_upload + srm_job_run(banded) + srm_job_download):
the same results and transcripts, bit for bit, on two calls in a row, the second
with a smaller input; and the oracle's records.  The test's own device buffers
come from hipMalloc (the caller's inputs and outputs); the entry point takes
its scratch from its static device array."""
import ctypes as C
import math

import numpy as np
import pytest

import saruman_amd as sa
from saruman_amd import _lib
from saruman_amd.batch import Job, _layout, context, format_records, pack_codes
from saruman_amd.testing import random_batch

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.live = []

    def put(self, arr):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(arr.nbytes, 1)) == 0
        self.live.append(p)
        assert self.h.hipMemcpy(p, arr.ctypes.data, arr.nbytes, H2D) == 0
        return p

    def zeros(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(nbytes, 1)) == 0
        self.live.append(p)
        assert self.h.hipMemset(p, 0, max(nbytes, 1)) == 0
        return p

    def get(self, p, arr):
        assert self.h.hipMemcpy(arr.ctypes.data, p, arr.nbytes, D2H) == 0
        return arr

    def free(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live = []


def hit_words(res, cand, ks):
    """(candidate, first word, word count) of every hit's transcript."""
    out = []
    for t in range(len(res)):
        if int(res["score"][t]) <= int(ks[t]):
            out.append((t, int(cand["ops_off"][t]), (int(res["n_ops"][t]) + 15) // 16))
    return out


def run_static(hip, lib, batch, lay):
    rw, ww = pack_codes(lay["rcat"]), pack_codes(lay["wcat"])
    cand = lay["cand"]
    n = len(cand)
    d_rw, d_ww = hip.put(rw), hip.put(ww)
    d_res = hip.zeros(n * 8)
    d_ops = hip.zeros(lay["ops_total"] * 4)
    rc = lib.srm_align_static(0, None, batch.match, batch.mismatch, cand.ctypes.data, n, d_rw,
                              rw.size, d_ww, ww.size, d_res, d_ops, lay["ops_total"])
    assert rc == 0, lib.srm_last_error()
    res = hip.get(d_res, np.zeros(n, _lib.RES_DTYPE))
    ops = hip.get(d_ops, np.zeros(max(lay["ops_total"], 1), np.uint32))
    return res, ops


def run_job(batch, lay):
    job = Job(context(0))
    try:
        job.upload(batch.match, batch.mismatch, lay["cand"], pack_codes(lay["rcat"]),
                   pack_codes(lay["wcat"]), ops_total=lay["ops_total"])
        job.run("banded")
        assert job.paths()[0] == batch.n
        return job.download()
    finally:
        job.close()


def test_static_entry_point_matches_the_job_path(orc):
    print("[gpu] static entry point: two calls", flush=True)
    lib = _lib.load()
    hip = Hip()
    rng = np.random.default_rng(0x57A7)
    first = random_batch(rng, 3000, L=(50, 250), k=13, per_cand_k=lambda L: math.ceil(0.05 * L))
    second = random_batch(rng, 700, L=(30, 120), k=4, flags=(0, 1, 2, 3, 4, 5, 6, 7),
                          match=0, mismatch=2, per_cand_k=lambda L: 2 + L // 40)
    try:
        for batch in (first, second):
            lay = _layout(batch)
            res_s, ops_s = run_static(hip, lib, batch, lay)
            res_j, ops_j = run_job(batch, lay)
            assert (res_s == res_j).all()
            ks = batch.kvec()
            for t, w0, nw in hit_words(res_j, lay["cand"], ks):
                assert (ops_s[w0:w0 + nw] == ops_j[w0:w0 + nw]).all(), t
            assert format_records(batch, res_s, ops_s, lay) == orc.align_batch(batch).records
            hip.free()
    finally:
        hip.free()


def test_static_entry_point_refuses_what_it_cannot_hold():
    """Refusals happen on the host, before anything is copied or launched."""
    print("[gpu] static entry point: refusals", flush=True)
    lib = _lib.load()
    rng = np.random.default_rng(7)
    rw = np.zeros(64, np.uint32)
    # more candidates than the host plan holds
    big = np.zeros(70_000, _lib.CAND_DTYPE)
    assert lib.srm_align_static(0, None, 0, 1, big.ctypes.data, len(big), None, 0, None, 0, None,
                                None, 0) == -1
    # a candidate outside the exact band (k above SRM_BAND_KMAX)
    batch = random_batch(rng, 10, L=60, k=20)
    lay = _layout(batch)
    w, r = pack_codes(lay["wcat"]), pack_codes(lay["rcat"])
    hip = Hip()
    try:
        d_r, d_w = hip.put(r), hip.put(w)
        d_res, d_ops = hip.zeros(10 * 8), hip.zeros(lay["ops_total"] * 4)
        assert lib.srm_align_static(0, None, 0, 1, lay["cand"].ctypes.data, 10, d_r, r.size, d_w,
                                    w.size, d_res, d_ops, lay["ops_total"]) == -1
        assert b"plain kernel" in lib.srm_last_error()
        # a batch whose band scratch exceeds the static array: long reads, wide bands
        many = random_batch(rng, 20_000, L=850, k=14, flags=(3,), p_random=0.0,
                            edge_cases=False)
        lay2 = _layout(many)
        assert lib.srm_align_static(0, None, 0, 1, lay2["cand"].ctypes.data, many.n, d_r,
                                    1 << 40, d_w, 1 << 40, d_res, d_ops, 1 << 40) == -2
        assert b"static scratch" in lib.srm_last_error()
    finally:
        hip.free()
    del rw
