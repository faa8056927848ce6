"""Failure paths forced from the host.  The reference has none of these: it
silently drops a batch when a host allocation fails (cuda-functions.cu:626-638),
never checks a batch against its device pool (:648-665), and exits the process on
any CUDA error.  Here every failure is an error code for one call or one request,
and the next call or the neighbouring requests are unaffected:
  - code arrays too short for the candidates' offsets are refused by the upload,
    before anything is launched;
  - an allocation that fails (srm_debug_fail_allocs) fails that upload or that
    request only;
  - a serving run whose HBM budget is tight keeps what the slots hold within it,
    and every request still gives the records it gives alone.
"""
import ctypes as C

import numpy as np
import pytest

import saruman_amd as sa
from saruman_amd import _lib, serve, workload as wl
from saruman_amd.batch import Context, Job, _layout, format_records, pack_codes
from saruman_amd.testing import random_batch

pytestmark = pytest.mark.gpu

SRM_EINVAL, SRM_ENOMEM = -1, -2


def banner(msg):
    print(msg, flush=True)


def upload_rc(job, lay, rw, ww):
    lib = job.lib
    cand = lay["cand"]
    return lib.srm_job_upload(job.h, 0, 1, cand.ctypes.data, len(cand), rw.ctypes.data, rw.size,
                              ww.ctypes.data, ww.size, None, 0, None, 0)


def test_upload_rejects_short_code_arrays(orc):
    """srm_job_upload checks every offset against the arrays it was given,
    padding included: arrays shorter than the padding are refused, not wrapped
    around."""
    banner("[gpu] failures: short code arrays")
    lib = _lib.load()
    rng = np.random.default_rng(31)
    batch = random_batch(rng, 300, L=36, k=2)
    lay = _layout(batch)
    rw, ww = pack_codes(lay["rcat"]), pack_codes(lay["wcat"])
    ctx = Context(0)
    job = Job(ctx)
    try:
        for bad_r, bad_w in ((rw[:8], ww), (rw, ww[:8]), (rw[:-40], ww), (rw, ww[:-40])):
            assert upload_rc(job, lay, np.ascontiguousarray(bad_r), np.ascontiguousarray(bad_w)) \
                == SRM_EINVAL
            assert b"out of bounds" in lib.srm_last_error()
        # a well-formed upload on the same job then runs normally
        assert upload_rc(job, lay, rw, ww) == 0
        job.n, job.ops_total = batch.n, lay["ops_total"]
        job.run("banded")
        res, ops = job.download()
    finally:
        job.close()
        ctx.close()
    got = format_records(batch, res, ops, lay)
    want = orc.align_batch(batch)
    assert got == want.records


def test_injected_allocation_failure_fails_one_upload(orc):
    banner("[gpu] failures: injected allocation failure (job)")
    lib = _lib.load()
    rng = np.random.default_rng(32)
    batch = random_batch(rng, 500, L=100, k=5)
    lay = _layout(batch)
    rw, ww = pack_codes(lay["rcat"]), pack_codes(lay["wcat"])
    ctx = Context(0)
    job = Job(ctx)
    try:
        lib.srm_debug_fail_allocs(1)
        try:
            assert upload_rc(job, lay, rw, ww) == SRM_ENOMEM
            assert b"injected" in lib.srm_last_error()
        finally:
            lib.srm_debug_fail_allocs(0)
        assert upload_rc(job, lay, rw, ww) == 0
        job.n, job.ops_total = batch.n, lay["ops_total"]
        job.run("auto")
        res, ops = job.download()
        # a released job allocates again on its next upload
        assert lib.srm_job_release(job.h) == 0
        assert upload_rc(job, lay, rw, ww) == 0
        job.run("auto")
        res2, ops2 = job.download()
    finally:
        job.close()
        ctx.close()
    want = orc.align_batch(batch)
    assert format_records(batch, res, ops, lay) == want.records
    assert format_records(batch, res2, ops2, lay) == want.records


@pytest.fixture(scope="module")
def data():
    genome = wl.synth_genome(0x5A2_0015, [200_000], 0.5)
    pool = wl.synth_reads(0x5A2_0016, genome, 30_000, 36, 250, 1, p_sub=0.01, p_ins=0.001,
                          indel_max=3, k_rate=0.05)
    return genome, pool


def alone(genome, qs, pool, parts):
    out = []
    with serve.Server(genome, qs, n_slots=1, chunk_value=5000) as S:
        for lo, hi in parts:
            i = S.submit(pool, lo, hi)
            S.wait(i)
            out.append(S.result(i))
    return out


def test_injected_allocation_failure_fails_one_request(data):
    """One request whose device allocation fails ends with SRM_ENOMEM and no
    records; the requests before and after it give their records as alone."""
    banner("[gpu] failures: injected allocation failure (serving)")
    genome, pool = data
    lib = _lib.load()
    qs = serve.qs_for_lengths(36, 250)
    parts = [(0, 2000), (2000, 14000), (14000, 16000)]
    want = alone(genome, qs, pool, [parts[0], parts[2]])
    with serve.Server(genome, qs, n_slots=1, chunk_value=5000) as S:
        a = S.submit(pool, *parts[0])
        S.wait(a)
        ra = S.result(a)
        # the larger request must grow the slot's buffers: that allocation fails
        lib.srm_debug_fail_allocs(1)
        try:
            b = S.submit(pool, *parts[1])
            S.wait(b)
        finally:
            lib.srm_debug_fail_allocs(0)
        st = serve.RequestStats()
        need = lib.srm_server_result(S.h, b, None, 0, C.byref(st), 0)
        assert st.status == SRM_ENOMEM and need == 0
        c = S.submit(pool, *parts[2])
        S.wait(c)
        rc = S.result(c)
    assert ra[0] == want[0][0] and rc[0] == want[1][0]


def test_slots_stay_within_hbm_budget(data):
    """Requests of different sizes under a budget that fits about one large
    request beside the small ones: every request succeeds with the records it
    gives alone, and the bytes the slots account for (buffers kept by idle slots,
    reservations of busy ones, each enforced as the context's pool cap) never
    exceed the budget."""
    banner("[gpu] failures: HBM budget")
    genome, pool = data
    qs = serve.qs_for_lengths(36, 250)
    sizes = [1000, 8000, 2000, 5000, 1500, 8000, 1000, 3000]   # 29,500 of the 30,000 reads
    parts, lo = [], 0
    for sz in sizes:
        parts.append((lo, lo + sz))
        lo += sz
    want = alone(genome, qs, pool, parts)
    biggest = max(st["dev_bytes"] for _, st in want)
    budget = int(1.6 * biggest) + (4 << 20)
    with serve.Server(genome, qs, n_slots=3, chunk_value=5000, hbm_budget=budget) as S:
        ids = [S.submit(pool, a, b) for a, b in parts]
        got = []
        for i in ids:
            S.wait(i)
            got.append(S.result(i))
        now, peak = S.memory()
    assert [g[0] for g in got] == [w[0] for w in want]
    assert peak <= budget and now <= budget
    assert peak >= biggest


@pytest.mark.parametrize("shape", ["unit_mixed", "costs", "plain_kernel", "bytes"])
def test_bytes_bound_is_enough(orc, shape):
    """srm_job_bytes_bound: a context capped at what it holds plus the bound runs
    the upload, the record stage and the run of the batch (twice in a row, the
    second batch smaller), and never holds more than the cap."""
    banner(f"[gpu] failures: bytes bound {shape}")
    import math
    lib = _lib.load()
    rng = np.random.default_rng(40 + len(shape))
    kernel = "plain" if shape == "plain_kernel" else "auto"

    def make(n):
        if shape == "costs":
            return random_batch(rng, n, L=(30, 120), k=4, flags=(0, 1, 3, 5, 7), match=1,
                                mismatch=3, per_cand_k=lambda L: 2 + L // 30)
        b = random_batch(rng, n, L=(50, 250), k=13, per_cand_k=lambda L: math.ceil(0.05 * L))
        if shape == "bytes":
            for t in range(0, b.n, 4):
                w = bytearray(b.windows[t])
                w[0] = ord("N")
                b.windows[t] = bytes(w)
        return b

    ctx = Context(0)
    job = Job(ctx)
    try:
        for n in (1500, 600):
            batch = make(n)
            lay = _layout(batch)
            rw, ww = pack_codes(lay["rcat"]), pack_codes(lay["wcat"])
            rb = lay["rcat"] if lay["any_bytes"] else None
            wb = lay["wcat"] if lay["any_bytes"] else None
            cand = lay["cand"]
            bound = lib.srm_job_bytes_bound(
                ctx.h, batch.match, batch.mismatch, cand.ctypes.data, len(cand), rw.size, ww.size,
                0 if rb is None else len(rb), 0 if wb is None else len(wb),
                _lib.SRM_KERNEL[kernel], 1, 60 * batch.n, 59)
            cap = lib.srm_ctx_pool_used(ctx.h) + bound
            assert lib.srm_ctx_set_pool_cap(ctx.h, cap) == 0
            job.upload(batch.match, batch.mismatch, cand, rw, ww, rb, wb,
                       ops_total=lay["ops_total"])
            job.set_records(batch.pos, read_id=np.arange(batch.n, dtype=np.int32),
                            names=sa.batch.name_table(batch.names))
            job.run(kernel)
            text, hits = job.records_text()
            assert lib.srm_ctx_pool_used(ctx.h) <= cap
            want = orc.align_batch(batch)
            assert text == "".join(r for r in want.records if r).encode()
    finally:
        job.close()
        ctx.close()
