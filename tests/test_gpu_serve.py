"""Serving (continuous batching over a resident genome): every request's records
are the same whether it runs alone or beside others, match the oracle pipeline,
and admission by HBM budget changes timing only."""
import time

import numpy as np
import pytest

from saruman_amd import serve, workload as wl

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def data():
    genome = wl.synth_genome(0x5A2_0005, [300_000], 0.5)
    pool = wl.synth_reads(0x5A2_0006, genome, 40_000, 36, 250, 1, p_sub=0.01, p_ins=0.001,
                          indel_max=3, k_rate=0.05)
    return genome, pool


def slices(sizes):
    lo = 0
    out = []
    for s in sizes:
        out.append((lo, lo + s))
        lo += s
    return out


def test_mix_equals_alone(data):
    print("[gpu] serve mix vs alone", flush=True)
    genome, pool = data
    qs = serve.qs_for_lengths(36, 250)
    sizes = [1000, 7000, 2500, 12000, 1500, 6000]
    parts = slices(sizes)
    with serve.Server(genome, qs, n_slots=3, chunk_value=5000) as S:
        # hold admission until every request is filtered, so that the first
        # three are admitted together and share the device
        S.hold(True)
        ids = [S.submit(pool, lo, hi) for lo, hi in parts]
        deadline = time.monotonic() + 60
        while S.queue()[1] < len(parts):
            assert time.monotonic() < deadline, S.queue()
            time.sleep(0.005)
        S.hold(False)
        mix = [S.result(i) for i in ids if S.wait(i)]
    with serve.Server(genome, qs, n_slots=1, chunk_value=5000) as S:
        alone = []
        for lo, hi in parts:
            i = S.submit(pool, lo, hi)
            S.wait(i)
            alone.append(S.result(i))
    for (rm, sm), (ra, sa) in zip(mix, alone):
        assert sm["status"] == 0 and sa["status"] == 0
        assert rm == ra
        assert sm["n_mapped"] == sa["n_mapped"] and sm["n_candidates"] == sa["n_candidates"]
    # requests were in flight together: slot occupancy (admission -> done) intersects
    iv = sorted((s["t_admitted"], s["t_done"]) for _, s in mix)
    overlap = any(iv[j][0] < iv[j - 1][1] for j in range(1, len(iv)))
    assert overlap, iv
    assert len({s["slot"] for _, s in mix}) > 1


@pytest.mark.parametrize("lo,hi,device", [(20_000, 23_000, False), (20_000, 36_000, True)])
def test_request_matches_oracle_pipeline(data, orc, lo, hi, device):
    """One request through the server vs filter -> oracle alignment -> oracle
    records, compared as multisets (record order differs between pipelines).  The
    16,000-read request (2.3 MB of bases, device filter and records) takes the
    server's multi-threaded host passes: reverse complements, code packing,
    perfect-hit lines and the record text's copy-out."""
    print(f"[gpu] serve vs oracle: reads {lo}..{hi}, device {device}", flush=True)
    from saruman_amd.batch import Batch
    genome, pool = data
    sub = wl.Reads(pool.seq, pool.off[lo:hi], pool.length[lo:hi], pool.true_pos[lo:hi],
                   pool.strand[lo:hi], pool.k[lo:hi])
    # request-local read ids and names r<i>
    qs = serve.qs_for_lengths(36, 250)
    with serve.Server(genome, qs, n_slots=2, chunk_value=5000, gpu_filter=device,
                      gpu_records=device) as S:
        i = S.submit(pool, lo, hi)
        S.wait(i)
        got, st = S.result(i)
    local = wl.Reads(b"".join(pool.seq[int(o):int(o) + int(n)] for o, n in zip(sub.off, sub.length)),
                     np.concatenate([[0], np.cumsum(sub.length.astype(np.int64))[:-1]]).astype(np.uint64),
                     sub.length, sub.true_pos, sub.strand, sub.k)
    idx = wl.Index(genome, wl.needed_qs(local))
    strands = [wl.run_filter(idx, genome, local, dr, 5000) for dr in (0, 4)]
    rev = wl.revcomp_all(local).tobytes()
    want = []
    g = genome.seq
    for s in strands:
        for r, S_ in zip(s.perfect_read, s.perfect_pos):
            o, L = int(local.off[r]), int(local.length[r])
            rd = (rev if s.do_reverse else local.seq)[o:o + L]
            want.append(f"r{r}\t{S_}\t{S_ + L - 1}\t{'<<' if s.do_reverse else '>>'}\t"
                        f"{g[S_:S_ + L].decode()}\t{rd.decode()}\t0\n")
    db = wl.device_batch(local, strands)
    wins, rds = [], []
    for t in range(db.n):
        p, w, r = int(db.pos[t]), int(db.w[t]), int(db.read_id[t])
        o, L = int(local.off[r]), int(local.length[r])
        wins.append(g[p:p + w])
        rds.append((rev if db.flag[t] > 3 else local.seq)[o:o + L])
    b = Batch(0, 1, int(db.k.max()), wins, rds, db.flag, db.pos, [f"r{int(x)}" for x in db.read_id],
              ks=db.k)
    want += [x for x in orc.align_batch(b).records if x]
    got_lines = sorted(x + "\n" for x in got.splitlines())
    assert got_lines == sorted(want)
    assert st["n_candidates"] == db.n


def test_budget_forces_queueing(data):
    print("[gpu] serve budget", flush=True)
    genome, pool = data
    qs = serve.qs_for_lengths(36, 250)
    sizes = [3000, 3000, 3000, 3000]
    parts = slices(sizes)
    with serve.Server(genome, qs, n_slots=4, chunk_value=5000) as S:
        free = [S.result(i) for i in [S.submit(pool, lo, hi) for lo, hi in parts] if S.wait(i)]
    # a budget that admits one such request at a time
    one = int(1.5 * max(s["dev_bytes"] for _, s in free))
    assert one < 2 * min(s["dev_bytes"] for _, s in free)
    with serve.Server(genome, qs, n_slots=4, chunk_value=5000, hbm_budget=one) as S:
        ids = [S.submit(pool, lo, hi) for lo, hi in parts]
        tight = [S.result(i) for i in ids if S.wait(i)]
    for (a, _), (b, _) in zip(free, tight):
        assert a == b
    iv = sorted((s["t_admitted"], s["t_done"]) for _, s in tight)
    assert all(iv[j][0] >= iv[j - 1][1] - 1e-6 for j in range(1, len(iv))), iv


def test_over_budget_request_fails_cleanly(data):
    print("[gpu] serve over budget", flush=True)
    genome, pool = data
    qs = serve.qs_for_lengths(36, 250)
    # small requests need about 5.5 MiB, the 30,000-read one about 37.5 MiB
    with serve.Server(genome, qs, n_slots=2, chunk_value=5000, hbm_budget=24 << 20) as S:
        small = S.submit(pool, 0, 200)
        big = S.submit(pool, 1000, 31000)
        small2 = S.submit(pool, 200, 400)
        for i in (small, big, small2):
            assert S.wait(i, 120_000)
        assert S.result(small)[1]["status"] == 0
        assert S.result(small2)[1]["status"] == 0
        with pytest.raises(RuntimeError):
            S.result(big)


def test_device_filter_and_records_match_host(data):
    """The server with the q-gram filter and the record formatting on the device
    returns the same records and counts as with both on the host."""
    print("[gpu] serve: device filter and records", flush=True)
    genome, pool = data
    qs = serve.qs_for_lengths(36, 250)
    parts = slices([3000, 9000, 1500, 6000])
    out = {}
    for mode in (False, True):
        with serve.Server(genome, qs, n_slots=2, chunk_value=5000, gpu_filter=mode,
                          gpu_records=mode) as S:
            ids = [S.submit(pool, lo, hi) for lo, hi in parts]
            out[mode] = [S.result(i) for i in ids if S.wait(i)]
    assert len(out[True]) == len(out[False]) == len(parts)
    for (rd, sd), (rh, sh) in zip(out[True], out[False]):
        assert rd == rh
        for f in ("n_candidates", "n_hits", "n_perfect", "n_mapped"):
            assert sd[f] == sh[f], f
