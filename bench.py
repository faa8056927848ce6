#!/usr/bin/env python3
"""Benchmark: reads mapped/s of the batched persistent banded kernel vs the plain
one-alignment-per-thread kernel, on one MI355X.

A step = one pass of the alignment hot path (DP fill + traceback of every
candidate of the workload, results left in HBM) over the whole candidate set of
the workload, with the packed genome resident and the candidate batch already
uploaded.  Default workload: the mixed-length one (4 Mbp genome with 2% repeat
copies, 2,000,000 reads of 50-250 bp, k = ceil(5% L), both strands).

Prints one JSON line (rank 0).  `value` = reads mapped per second of the batched
kernel: distinct reads with at least one record (perfect hit from the filter or
aligned hit with score <= k) divided by the step time.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_INT32_TOPS = 39.3      # 256 CU x 64 lanes x 2.4 GHz, int32 VALU ops/s
PEAK_HBM_GBS = 8000.0
OPS_PER_BAND_CELL = 6       # compare + select (substitution), 3 adds, min3

# HBM bytes of one step of the band phase (bits_band_kernel and bits_trace_kernel of every
# class), from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (scaled x2.000 and x1.000
# by a 256 MiB copy kernel), for the exact configuration profiled: workload,
# candidates, launch tuning and the sha256 prefix of the kernel source.
# Source: profiles/bits_band_and_trace_kernels_mixed_2M_reads.md
PMC_TRAFFIC = [
    {"workload": "mixed", "candidates": 1774238, "tuning": "defaults",
     "kernel_sha": "144bad373c94", "bytes": 4343037562},
]


def kernel_sha() -> str:
    """sha256 prefix of the alignment kernels' sources (both band kernel files and
    their shared helpers)."""
    import hashlib
    h = hashlib.sha256()
    for name in ("band_bits.hip", "nw_kernels.hip", "nw_seq.h"):
        with open(os.path.join(ROOT, "saruman_amd", "csrc", name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:12]


def tuning_key(tuning):
    """The launch tuning as reported: "defaults" or {blocks_per_cu, merge_below}."""
    if not tuning or (not any(tuning["blocks_per_cu"]) and tuning.get("merge_below", 0) <= 1
                      and tuning.get("class_streams", 0) == 0):
        return "defaults"
    return {k: tuning.get(k, 0) for k in ("blocks_per_cu", "merge_below", "class_streams")}


def profiled_traffic(workload, candidates, tuning):
    tun = tuning_key(tuning)
    sha = kernel_sha()
    for e in PMC_TRAFFIC:
        if (e["workload"] == workload and e["candidates"] == candidates and e["tuning"] == tun
                and e["kernel_sha"] == sha):
            return e["bytes"]
    return None

WORKLOAD_TEXT = {
    "small": "1 Mbp uniform genome, 100,000 x 36 bp reads, 0-2 errors, k=2, both strands",
    "single": "5 Mbp 50% GC genome, 1,000,000 x 100 bp reads (sub 1%, indel 0.1%, 5% random), "
              "k=5, both strands",
    "mixed": "4 Mbp genome with 200 repeat copies (2%, 0.5% divergence), 2,000,000 reads "
             "50-250 bp (sub 1%, indel 0.1%), k=ceil(5% L), both strands",
    "scale": "12 Mbp genome in 2 contigs (8+4 Mbp, 45% GC), 10,000,000 x 150 bp reads "
             "(sub 1%, indel 0.2%), k=8, both strands",
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def band_cells(L, k, flag):
    """Algorithmic band cells: L(2k+1) for start-anchored flags, L(4k+1) for flags
    with a free start (s = k)."""
    s_free = np.isin(flag, (2, 3, 6, 7))
    return np.where(s_free, L * (4 * k + 1), L * (2 * k + 1)).astype(np.float64)


def hbm_bytes(L, w, k):
    """Algorithmic HBM bytes per candidate: 2-bit read and window, 16 B descriptor,
    2-bit transcript, 8 B result."""
    return (np.ceil(L / 4) + np.ceil(w / 4) + 16 + np.ceil(2 * (L + 3 * k) / 8) + 8).astype(
        np.float64)


def cpu_baseline(genome, reads, db, n_sample, mapped_per_cand):
    """The C oracle (full int16 matrix, scalar) over the first n_sample candidates,
    OpenMP over all host cores."""
    from oracle import oracle as orc
    from saruman_amd.batch import Batch
    from saruman_amd.workload import revcomp_all
    n = min(n_sample, db.n)
    g = genome.seq
    fwd = reads.seq
    rev = revcomp_all(reads).tobytes()
    wins, rds = [], []
    for t in range(n):
        p, w = int(db.pos[t]), int(db.w[t])
        r = int(db.read_id[t])
        o, L = int(reads.off[r]), int(reads.length[r])
        wins.append(g[p:p + w])
        rds.append((rev if db.flag[t] > 3 else fwd)[o:o + L])
    b = Batch(0, 1, int(db.k[:n].max()), wins, rds, db.flag[:n], db.pos[:n],
              [f"r{int(x)}" for x in db.read_id[:n]], ks=db.k[:n])
    res = orc.align_batch(b, with_records=False)
    cand_rate = n / res.seconds
    return dict(value=cand_rate * mapped_per_cand, unit="reads/s", cores=res.threads,
                kind="port",
                sample=f"oracle (full int16 matrix) over the first {n} candidates of the same "
                       f"workload in {res.seconds:.2f} s on {res.threads} threads; "
                       f"{cand_rate:.0f} candidates/s scaled by the workload's reads mapped "
                       f"per candidate ({mapped_per_cand:.3f})",
                candidates_per_s=cand_rate, seconds=res.seconds), res


def serve_leg(args):
    """Continuous batching over the resident 5 Mbp genome: Poisson arrivals, request
    sizes log-uniform in 1,000-200,000 reads of 36-250 bp, k = ceil(5% L).  The
    filter and the record formatting run on the device (the host at the first
    rate as well, for comparison); arrival rates are swept."""
    from saruman_amd import serve
    from saruman_amd import workload as wl
    cfg, genome, pool = wl.make_workload("serve", n_reads=400_000)
    rng = np.random.default_rng(cfg["seed"])
    sizes = serve.request_sizes(rng, args.serve_requests)
    qs = serve.qs_for_lengths(36, 250)
    rates = [float(x) for x in args.serve_rates.split(",")]

    def one(rate, device):
        with serve.Server(genome, qs, n_slots=args.serve_slots, chunk_value=cfg["chunk"],
                          gpu_filter=device, gpu_records=device) as S:
            serve.poisson_load(S, pool, 50.0, [2000, 2000], seed=1)     # warm the slots
            res = serve.poisson_load(S, pool, rate, sizes, seed=cfg["seed"])
        st = res.stats
        dev = sorted((s["t_admitted"], s["t_aligned"]) for s in st)
        overlapped = sum(1 for j in range(1, len(dev)) if dev[j][0] < dev[j - 1][1])
        out = {"rate_hz": rate, "filter_and_records": "device" if device else "host",
               "p50_ms": res.p50 * 1e3, "p99_ms": res.p99 * 1e3,
               "reads_per_s": res.reads / res.wall, "mapped_reads_per_s": res.mapped / res.wall,
               "reads": res.reads, "wall_s": res.wall,
               "device_intervals_overlapping": overlapped,
               "mean_filter_ms": float(np.mean([(s["t_filtered"] - s["t_submit"]) for s in st])) * 1e3,
               "mean_device_ms": float(np.mean([s["device_ms"] for s in st]))}
        log(f"[bench] serving ({out['filter_and_records']} filter/records) at {rate:g}/s: "
            f"p50 {out['p50_ms']:.1f} ms, p99 {out['p99_ms']:.1f} ms, "
            f"{out['reads_per_s']:.0f} reads/s over {len(st)} requests")
        return out

    runs = [one(r, True) for r in rates]
    if not args.serve_device_only:
        runs.append(one(rates[0], False))
    return {"workload": "5 Mbp resident genome; Poisson arrivals; request sizes log-uniform in "
                        "1,000-200,000 reads of 36-250 bp; k=ceil(5% L)",
            "requests": args.serve_requests, "slots": args.serve_slots, "runs": runs}


def records_leg(job, db, reads, genome, res, ops):
    """TSV records of every hit of the step just run: srm_job_records (device,
    including the copy of the text to the host) against srm_format_records on the
    host cores, byte for byte."""
    import ctypes as C
    import saruman_amd as sa
    from saruman_amd import workload as wl
    from saruman_amd._lib import ptr
    lib = sa.load()
    rid = np.ascontiguousarray(db.read_id, np.int32)
    pos = np.ascontiguousarray(db.pos, np.int64)
    text, hits = job.records(pos, read_id=rid, prefix="r")          # first call sizes buffers
    pbuf = wl.Pinned(len(text) + 1)                                  # page-locked output
    t0 = time.perf_counter()
    got, hits = job.records(pos, read_id=rid, prefix="r", out=pbuf.array, copy=False)
    dev_ms = (time.perf_counter() - t0) * 1e3
    text = bytes(pbuf.array[:got])
    pbuf.free()
    fwd = np.frombuffer(reads.seq, np.uint8)
    rev = wl.revcomp_all(reads)
    roff = np.ascontiguousarray(reads.off[db.read_id], np.uint64)
    g = np.frombuffer(genome.seq, np.uint8)
    out = np.empty(len(text) + 1, np.uint8)
    nh = C.c_longlong()
    threads = 16
    t0 = time.perf_counter()
    need = lib.srm_format_records(int(db.n), ptr(res), ptr(ops), ptr(db.cand), ptr(g), ptr(pos),
                                  ptr(fwd), ptr(rev), ptr(roff), ptr(rid), b"r", ptr(out),
                                  out.size, C.byref(nh), threads)
    host_ms = (time.perf_counter() - t0) * 1e3
    same = need == len(text) and out[:need].tobytes() == text and nh.value == hits
    return {"device_ms": dev_ms, "host_ms": host_ms,
            "host_threads": threads, "records": int(hits), "mb": len(text) / 1e6,
            "identical": bool(same)}


def device_filter_leg(ctx, index, genome, reads, chunk, host_strands):
    """Both strands through the GPU q-gram filter (one untimed pass, one timed),
    checked field by field against the host filter's output."""
    from saruman_amd import workload as wl
    df = wl.DeviceFilter(ctx, index, genome)
    preads, buf = wl.pinned_reads(reads)      # bases in page-locked memory
    try:
        df.run_both(preads, chunk)
        t0 = time.perf_counter()
        out = df.run_both(preads, chunk)
        wall = time.perf_counter() - t0
        st = df.stats()
    finally:
        df.close()
        buf.free()
    fields = ("read", "pos", "flag", "w", "batch_end", "perfect_read", "perfect_pos")
    same = all(np.array_equal(getattr(h, f), getattr(d, f))
               for h, d in zip(host_strands, out) for f in fields)
    return {"wall_s": wall, "device_ms": st["device_ms"], "identical": bool(same),
            "host_s": sum(h.seconds for h in host_strands), "lists": st["lists"],
            "list_entries": st["list_entries"], "overflow_reads": st["overflow_reads"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="mixed", choices=sorted(WORKLOAD_TEXT))
    ap.add_argument("--reads", type=int, default=None, help="override the read count")
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-tuning", action="store_true", help="ignore saruman_amd/tuning/*.json")
    ap.add_argument("--no-dfilter", action="store_true", help="skip the GPU q-gram filter leg")
    ap.add_argument("--no-records", action="store_true", help="skip the record formatting leg")
    ap.add_argument("--class-streams", type=int, default=None,
                    help="run the band classes concurrently on N streams (overrides the tuning)")
    ap.add_argument("--serve-requests", type=int, default=24)
    ap.add_argument("--serve-rates", default="4,16,64", help="Poisson arrivals per second (sweep)")
    ap.add_argument("--serve-device-only", action="store_true",
                    help="skip the host-filter serving run")
    ap.add_argument("--serve-slots", type=int, default=4)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported "
              "(candidates are independent; multi-GPU sharding is described in DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    import saruman_amd as sa
    from saruman_amd import workload as wl

    t0 = time.perf_counter()
    cfg, genome, reads = wl.make_workload(args.workload, n_reads=args.reads)
    t_gen = time.perf_counter() - t0
    log(f"[bench] workload {args.workload}: genome {genome.n} bp, {reads.n} reads ({t_gen:.1f}s)")
    t0 = time.perf_counter()
    index = wl.Index(genome, wl.needed_qs(reads))
    t_index = time.perf_counter() - t0
    t0 = time.perf_counter()
    strands = [wl.run_filter(index, genome, reads, dr, cfg["chunk"]) for dr in (0, 4)]
    t_filter = time.perf_counter() - t0
    db = wl.device_batch(reads, strands)
    perfect_reads = np.unique(np.concatenate([s.perfect_read for s in strands]))
    n_batches = sum(len(s.batch_end) for s in strands)
    log(f"[bench] index {t_index:.1f}s, filter {t_filter:.1f}s: {db.n} candidates in "
        f"{n_batches} reference batches, {len(perfect_reads)} reads with perfect hits")

    ctx = sa.Context(0)
    gpu_filter = None
    if not args.no_dfilter:
        gpu_filter = device_filter_leg(ctx, index, genome, reads, cfg["chunk"], strands)
        log(f"[bench] GPU filter: {gpu_filter['wall_s']:.3f}s for both strands "
            f"(device {gpu_filter['device_ms']:.1f} ms), identical to the host filter: "
            f"{gpu_filter['identical']}")
    index.close()
    from saruman_amd import tune
    tuning = None if args.no_tuning else tune.apply_cached(ctx)
    if args.class_streams is not None:
        tuning = dict(tuning or {"blocks_per_cu": [0] * 9, "merge_below": 0})
        tuning["class_streams"] = args.class_streams
        tune.apply(ctx, tuning)
    log(f"[bench] launch tuning: {tuning if tuning else 'defaults'}")
    t0 = time.perf_counter()
    ctx.set_genome(wl.genome_words(genome))
    job = sa.Job(ctx)
    job.upload(0, 1, db.cand, db.read_words, None, ops_total=db.ops_total)
    t_upload = time.perf_counter() - t0
    paths = job.paths()
    log(f"[bench] uploaded in {t_upload:.2f}s; paths banded/plain/bytes = {paths}")
    lib = sa.load()

    # ---- batched persistent kernel: untimed warmup, then exactly K timed steps
    if not args.no_graph:
        job.use_graph(True)
    for _ in range(args.warmup):
        job.run("banded")
    lib.srm_device_sync(0)
    t0 = time.perf_counter()
    dev_ms = []
    for _ in range(args.steps):
        dev_ms.append(job.run("banded"))
    lib.srm_device_sync(0)
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    # event-timed band kernels (non-graph launch, same stream)
    job.use_graph(False)
    job.run("banded")
    band_ms, _ = job.kernel_ms()
    t0 = time.perf_counter()
    res, ops = job.download()
    t_d2h = time.perf_counter() - t0
    hit = res["score"].astype(np.int64) <= db.k
    mapped = np.union1d(np.unique(db.read_id[hit]), perfect_reads)
    n_mapped = len(mapped)
    log(f"[bench] banded: {ms_per_step:.3f} ms/step (device {np.median(dev_ms):.3f} ms, "
        f"band kernels {band_ms:.3f} ms), hits {int(hit.sum())}, reads mapped {n_mapped}")

    # ---- records of every hit: formatted on the device vs on 16 host threads
    records = None
    if not args.no_records:
        records = records_leg(job, db, reads, genome, res, ops)
        log(f"[bench] records: device {records['device_ms']:.1f} ms (incl. the copy to the host), "
            f"host {records['host_ms']:.1f} ms on {records['host_threads']} threads, "
            f"{records['mb']:.0f} MB, identical: {records['identical']}")

    # ---- plain kernel on the same job
    plain = None
    if not args.no_plain:
        job.run("plain")   # one untimed pass (first-launch costs)
        lib.srm_device_sync(0)
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            job.run("plain")
        lib.srm_device_sync(0)
        p_ms = (time.perf_counter() - t0) * 1e3 / args.plain_steps
        pres, pops = job.download()
        same = bool((np.minimum(pres["score"], db.k + 1) == np.minimum(res["score"], db.k + 1)).all()
                    and (pres["n_ops"] == res["n_ops"]).all()
                    and (pres["end_row"] == res["end_row"]).all()
                    and (pres["end_col"] == res["end_col"]).all()
                    and (pops == ops).all())
        plain = dict(ms_per_step=p_ms, value=n_mapped / (p_ms / 1e3), results_identical=same)
        log(f"[bench] plain: {p_ms:.1f} ms/step, identical results: {same}")
    job.close()

    L = db.L.astype(np.float64)
    W = db.w.astype(np.float64)
    K = db.k.astype(np.float64)
    ref_cells = float((L * W).sum())
    bcells = float(band_cells(L, K, db.flag).sum())
    achieved = OPS_PER_BAND_CELL * bcells / (band_ms / 1e3) / 1e12
    hbm = float(hbm_bytes(L, W, K).sum())

    cpu = None
    if not args.no_cpu:
        cpu, _ = cpu_baseline(genome, reads, db, args.cpu_sample, n_mapped / max(db.n, 1))
        log(f"[bench] cpu oracle: {cpu['candidates_per_s']:.0f} candidates/s on {cpu['cores']} threads")

    serving = None
    if not args.no_serve:
        serving = serve_leg(args)

    value = n_mapped / (ms_per_step / 1e3)
    out = {
        "metric": "reads mapped/sec (batched persistent banded kernel; GCUPS and the plain "
                  "kernel alongside)",
        "value": value,
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {WORKLOAD_TEXT[args.workload]}",
                   "reads": int(reads.n), "candidates": int(db.n), "reads_mapped": int(n_mapped),
                   "hits": int(hit.sum()), "match": 0, "mismatch": 1, "parallelism": "1 GPU",
                   "graph": not args.no_graph,
                   "tuning": tuning_key(tuning)},
        "gcups": ref_cells / (ms_per_step / 1e3) / 1e9,
        "band_gcups": bcells / (band_ms / 1e3) / 1e9,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_INT32_TOPS,
                     "unit": "TOP/s", "frac": achieved / PEAK_INT32_TOPS,
                     "traffic": profiled_traffic(args.workload, int(db.n), tuning),
                     "kernel": "bits_band_kernel + bits_trace_kernel of every band class of one "
                               "step (HIP events around the band phase)",
                     "ops_per_step": OPS_PER_BAND_CELL * bcells,
                     "kernel_ms": band_ms},
        "roofline_hbm": {"achieved": hbm / (band_ms / 1e3) / 1e9, "peak": PEAK_HBM_GBS,
                         "unit": "GB/s", "algorithmic_bytes": hbm},
        "plain": plain,
        "speedup_vs_plain": (plain["ms_per_step"] / ms_per_step) if plain else None,
        "cpu_baseline": cpu,
        "host": {"generate_s": t_gen, "index_s": t_index, "filter_s": t_filter,
                 "upload_s": t_upload, "download_s": t_d2h},
        "e2e_reads_per_s_pcie": n_mapped / (t_upload + ms_per_step / 1e3 + t_d2h),
        "gpu_filter": gpu_filter,
        "records": records,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
