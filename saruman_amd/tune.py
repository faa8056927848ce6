"""Autotuner for the banded kernels' launch shape on this GPU.

Unit costs (match 0 / mismatch 1, every SARUMAN run): the tile shape and the
launch order of the bit-parallel kernels, by coordinate descent over
  class_streams   band classes side by side on 1, 2, 4 or 8 streams
  block_waves     waves per block (1, 2, 4; one wave per tile)
  tile_cands      candidates per tile (64, or 32 with half the lanes idle)
  split_len       each band class split by read length at a quartile of the
                  batch's lengths (two launches, the short reads with less LDS)
  merge_below     narrow classes run on a wider class's kernel (fewer launches)
  walk_layout     the forward pass stores the D bits for the walk (0), or every
                  32nd column's state from which the walk recomputes (1)
each point timed as the median of whole band phases (HIP events).
Other costs: the persistent nibble kernel with 1..max resident blocks per CU per
band class, then merging and streams.  The result is kept only if it beats the
default launch by 2% on a fresh measurement of both; the table -- with every
point of the search -- is written to saruman_amd/tuning/<arch>.json and applied
with srm_ctx_set_tuning().

    python -m saruman_amd.tune [--reads N]     (on the GPU; default: the 2M-read mixed workload)
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import time
from pathlib import Path

import numpy as np

from . import _lib

TUNING_DIR = Path(__file__).resolve().parent / "tuning"
ARCH = "gfx950"


class Tuning(C.Structure):
    _fields_ = [("blocks_per_cu", C.c_int * 9), ("merge_below", C.c_int),
                ("class_streams", C.c_int), ("tile_cands", C.c_int), ("block_waves", C.c_int),
                ("split_len", C.c_int), ("walk_layout", C.c_int)]


KNOBS = ("merge_below", "class_streams", "tile_cands", "block_waves", "split_len", "walk_layout")


def defaults() -> dict:
    return {"blocks_per_cu": [0] * 9, **{k: 0 for k in KNOBS}}


def table_path(arch: str = ARCH) -> Path:
    return TUNING_DIR / f"{arch}.json"


def load_table(arch: str = ARCH):
    p = table_path(arch)
    if not p.exists():
        return None
    return json.loads(p.read_text())


def apply(ctx, table) -> None:
    """Applies a tuning table (dict) to a saruman_amd.Context; None = defaults."""
    lib = _lib.load()
    if table is None:
        _lib.check(lib.srm_ctx_set_tuning(ctx.h, None), "srm_ctx_set_tuning")
        return
    t = Tuning()
    for i, b in enumerate(table["blocks_per_cu"][:9]):
        t.blocks_per_cu[i] = int(b)
    for k in KNOBS:
        setattr(t, k, int(table.get(k, 0)))
    _lib.check(lib.srm_ctx_set_tuning(ctx.h, C.byref(t)), "srm_ctx_set_tuning")


def apply_cached(ctx, arch: str = ARCH):
    table = load_table(arch)
    if table is not None:
        apply(ctx, table)
    return table


def _class_times(job, reps):
    # per-class times are measured with the classes one after another (class_streams 1)
    lib = _lib.load()
    per = {nw: [] for nw in range(1, 9)}
    totals = []
    for _ in range(reps):
        job.run("banded")
        b_ms, _ = job.kernel_ms()
        totals.append(b_ms)
        for nw in range(1, 9):
            blocks, ntiles, ms = C.c_int(), C.c_int(), C.c_float()
            lib.srm_job_band_info(job.h, nw, C.byref(blocks), C.byref(ntiles), C.byref(ms))
            if ntiles.value:
                per[nw].append(ms.value)
    return float(np.median(totals)), {nw: float(np.median(v)) for nw, v in per.items() if v}


def _total(job, reps):
    """Median device time of the whole step (HIP events around all launches)."""
    return float(np.median([job.run("banded") for _ in range(reps)]))


def _search_unit(ctx, job, upload, reps, log, lengths):
    """Coordinate descent over the unit-cost launch knobs; returns (best table,
    every point measured)."""
    q = [int(x) for x in np.quantile(lengths, (0.25, 0.5, 0.75))] if len(lengths) else []
    space = [("walk_layout", [0, 1]), ("class_streams", [4, 1, 2, 8]), ("block_waves", [4, 2, 1]),
             ("tile_cands", [64, 32]), ("split_len", [0] + sorted(set(q))),
             ("merge_below", [0, 2, 3])]
    best = defaults()
    points = []

    def measure(t):
        apply(ctx, t)
        upload(job)
        ms = _total(job, reps)
        points.append({**{k: t[k] for k in KNOBS}, "band_ms": ms})
        log(f"[tune] {' '.join(f'{k}={t[k]}' for k in KNOBS)}: {ms:.3f} ms")
        return ms

    best_ms = measure(best)
    for _ in range(2):                       # two sweeps over the knobs
        changed = False
        for knob, values in space:
            for v in values:
                if v == best[knob] or (knob == "class_streams" and v == 4 and best[knob] == 0):
                    continue
                t = {**best, knob: v}
                ms = measure(t)
                if ms < best_ms:
                    best, best_ms, changed = t, ms, True
        if not changed:
            break
    return best, points


def tune(ctx, upload, reps: int = 5, log=print, unit: bool = True, lengths=()):
    """upload(job) uploads the representative candidate batch into a Job; unit:
    the batch runs match 0 / mismatch 1; lengths: its read lengths (split points)."""
    from .batch import Job
    lib = _lib.load()
    job = Job(ctx)
    try:
        points = []
        if unit:
            cand, points = _search_unit(ctx, job, upload, reps, log, np.asarray(lengths))
        else:
            apply(ctx, {**defaults(), "class_streams": 1})
            upload(job)
            base_total, base = _class_times(job, reps)
            log(f"[tune] default launch: band kernels {base_total:.3f} ms "
                + " ".join(f"nw{nw}={ms:.3f}" for nw, ms in base.items()))
            occ = {nw: lib.srm_band_occupancy(nw, 0) for nw in range(1, 9)}
            best = {nw: (ms, 0) for nw, ms in base.items()}
            for b in range(1, max(occ.values()) + 1):
                t = {**defaults(), "blocks_per_cu": [0] + [min(b, occ[nw]) for nw in range(1, 9)],
                     "class_streams": 1}
                apply(ctx, t)
                upload(job)
                tot, per = _class_times(job, reps)
                log(f"[tune] {b} block(s)/CU: {tot:.3f} ms " +
                    " ".join(f"nw{nw}={ms:.3f}" for nw, ms in per.items()))
                for nw, ms in per.items():
                    if b <= occ[nw] and ms < best.get(nw, (1e30, 0))[0]:
                        best[nw] = (ms, b)
            blocks = [0] + [best.get(nw, (0, 0))[1] for nw in range(1, 9)]
            results = {}
            for mb in (0, 2, 3):
                for cs in (1, 2, 4):
                    t = {**defaults(), "blocks_per_cu": blocks, "merge_below": mb,
                         "class_streams": cs}
                    apply(ctx, t)
                    upload(job)
                    results[(mb, cs)] = _total(job, reps)
                    log(f"[tune] merge_below={mb} class_streams={cs}: {results[(mb, cs)]:.3f} ms")
            mb, cs = min(results, key=results.get)
            cand = {**defaults(), "blocks_per_cu": blocks, "merge_below": mb, "class_streams": cs}
        # keep the default launch unless the tuned one is at least 2% faster on a
        # fresh measurement of both
        apply(ctx, None)
        upload(job)
        d_tot = _total(job, 2 * reps)
        apply(ctx, cand)
        upload(job)
        t_tot = _total(job, 2 * reps)
        log(f"[tune] default {d_tot:.3f} ms vs tuned {t_tot:.3f} ms")
        if t_tot > 0.98 * d_tot:
            cand = defaults()
        table = {"arch": ARCH, **cand, "band_ms_default": d_tot, "band_ms_tuned": t_tot,
                 "kept": "tuned" if cand != defaults() else "defaults",
                 "tuned_at": time.strftime("%Y-%m-%d %H:%M:%S"), "search": points}
        apply(ctx, table)
        return table
    finally:
        job.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reads", type=int, default=2_000_000)
    ap.add_argument("--workload", default="mixed")
    ap.add_argument("--out", default=str(table_path()))
    args = ap.parse_args()
    from . import workload as wl
    from .batch import Context
    cfg, genome, reads = wl.make_workload(args.workload, n_reads=args.reads)
    index = wl.Index(genome, wl.needed_qs(reads))
    strands = [wl.run_filter(index, genome, reads, dr, cfg["chunk"]) for dr in (0, 4)]
    db = wl.device_batch(reads, strands)
    ctx = Context(0)
    ctx.set_genome(wl.genome_words(genome))

    def upload(job):
        job.upload(0, 1, db.cand, db.read_words, None, ops_total=db.ops_total)

    table = tune(ctx, upload, lengths=db.L)
    table["workload"] = f"{args.workload}, {args.reads} reads, {db.n} candidates"
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    Path(args.out).write_text(json.dumps(table, indent=1) + "\n")
    print(json.dumps(table), flush=True)


if __name__ == "__main__":
    main()
