"""ctypes wrapper of the C oracle (oracle/nw_oracle.c).

TEST INFRASTRUCTURE ONLY -- imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker; never by saruman_amd.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
_lib = None
_vp = C.c_void_p


def load():
    global _lib
    if _lib is not None:
        return _lib
    src = HERE / "nw_oracle.c"
    if not LIB.exists() or (src.exists() and LIB.stat().st_mtime < src.stat().st_mtime):
        tmp = LIB.with_suffix(".so.tmp")
        subprocess.run(["gcc", "-O2", "-std=c11", "-fopenmp", "-shared", "-fPIC", "-o",
                        str(tmp), str(src)], check=True)
        os.replace(tmp, LIB)
    lib = C.CDLL(str(LIB))
    lib.oracle_align_one.restype = C.c_int
    lib.oracle_align_one.argtypes = [C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_int, C.c_char_p,
                                     C.c_int, C.c_int, C.c_int, _vp, _vp, C.c_int,
                                     C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    lib.oracle_matrix.restype = C.c_int
    lib.oracle_matrix.argtypes = [C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_int, C.c_char_p,
                                  C.c_int, C.c_int, _vp]
    lib.oracle_record.restype = C.c_int
    lib.oracle_record.argtypes = [C.c_char_p, C.c_uint, C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.c_int, _vp, _vp, C.c_char_p, C.c_int]
    lib.oracle_batch.restype = None
    lib.oracle_batch.argtypes = [C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp, C.c_int,
                                 _vp, C.c_int, _vp, _vp, _vp]
    lib.oracle_threads.restype = C.c_int
    _lib = lib
    return lib


@dataclass
class OneResult:
    score: int
    R: bytes
    G: bytes
    n: int
    end_row: int
    end_col: int


def align_one(match, mismatch, k, win: bytes, read: bytes, flag: int,
              stop_at_row0: bool = False) -> OneResult:
    lib = load()
    w, L = len(win), len(read)
    cap = w + k
    R = C.create_string_buffer(cap + 1)
    G = C.create_string_buffer(cap + 1)
    n, er, ec = C.c_int(), C.c_int(), C.c_int()
    sc = lib.oracle_align_one(match, mismatch, k, win, w, read, L, flag, int(stop_at_row0),
                              C.addressof(R), C.addressof(G), cap, C.byref(n), C.byref(er),
                              C.byref(ec))
    return OneResult(sc, R.raw[:cap], G.raw[:cap], n.value, er.value, ec.value)


def matrix(match, mismatch, k, win: bytes, read: bytes, flag: int) -> np.ndarray:
    lib = load()
    M = np.zeros((len(win) + 1, len(read) + 1), np.int16)
    lib.oracle_matrix(match, mismatch, k, win, len(win), read, len(read), flag, M.ctypes.data)
    return M


def record(name: str, pos: int, flag: int, k: int, L: int, w: int, score: int, R: bytes,
           G: bytes) -> str:
    lib = load()
    buf = C.create_string_buffer(8192)
    Rb = C.create_string_buffer(R, len(R))
    Gb = C.create_string_buffer(G, len(G))
    n = lib.oracle_record(name.encode(), pos & 0xFFFFFFFF, flag, k, L, w, score,
                          C.addressof(Rb), C.addressof(Gb), buf, len(buf))
    assert 0 < n < len(buf)
    return buf.value.decode()


@dataclass
class BatchResult:
    scores: np.ndarray
    records: list
    seconds: float
    threads: int


def align_batch(batch, with_records: bool = True) -> BatchResult:
    """Oracle over a saruman_amd.batch.Batch (same fields), OpenMP over candidates."""
    lib = load()
    n = batch.n
    wl = np.fromiter((len(x) for x in batch.windows), np.int64, n)
    rl = np.fromiter((len(x) for x in batch.reads), np.int64, n)
    genome = b"".join(batch.windows)
    reads = b"".join(batch.reads)
    goff = np.cumsum(wl).astype(np.int32)
    roff = np.cumsum(rl).astype(np.int32)
    ks = batch.kvec().astype(np.int32)
    flags = np.asarray(batch.flags, np.int8)
    slot = np.concatenate([[0], np.cumsum(wl + ks)]).astype(np.int64)
    total = int(slot[-1])
    R = np.zeros(max(total, 1), np.uint8)
    G = np.zeros(max(total, 1), np.uint8)
    scores = np.zeros(n, np.int32)
    gb = C.create_string_buffer(genome, len(genome) + 1)
    rb = C.create_string_buffer(reads, len(reads) + 1)
    t0 = time.perf_counter()
    lib.oracle_batch(batch.match, batch.mismatch, int(batch.k), ks.ctypes.data,
                     C.addressof(gb), C.addressof(rb), goff.ctypes.data, roff.ctypes.data, n,
                     flags.ctypes.data, int(batch.first_row0), scores.ctypes.data,
                     R.ctypes.data, G.ctypes.data)
    secs = time.perf_counter() - t0
    recs = []
    if with_records:
        for t in range(n):
            if scores[t] > ks[t]:
                recs.append("")
                continue
            s0, s1 = int(slot[t]), int(slot[t + 1])
            recs.append(record(batch.names[t], int(batch.pos[t]), int(batch.flags[t]),
                               int(ks[t]), int(rl[t]), int(wl[t]), int(scores[t]),
                               R[s0:s1].tobytes(), G[s0:s1].tobytes()))
    return BatchResult(scores, recs, secs, int(lib.oracle_threads()))
