"""Seeded random candidate batches for kernel-level tests and smoke runs.

Each candidate is a genome window of w = L + s + e bytes (s, e from the flag as in
src/cuda-functions.cu:383-410) and a read cut from it with a few random edits, so
most candidates are hits; a share are unrelated reads (non-hits) and a share use a
window longer than L + s + e, as the batch builder produces after an
out-of-genome skip (src/hash.c:982-984, Q10).
"""
from __future__ import annotations

import numpy as np

from .batch import Batch

ACGT = np.frombuffer(b"ACGT", np.uint8)


def _flag_se(flag: int, k: int):
    s = k if flag in (2, 3, 6, 7) else 0
    e = k if flag in (1, 3, 5, 7) else 0
    return s, e


def mutate(rng, seq: np.ndarray, n_err: int, p_sub=0.8, p_ins=0.1) -> np.ndarray:
    s = list(seq)
    for _ in range(n_err):
        if not s:
            break
        u = rng.random()
        i = int(rng.integers(0, len(s)))
        if u < p_sub:
            s[i] = int(ACGT[(np.searchsorted(ACGT, s[i]) + rng.integers(1, 4)) % 4])
        elif u < p_sub + p_ins:
            s.insert(i, int(ACGT[rng.integers(0, 4)]))
        else:
            del s[i]
    return np.array(s, np.uint8)


def random_batch(rng, n: int, L=36, k=2, flags=(1, 3, 5, 7), match=0, mismatch=1,
                 p_random=0.1, p_long=0.03, per_cand_k=None, edge_cases=True) -> Batch:
    """L may be an int or a (lo, hi) range; per_cand_k: callable L -> k (extended API)."""
    wins, reads, fl, ks = [], [], [], []
    for t in range(n):
        Lt = int(rng.integers(L[0], L[1] + 1)) if isinstance(L, tuple) else int(L)
        kt = int(per_cand_k(Lt)) if per_cand_k else int(k)
        f = int(flags[rng.integers(0, len(flags))])
        s, e = _flag_se(f, kt)
        w = Lt + s + e
        if rng.random() < p_long:
            w = Lt + 2 * kt                      # window longer than L + s + e (Q10)
        win = ACGT[rng.integers(0, 4, w)]
        mode = rng.random()
        if mode < p_random:
            read = ACGT[rng.integers(0, 4, Lt)]
        else:
            off = int(rng.integers(0, max(1, w - Lt + 1)))
            if edge_cases and rng.random() < 0.15:
                off = int(rng.choice([0, max(0, w - Lt)]))
            src = win[off:off + Lt + 4]
            if len(src) < Lt + 4:
                src = np.concatenate([src, ACGT[rng.integers(0, 4, Lt + 4 - len(src))]])
            nerr = int(rng.integers(0, kt + 2))
            read = mutate(rng, src, nerr)
            if edge_cases and rng.random() < 0.05:
                # leading insertions: read chars before the window start
                read = np.concatenate([ACGT[rng.integers(0, 4, int(rng.integers(1, max(kt, 1) + 1)))],
                                       read])
            if len(read) < Lt:
                read = np.concatenate([read, ACGT[rng.integers(0, 4, Lt - len(read))]])
            read = read[:Lt]
        wins.append(win.tobytes())
        reads.append(read.tobytes())
        fl.append(f)
        ks.append(kt)
    kvec = np.array(ks, np.int64)
    return Batch(match=match, mismatch=mismatch, k=int(kvec.max()) if n else int(k),
                 windows=wins, reads=reads, flags=np.array(fl, np.uint8),
                 pos=rng.integers(1, 1 << 24, n).astype(np.int64),
                 names=[f"r{t}" for t in range(n)],
                 ks=kvec if per_cand_k else None)
