"""Pins the C oracle (oracle/nw_oracle.c) without the reference binary.

The reference cannot be built here (nvcc, CUDA samples headers and argtable2 are
absent) and ships no test vectors, so the oracle is checked against what the
mathematics fixes and against traces worked by hand from
src/cuda-functions.cu:317-536 and :828-994:
  * flags 0/4 (s = e = 0), match 0 / mismatch 1: textbook Levenshtein distance;
  * flags 1-7: the semiglobal closed form
        min_{0<=a<=b<=w} max(0, a-2s) + lev(read, g[a:b]) + max(0, (w-2e) - b);
  * any (match, mismatch): minimum over explicitly enumerated lattice paths;
  * transcript invariants (rows are a read suffix / window suffix, equal length);
  * hand-worked traces and records: exact hit, substitution, the Q4 column-0
    start clause, the Q1 first-candidate stop at row 0.
"""
import functools
import itertools

import numpy as np
import pytest

ACGT = b"ACGT"


def lev(a: bytes, b: bytes) -> int:
    """Textbook Levenshtein distance (Wagner-Fischer, written from the definition)."""
    prev = list(range(len(b) + 1))
    for i in range(1, len(a) + 1):
        cur = [i] + [0] * len(b)
        for j in range(1, len(b) + 1):
            cur[j] = min(prev[j] + 1, cur[j - 1] + 1, prev[j - 1] + (a[i - 1] != b[j - 1]))
        prev = cur
    return prev[-1]


def flag_se(flag, k):
    s = k if flag in (2, 3, 6, 7) else 0
    e = k if flag in (1, 3, 5, 7) else 0
    return s, e


def rand_seq(rng, n):
    return bytes(ACGT[i] for i in rng.integers(0, 4, n))


def test_levenshtein_for_unflagged(orc):
    rng = np.random.default_rng(1)
    for _ in range(400):
        L = int(rng.integers(1, 9))
        w = int(rng.integers(1, 11))
        read, win = rand_seq(rng, L), rand_seq(rng, w)
        for flag in (0, 4):
            M = orc.matrix(0, 1, 3, win, read, flag)
            assert int(M[-1, -1]) == lev(read, win)


def test_semiglobal_closed_form(orc):
    rng = np.random.default_rng(2)
    for _ in range(300):
        k = int(rng.integers(0, 4))
        L = int(rng.integers(1, 8))
        flag = int(rng.choice([1, 2, 3, 5, 6, 7]))
        s, e = flag_se(flag, k)
        w = L + s + e + int(rng.integers(-1, 2))
        if w < 1:
            continue
        read, win = rand_seq(rng, L), rand_seq(rng, w)
        if rng.random() < 0.5:   # make hits likely
            off = int(rng.integers(0, max(1, w - L + 1)))
            read = (win[off:off + L] + read)[:L]
        want = min(max(0, a - 2 * s) + lev(read, win[a:b]) + max(0, (w - 2 * e) - b)
                   for a in range(w + 1) for b in range(a, w + 1))
        got = orc.matrix(0, 1, k, win, read, flag)[-1, -1]
        assert int(got) == want, (read, win, flag, k)


def brute_force_score(ma, mm, k, win: bytes, read: bytes, flag: int) -> int:
    """Minimum cost over explicitly enumerated monotone lattice paths.

    Cells (h, c), h in 0..w (window rows), c in 0..L (read columns).  Boundary cells
    carry their initial value (row 0: c; column 0: h for h > 2s counted in unit
    steps from 0 at h = 2s, 0 above) and are never re-entered.  A path starts at a
    boundary cell and every later cell is interior.  Move costs: diagonal into
    (h+1, c+1): match/mismatch of win[h] vs read[c]; right: mismatch; down:
    mismatch, except down moves inside the last column from a row h >= w - 2e,
    which are free.
    """
    s, e = flag_se(flag, k)
    w, L = len(win), len(read)

    def init(h, c):
        if h == 0:
            return c
        return 0 if h <= 2 * s else h - 2 * s

    best = None

    def walk(h, c, cost):
        nonlocal best
        if (h, c) == (w, L):
            best = cost if best is None else min(best, cost)
            return
        if h < w and c < L:
            walk(h + 1, c + 1, cost + (mm if win[h] != read[c] else ma))
        if c < L and h >= 1:
            walk(h, c + 1, cost + mm)
        if h < w and c >= 1:
            walk(h + 1, c, cost + (0 if (c == L and h >= w - 2 * e) else mm))

    for h in range(w + 1):
        for c in range(L + 1):
            if h == 0 or c == 0:
                st = init(h, c)
                if (h, c) == (w, L):
                    best = st if best is None else min(best, st)
                if h < w and c < L:
                    walk(h + 1, c + 1, st + (mm if win[h] != read[c] else ma))
                if h >= 1 and c < L:
                    walk(h, c + 1, st + mm)
                if c >= 1 and h < w:
                    walk(h + 1, c, st + (0 if (c == L and h >= w - 2 * e) else mm))
    return best


@pytest.mark.parametrize("ma,mm", [(0, 1), (0, 2), (1, 3), (0, 3), (2, 5)])
def test_lattice_paths(orc, ma, mm):
    rng = np.random.default_rng(100 + 7 * ma + mm)
    for _ in range(60):
        k = int(rng.integers(0, 3))
        L = int(rng.integers(1, 5))
        flag = int(rng.choice([0, 1, 2, 3, 4, 5, 6, 7]))
        s, e = flag_se(flag, k)
        w = max(1, L + s + e + int(rng.integers(-1, 2)))
        read, win = rand_seq(rng, L), rand_seq(rng, w)
        if rng.random() < 0.5:
            off = int(rng.integers(0, max(1, w - L + 1)))
            read = (win[off:off + L] + read)[:L]
        got = int(orc.matrix(ma, mm, k, win, read, flag)[-1, -1])
        assert got == brute_force_score(ma, mm, k, win, read, flag), (ma, mm, read, win, flag, k)


def test_transcript_invariants(orc):
    rng = np.random.default_rng(3)
    nhits = 0
    for _ in range(1500):
        k = int(rng.integers(1, 5))
        L = int(rng.integers(k * 3 + 1, 40))
        flag = int(rng.choice([1, 3, 5, 7]))
        s, e = flag_se(flag, k)
        w = L + s + e
        win = rand_seq(rng, w)
        off = int(rng.integers(0, w - L + 1))
        src = bytearray(win[off:off + L])
        for _ in range(int(rng.integers(0, k + 1))):
            i = int(rng.integers(0, len(src)))
            u = rng.random()
            if u < 0.6:
                src[i] = ACGT[(ACGT.index(src[i]) + 1) % 4]
            elif u < 0.8:
                src.insert(i, ACGT[int(rng.integers(0, 4))])
            else:
                del src[i]
        read = bytes((src + b"A" * L)[:L])
        r = orc.align_one(0, 1, k, win, read, flag)
        if r.score > k:
            continue
        nhits += 1
        R, G = r.R[:r.n], r.G[:r.n]
        assert r.n <= w + k
        assert all(x != 0 for x in R + G)
        assert not any(a == ord("_") and b == ord("_") for a, b in zip(R, G))
        assert R.replace(b"_", b"") == read[r.end_col:]
        assert G.replace(b"_", b"") == win[r.end_row:]
        # a normal walk ends on the free start region of column 0 or at (0, 0)
        assert r.end_col == 0 and r.end_row <= 2 * s + 3
    assert nhits > 500


def test_hand_exact_match_flag1(orc):
    # traced by hand: diagonal matches, then two free verticals in column L
    r = orc.align_one(0, 1, 2, b"ACGTAA", b"ACGT", 1)
    assert r.score == 0
    assert r.R == b"ACGT__\0\0" and r.G == b"ACGTAA\0\0"
    rec = orc.record("r0", 1000, 1, 2, 4, 6, r.score, r.R, r.G)
    assert rec == "r0\t1000\t1003\t>>\tACGT\tACGT\t0\n"


def test_hand_substitution(orc):
    # M worked by hand; the walk takes the mismatch branch at (3, 3)
    r = orc.align_one(0, 1, 2, b"ACGTAA", b"ACTT", 1)
    assert r.score == 1
    assert r.R == b"ACTT__\0\0" and r.G == b"ACGTAA\0\0"
    rec = orc.record("q", 7, 5, 2, 4, 6, r.score, r.R, r.G)
    assert rec == "q\t7\t10\t<<\tACTT\tACGT\t1\n"


def test_hand_q4_start_clause(orc):
    # flag 3, k = 1: the walk reaches (1, 0) and moves up for free because of the
    # start clause of the traceback's correction (cuda-functions.cu:182), which
    # the forward pass does not have
    r = orc.align_one(0, 1, 1, b"ACGTA", b"CGT", 3)
    assert r.score == 0
    assert r.R == b"_CGT_\0" and r.G == b"ACGTA\0"
    assert (r.end_row, r.end_col) == (0, 0)
    rec = orc.record("x", 50, 3, 1, 3, 5, r.score, r.R, r.G)
    assert rec == "x\t51\t53\t>>\tCGT\tCGT\t0\n"


def test_hand_row0_insertion_and_q1(orc):
    # a leading read character inserted before the window: canonical walks
    # continue along row 0 (every candidate but the first of a launch)
    r = orc.align_one(0, 1, 1, b"ACGAA", b"TACG", 1)
    assert r.score == 1
    assert r.R == b"TACG__" and r.G == b"_ACGAA"
    assert orc.record("t", 10, 1, 1, 4, 5, 1, r.R, r.G) == "t\t10\t12\t>>\tTACG\t_ACG\t1\n"
    # the first candidate of a launch stops at row 0 (cuda-functions.cu:129)
    r0 = orc.align_one(0, 1, 1, b"ACGAA", b"TACG", 1, stop_at_row0=True)
    assert r0.R == b"ACG__\0" and r0.G == b"ACGAA\0"
    assert orc.record("t", 10, 1, 1, 4, 5, 1, r0.R, r0.G) == "t\t10\t13\t>>\tACG_\tACGA\t1\n"


def test_hand_q5_column0_stop(orc):
    """Q5, worked by hand (cuda-functions.cu:181-190, 239-276).  Flag 3, k = 2, so
    s = e = 2 and column 0 holds max(0, h - 4).  The read GATTACAC sits at window
    offset 5 with its fourth base inserted: score 1 (column 0, row 5) + 1 = 2.
    Walk from (12, 8): five diagonal matches down to (8, 4), the left step of the
    inserted T at (7, 3) (top gives 3 + 1 != 2), diagonals to (6, 1) and (5, 0).
    At (5, 0): diag/left read 999, top = 0, and the start clause (column <= 1,
    row - 1 = 4 <= 2(s + 1)) makes the top cost 0, so 0 != 1 and the walk stops.
    The record then reports the window start (200), not the alignment start 205,
    and stop = 8 chars - 1 insertion + 200 - 1 = 206."""
    r = orc.align_one(0, 1, 2, b"CCCCCGATACAC", b"GATTACAC", 3)
    assert r.score == 2
    assert r.R == b"GATTACAC" + b"\0" * 6 and r.G == b"GA_TACAC" + b"\0" * 6
    assert (r.end_row, r.end_col) == (5, 0)
    rec = orc.record("z", 200, 3, 2, 8, 12, r.score, r.R, r.G)
    assert rec == "z\t200\t206\t>>\tGATTACAC\tGA_TACAC\t2\n"


def test_record_end_of_read_branch(orc):
    """The eor > 0 branch (cuda-functions.cu:974-976), worked by hand on a slot:
    w = 6, k = 1, R = "AC_G" then NUL: the scan bound w - stc - sc + z exceeds
    the emitted length, the NUL at i = 4 ends the loop, chars = 4 - 0."""
    R = b"AC_G\0\0\0"
    G = b"ACTG\0\0\0"
    rec = orc.record("e", 3, 1, 1, 3, 6, 1, R, G)
    assert rec == "e\t3\t6\t>>\tAC_G\tACTG\t1\n"
