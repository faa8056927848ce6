"""Hand-worked cases of SARUMAN's candidate dedup and pairing, pinned by literals.

Every expected value below was worked by hand from the reference text:
add_start_position (src/hash.c:40-178) and run_seashore's per-read body
(src/hash.c:502-834). None was produced by running any implementation. Each case
runs through three implementations, which must all give the literal:
  - the oracle: oracle/filter_oracle.py (Strand.add_start_position, seashore_read);
  - the host filter's add_start and seashore_read (saruman_amd/csrc/filter.cpp);
  - the device filter's add_start and seashore (saruman_amd/csrc/filter_pair.h,
    the per-read code the kernels of filter_gpu.hip run).
The two product implementations are compiled for the host, with AddressSanitizer
and UBSan, by tests/cpu_build/seashore_cases.cpp.

The seashore cases give each read's q-gram position lists directly: the
data_array of hash.c:396-461, each list 0-terminated, the lists of a chunk of
chunk_value reads back to back. Then the Q12 scan, which starts list b at a's index
(hash.c:628, 688-689), can be aimed at a chosen element. Lists hold positions as
the index stores them: the 0-based genome offset of the q-gram. The genome is only
read by the perfect check.
"""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

from oracle import filter_oracle as fo

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "tests" / "cpu_build" / "seashore_cases.cpp"


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("cpu_build") / "seashore_cases"
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer",
           "-I", str(ROOT / "include"), "-I", str(ROOT / "saruman_amd" / "csrc"), str(SRC),
           "-lpthread", "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def _run(exe, text):
    r = subprocess.run([str(exe)], input=text, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout.splitlines()


def _fields(line):
    head, *kv = line.split()
    d = {"impl": head}
    for item in kv:
        key, _, val = item.partition("=")
        d[key] = val
    return d


def _ents(val):
    return [tuple(int(x) for x in e.split(":")) for e in val.split(",") if e]


def _ints(val):
    return [int(x) for x in val.split(",") if x]


# --------------------------------------------------------------------------------
# add_start_position (hash.c:40-178): (k, exact_tresh, calls (pos, flag, possible),
# return codes, final entries (pos, flag) in list order, queue count).  Return
# codes: 0 first entry, 1 rejected, 2 appended or flag lowered.
DEDUP = {
    # rule (i), hash.c:118-127: reject when pos + i hits an entry for some |i| <= k
    # and the entry's flag <= the new flag.  k = 2: 102 and 98 are within 2 of 100;
    # 103 and 97 are k + 1 away and are appended.
    "reject_within_k": (2, 2, [(100, 1, 2), (102, 1, 2), (98, 1, 2), (103, 1, 2), (97, 1, 2)],
                        [0, 1, 1, 2, 2], [(100, 1), (103, 1), (97, 1)], 3),
    # rule (ii), hash.c:130-136: same position and strand, entry flag above the new
    # one: the flag is lowered, return 2.  The queue does not count a lowering.
    "lower_flag": (2, 2, [(100, 3, 2), (100, 1, 2)], [0, 2], [(100, 1)], 1),
    # rule (i) needs flag(entry) <= flag: 3 <= 1 fails, and rule (ii) needs the
    # same position, so a lower flag one base away is appended.
    "lower_flag_elsewhere_appends": (2, 2, [(100, 3, 2), (101, 1, 2)], [0, 2],
                                     [(100, 3), (101, 1)], 2),
    # equal flags at the same position: rule (i) with i = 0
    "same_pos_same_flag": (2, 2, [(100, 3, 2), (100, 3, 2)], [0, 1], [(100, 3)], 1),
    # a higher flag at an existing position: rule (i) (1 <= 3)
    "same_pos_higher_flag": (2, 2, [(100, 1, 2), (100, 3, 2)], [0, 1], [(100, 1)], 1),
    # strands: flag 5 is reverse.  (100, 1) after (100, 5): rule (i) fails (5 <= 1
    # is false), rules (ii)/(iii) need the same strand -> appended.  (100, 7):
    # rule (i) on (100, 5) (5 <= 7) -> rejected.  (100, 3): rule (i) fails on
    # (100, 5), (ii)/(iii) strand differs; rule (i) on (100, 1) (1 <= 3) -> rejected.
    "strands": (2, 2, [(100, 5, 2), (100, 1, 2), (100, 7, 2), (100, 3, 2)], [0, 2, 1, 1],
                [(100, 5), (100, 1)], 2),
    # a perfect hit's entry (-S, flag 0) suppresses every later candidate within
    # +-k of S (pos + i == -p), any flag, either strand; S +- (k+1) is appended.
    # The perfect entry itself is not queued (pos <= 0, hash.c:87-89).
    "perfect_suppresses": (2, 2, [(-100, 0, 2), (102, 1, 2), (98, 5, 2), (103, 1, 2), (97, 5, 2)],
                           [0, 1, 1, 2, 2], [(-100, 0), (103, 1), (97, 5)], 2),
    # rule (iii), hash.c:138-144: an entry already queued (negated, -100 flag 7)
    # rejects the same position on the same strand; on the other strand it does not.
    "queued_same_strand": (2, 2, [(-100, 7, 2), (100, 5, 2)], [0, 1], [(-100, 7)], 0),
    "queued_other_strand": (2, 2, [(-100, 7, 2), (100, 1, 2)], [0, 2], [(-100, 7), (100, 1)], 1),
    # a perfect hit found after a candidate at the same S: (i) fails (1 <= 0 is
    # false), (ii) needs p == pos (100 != -100), (iii) p == -pos on the same strand
    # -> rejected, so it is not printed.
    "perfect_after_candidate": (2, 2, [(100, 1, 2), (-100, 0, 2)], [0, 1], [(100, 1)], 1),
    # entries are tried in list order and the first one a rule applies to decides:
    # (100, 3) is lowered by rule (ii) before (101, 1), which would reject (100, 1)
    # by rule (i), is reached.
    "first_entry_decides": (2, 2, [(100, 3, 2), (101, 1, 2), (100, 1, 2)], [0, 2, 2],
                            [(100, 1), (101, 1)], 2),
    # the gate, hash.c:46: nothing happens unless possible == exact_tresh
    "possible_gate": (2, 2, [(100, 1, 1), (100, 1, 3), (200, 1, 2)], [1, 1, 0], [(200, 1)], 1),
    # k = 0: the window is the position itself
    "k_zero": (0, 2, [(100, 1, 2), (101, 1, 2), (100, 1, 2)], [0, 2, 1], [(100, 1), (101, 1)], 2),
}


def _oracle_dedup(k, exact, calls):
    st = fo.Strand()
    rets = [st.add_start_position(0, p, f, poss, exact, k) for p, f, poss in calls]
    ent = st.results.get(0)
    ents = list(zip(ent.positions, ent.flags)) if ent else []
    return rets, ents, st.queue


@pytest.mark.parametrize("name", sorted(DEDUP))
def test_dedup_hand_worked(exe, name):
    k, exact, calls, rets, ents, queued = DEDUP[name]
    assert _oracle_dedup(k, exact, calls) == (rets, ents, queued), "oracle"
    text = f"dedup {k} {exact} {len(calls)}\n" + "".join(f"{p} {f} {q}\n" for p, f, q in calls)
    lines = _run(exe, text)
    assert len(lines) == 2
    for line in lines:
        d = _fields(line)
        got = (_ints(d["rets"]), _ents(d["ent"]), int(d["queued"]))
        assert got == (rets, ents, queued), d["impl"]


# --------------------------------------------------------------------------------
# run_seashore for one read over given position lists.  L 36, k 2 gives q 11,
# rest 3, n 6 lists (prefix q-grams at read offsets 0, 11, 22, suffix q-grams at
# 25, 14, 3), half 3, exact_tresh 2 (saruman.c:42-75, :119).  For a prefix list a
# and its element P at index `off`: S = P - 11a is the read start, gs = S - 1.
#   first pass, list b > a: match iff P_b == gs + 11b + 1 + err, err in [-k, k];
#     the scan reads D(b, off), then D(b, 1), D(b, 2), ... up to a 0 (Q12).
#   second pass, list 3 + j: match iff P_b == gs + (2 - j)*11 + 1 + 3 + aerr,
#     |aerr| <= |2 - j - a|.
#   flags: a = 0 gives (S, 1 + dr); a > 0 gives (S - k, 3 + dr) (hash.c:636-686,
#   :740-790); only the exact_tresh-th match of one (a, P) is added (Q13).
# A genome of A's and reads of C's has no perfect hit.
GENOME_A = "A" * 1000
READ_C36 = "C" * 36

SEASHORE = {
    # Q12, hash.c:628: list 0 = [50, 120, 200]; P = 200 sits at index 2.  List 1
    # holds one element, so D(1, 2) is past its terminator: index 0 of list 2,
    # 212 = gs + 12 + 1 (err = +1) -> (200, 1) is added.  The scan then reads
    # D(1, 1) = 0 and stops.  P = 50 and 120 and the lists 1, 2 match nothing
    # (P = 500 of list 1 looks for 511 +- 2 in list 2).
    "q12_first_element_from_next_list": dict(
        L=36, k=2, dr=0, cv=1, genome=GENOME_A, reads=[READ_C36],
        lists=[[[50, 120, 200], [500], [212, 900], [], [], []]],
        want=[([(200, 1)], [], 1)]),
    # Q12 continued: list 0 = [40, 200] (P = 200 at index 1), list 1 empty, so
    # D(1, 1), D(1, 2), ... run through list 2 = [150, 212]: 212 (err +1) is found
    # at D(1, 2), after the first element read twice -> (200, 1).
    "q12_scan_runs_through_next_list": dict(
        L=36, k=2, dr=0, cv=1, genome=GENOME_A, reads=[READ_C36],
        lists=[[[40, 200], [], [150, 212], [], [], []]],
        want=[([(200, 1)], [], 1)]),
    # Q12 across reads: read 0's list 0 = [10, 20, 30, 40, 300] (P = 300 at index
    # 4), its other lists empty.  For b = 2, D(2, 4) runs past read 0's lists 2-5
    # into read 1's list 0 = [322] = gs + 23 + 0 -> (300, 1), but only when both
    # reads share a chunk (chunk_value 2): with chunk_value 1 each chunk has its own
    # zero-padded data_array and the read is 0 -> nothing.
    "q12_into_next_read_same_chunk": dict(
        L=36, k=2, dr=0, cv=2, genome=GENOME_A, reads=[READ_C36, READ_C36],
        lists=[[[10, 20, 30, 40, 300], [], [], [], [], []], [[322], [], [], [], [], []]],
        want=[([(300, 1)], [], 1), ([], [], 0)]),
    "q12_not_into_next_chunk": dict(
        L=36, k=2, dr=0, cv=1, genome=GENOME_A, reads=[READ_C36, READ_C36],
        lists=[[[10, 20, 30, 40, 300], [], [], [], [], []], [[322], [], [], [], [], []]],
        want=[([], [], 0), ([], [], 0)]),
    # Perfect hits, forward strand.  The read sits at genome offset 100.  (a=0,
    # P=100): the perfect check adds (-100, 0) and prints 100; its pairs with lists
    # 1-5 (111, 122, 125, 114, 103 are each an exact match) give (100, 1), rejected
    # by rule (i) against -100 (or by the gate after the first).  (a=1, P=111) and
    # (a=2, P=122) are the same S = 100: the perfect add is rejected (not printed
    # again) and their (98, 3) candidates fall within k of 100.
    "perfect_forward": dict(
        L=36, k=2, dr=0, cv=1,
        genome="A" * 100 + "ACGTTGCAACGGTCAGTCCATGAGATCCTAGGACTA" + "A" * 64,
        reads=["ACGTTGCAACGGTCAGTCCATGAGATCCTAGGACTA"],
        lists=[[[100], [111], [122], [125], [114], [103]]],
        want=[([(-100, 0)], [100], 0)]),
    # The same on the reverse strand (do_reverse 4): the strand sees the reverse
    # complement, which the genome holds at 100.  The perfect entry keeps flag 0
    # (Q16, hash.c:574); the candidates would be (100, 5) and (98, 7).
    "perfect_reverse": dict(
        L=36, k=2, dr=4, cv=1,
        genome="A" * 100 + "ACGTTGCAACGGTCAGTCCATGAGATCCTAGGACTA" + "A" * 64,
        reads=["TAGTCCTAGGATCTCATGGACTGACCGTTGCAACGT"],
        lists=[[[100], [111], [122], [125], [114], [103]]],
        want=[([(-100, 0)], [100], 0)]),
    # Without the perfect copy the same lists give (a=0, P=100): (100, 1) at the
    # first pair.  (a=1, P=111) and (a=2, P=122) give (98, 3), rejected by rule (i)
    # against (100, 1): 98 + 2 = 100 and 1 <= 3.  So only (100, 1).
    "pairs_without_perfect": dict(
        L=36, k=2, dr=0, cv=1, genome=GENOME_A, reads=[READ_C36],
        lists=[[[100], [111], [122], [125], [114], [103]]],
        want=[([(100, 1)], [], 1)]),
    # a > 0 alone, reverse strand: list 0 empty, (a=1, P=111): S = 100, pair with
    # list 2's 122 -> (98, 3 + 4); its second-pass pair with 125 is past the gate.
    # (a=2, P=122): second pass with list 3's 125 (j = 0: 99 + 26 + 0) -> (98, 7)
    # again, rejected by rule (i) (7 <= 7).
    "start_free_candidate": dict(
        L=36, k=2, dr=4, cv=1, genome=GENOME_A, reads=[READ_C36],
        lists=[[[], [111], [122], [125], [], []]],
        want=[([(98, 7)], [], 1)]),
    # Q13 with a flag that differs.  L 30, k 4: q 5, rest 0, so n = 6 lists of the
    # prefix q-grams at 0, 5, ..., 25, half 3, exact_tresh floor(30/5) - 4 + 1 = 3.
    # (a=1, P=105): S = 100, gs = 99.  First pass, list 2 = [110, 111] against
    # gs + 11 + err: 110 at err 0 is the 2nd match (possible 2, below the gate),
    # 111 at err +1 the 3rd: (96, 3) is added.  Second pass j = 0, list 3 = [110]
    # against gs + 11 + aerr: a 4th match, with flag 2 (nin + q = 16 <= 30,
    # hash.c:748), past the gate: rejected, although rule (ii) would have lowered
    # (96, 3) to flag 2.  (a=2, P=110): one second-pass match (possible 2): nothing.
    "q13_later_pair_other_flag": dict(
        L=30, k=4, dr=0, cv=1, genome="A" * 400, reads=["C" * 30],
        lists=[[[], [105], [110, 111], [110], [], []]],
        want=[([(96, 3)], [], 1)]),
    # Q13 with exact_tresh 3: a single matching pair adds nothing.  (a=1, P=105)
    # has only 110 in list 2 (2nd match) -> no candidate.
    "q13_single_pair_below_threshold": dict(
        L=30, k=4, dr=0, cv=1, genome="A" * 400, reads=["C" * 30],
        lists=[[[], [105], [110], [], [], []]],
        want=[([], [], 0)]),
}


def _oracle_seashore(c):
    L, k, dr, cv = c["L"], c["k"], c["dr"], c["cv"]
    P = fo.derive_params(L, k)
    genome = c["genome"].encode()
    out = []
    for r, rd in enumerate(c["reads"]):
        lo = (r // cv) * cv
        data, offs = [], []
        for rr in range(lo, min(lo + cv, len(c["reads"]))):
            for lst in c["lists"][rr]:
                offs.append(len(data))
                data.extend(lst)
                data.append(0)
        tmp = rd.encode()
        if dr == 4:
            tmp = fo.reverse_complement(tmp)
        st = fo.Strand()
        printed = fo.seashore_read(st, data, offs, (r - lo) * P.n, r, tmp, genome, P, dr)
        ent = st.results.get(r)
        ents = list(zip(ent.positions, ent.flags)) if ent else []
        out.append((ents, printed, st.queue))
    return out


def _harness_input(c):
    P = fo.derive_params(c["L"], c["k"])
    parts = [f"seashore {c['L']} {c['k']} {c['dr']} {len(c['reads'])} {c['cv']}",
             f"{len(c['genome'])} {c['genome']}"]
    parts += c["reads"]
    for lists in c["lists"]:
        assert len(lists) == P.n
        for lst in lists:
            parts.append(" ".join(str(x) for x in [len(lst), *lst]))
    return "\n".join(parts) + "\n"


def test_case_parameters():
    """The derived parameters the seashore cases are worked with (saruman.c:42-119)."""
    p = fo.derive_params(36, 2)
    assert (p.q, p.rest, p.n, p.exact_tresh, p.positions[:p.n]) == (11, 3, 6, 2, [0, 11, 22, 25, 14, 3])
    p = fo.derive_params(30, 4)
    assert (p.q, p.rest, p.n, p.exact_tresh, p.positions[:p.n]) == (5, 0, 6, 3, [0, 5, 10, 15, 20, 25])


@pytest.mark.parametrize("name", sorted(SEASHORE))
def test_seashore_hand_worked(exe, name):
    c = SEASHORE[name]
    assert _oracle_seashore(c) == c["want"], "oracle"
    lines = _run(exe, _harness_input(c))
    assert len(lines) == 2 * len(c["reads"])
    for line in lines:
        d = _fields(line)
        r = int(d["read"])
        got = (_ents(d["ent"]), _ints(d["perf"]), int(d["queued"]))
        assert got == c["want"][r], (d["impl"], r)


def test_planted_mistakes_are_caught(monkeypatch):
    """The literals discriminate: an oracle with the gate of hash.c:46 read as
    possible >= exact_tresh, or with the rules of hash.c:118-144 applied rule by
    rule over all entries instead of entry by entry, fails cases above."""
    real = fo.Strand.add_start_position

    def gate_ge(self, read_id, pos, flag, possible, exact_tresh, k):
        return real(self, read_id, pos, flag, exact_tresh if possible >= exact_tresh else possible,
                    exact_tresh, k)

    monkeypatch.setattr(fo.Strand, "add_start_position", gate_ge)
    assert _oracle_seashore(SEASHORE["q13_later_pair_other_flag"]) != \
        SEASHORE["q13_later_pair_other_flag"]["want"]
    monkeypatch.setattr(fo.Strand, "add_start_position", real)

    def rule_by_rule(self, read_id, pos, flag, possible, exact_tresh, k):
        if possible != exact_tresh:
            return 1
        ent = self.results.get(read_id)
        if ent is None:
            return real(self, read_id, pos, flag, possible, exact_tresh, k)
        for p, f in zip(ent.positions, ent.flags):
            if any(pos + i in (p, -p) for i in range(-k, k + 1)) and f <= flag:
                return 1
        return real(self, read_id, pos, flag, possible, exact_tresh, k)

    monkeypatch.setattr(fo.Strand, "add_start_position", rule_by_rule)
    k, exact, calls, rets, ents, queued = DEDUP["first_entry_decides"]
    assert _oracle_dedup(k, exact, calls) != (rets, ents, queued)
