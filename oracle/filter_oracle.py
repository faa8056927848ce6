"""Pure-Python restatement of SARUMAN's host-side candidate generation.

TEST INFRASTRUCTURE ONLY: the checker for the product's C++ filter and batch
builder (saruman_amd/csrc/filter.cpp).  It is slow (pure Python loops) and is
meant for genomes of a few tens of kbp and a few thousand reads.

Restated reference code (read as text):
  - parameter derivation          src/saruman.c:42-75, :119
  - construct_qgram_index         src/util.c:456-499, src/hash.c:180-240
  - reverse_complement            src/util.c:22-57
  - qgram_to_read_mapping         src/hash.c:348-471  (0-terminated position lists)
  - run_seashore                  src/hash.c:473-839  (perfect check, 2 pairing passes)
  - add_start_position            src/hash.c:40-178   (dedup within +-k)
  - cuda_queue_check              src/hash.c:841-851
  - prepare_cuda_alignments       src/hash.c:898-1111 (batch layout)

-G chunks (src/saruman.c:139-415): build_index(base=) and run_strand(base=,
limit=) restate one pass of the chunk loop: q-gram positions are
base + offset (util.c:481), a read start must lie after the chunk's first base
(hash.c:545), the perfect check reads the chunk text (hash.c:552), a window must
end at or before total_bases_read (hash.c:982) and is read from the chunk at
position - base (hash.c:998).  A start-free window of chunk > 0 may begin up to
k bases before the chunk: the reference then indexes its chunk buffer with a
negative offset (in unsigned arithmetic), which is undefined; this restatement,
and the product, take those bases from the genome (window_genome).

Quirks kept on purpose: Q10 (s/e leak after an out-of-genome skip), Q11 (flag
branches exactly as written), Q12 (the first-pass scan of list b starts at the
index of P in list a, then 1, 2, ...), Q13 (only the exact_tresh-th matching
pair inserts), Q14 (position 0 is the list terminator), Q16 (reverse perfect
hits stored with flag 0).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

_COMP = {ord("A"): ord("T"), ord("C"): ord("G"), ord("G"): ord("C"), ord("T"): ord("A")}

# how often the quirk paths fired (read by the tests to show they are exercised)
STATS = {"spill": 0, "skip": 0, "leak": 0}


def reverse_complement(seq: bytes) -> bytes:
    """util.c:35-57 -- A<->T, C<->G, every other byte unchanged, then reversed."""
    return bytes(_COMP.get(b, b) for b in seq)[::-1]


@dataclass
class Params:
    L: int
    k: int
    q: int
    rest: int
    n: int
    positions: list
    exact_tresh: int


def derive_params(L: int, k: int) -> Params:
    """saruman.c:46-75 and :119 for a given read length and error bound k."""
    q = math.ceil(float(L) / (k + 1)) - 1
    rest = L % q
    n = L // q
    if rest != 0:
        n *= 2
    pos = [i * q for i in range(L // q)]
    pos += [L - z * q for z in range(1, L // q + 1)]
    exact_tresh = math.floor(float(L) / float(q)) - k + 1
    return Params(L, k, q, rest, n, pos, exact_tresh)


def build_index(genome: bytes, q: int, base: int = 0) -> dict:
    """util.c:474-481: every offset 0..len-q, ascending, keyed by the q bytes,
    stored as base + offset (base = chunk number * G)."""
    idx: dict = {}
    for p in range(0, len(genome) - q + 1):
        idx.setdefault(genome[p:p + q], []).append(p + base)
    return idx


@dataclass
class ResultEntry:
    positions: list = field(default_factory=list)
    flags: list = field(default_factory=list)


class Strand:
    """One strand worker's state: the result hash (insertion ordered) and queue."""

    def __init__(self):
        self.results: dict = {}   # read_id -> ResultEntry, insertion ordered
        self.queue = 0

    def add_start_position(self, read_id, pos, flag, possible, exact_tresh, k):
        """hash.c:40-178."""
        if possible != exact_tresh:
            return 1
        direction = 1 if flag > 3 else 0
        ent = self.results.get(read_id)
        if ent is None:
            self.results[read_id] = ResultEntry([pos], [flag])
            if pos > 0:
                self.queue += 1
            return 0
        for cp in range(len(ent.positions)):
            p = ent.positions[cp]
            f = ent.flags[cp]
            tdir = 1 if f > 3 else 0
            for i in range(-k, k + 1):
                if (pos + i == p or pos + i == -p) and f <= flag:
                    return 1
            if p == pos and tdir == direction and f > flag:
                ent.flags[cp] = flag
                return 2
            elif (p == pos and tdir == direction) or (p == -pos and tdir == direction):
                return 1
        ent.positions.append(pos)
        ent.flags.append(flag)
        if pos > 0:
            self.queue += 1
        return 2


@dataclass
class Candidate:
    read_id: int
    pos: int
    flag: int
    w: int
    window: bytes
    read: bytes


def _offsets(flag, k):
    # hash.c:938-979 (only the offsets the switch sets)
    s = e = None
    if flag in (1, 5):
        e = k
    elif flag in (2, 6):
        s = k
    elif flag in (3, 7):
        s = k
        e = k
    return s, e


def prepare_batch(strand: Strand, genome: bytes, reads: list, L: int, k: int, do_reverse: int,
                  total_bases: int, base: int = 0, window_genome: bytes | None = None):
    """hash.c:898-1111 -- returns the list of candidates of one call (may be empty)."""
    out = []
    if strand.queue <= 0:
        return out
    start_offset = 0
    end_offset = 0
    for rid, ent in strand.results.items():
        for i in range(len(ent.positions)):
            if ent.positions[i] > 0:
                f = ent.flags[i]
                s, e = _offsets(f, k)
                if s is not None:
                    start_offset = s
                if e is not None:
                    end_offset = e
                p = ent.positions[i]
                if p + L + start_offset + end_offset > total_bases:
                    STATS["skip"] += 1
                    continue          # Q10: skips the reset below
                if L + start_offset + end_offset != L + (s or 0) + (e or 0):
                    STATS["leak"] += 1
                w = L + start_offset + end_offset
                win = genome[p - base:p - base + w] if p >= base else window_genome[p:p + w]
                win = win.split(b"\x00")[0]
                win = win + b"\x00" * (w - len(win))   # strncpy zero pad
                rd = reads[rid]
                if f > 3:
                    # hash.c:1013-1034: the entry's flag selects the orientation
                    rd = reverse_complement(rd)
                out.append(Candidate(rid, p, f, w, win, rd))
                ent.positions[i] = -p
            end_offset = 0
            start_offset = 0
    strand.queue = 0
    return out


def seashore_read(strand: Strand, data: list, offs: list, counter: int, read_id: int,
                  tmp_read: bytes, genome: bytes, P: Params, do_reverse: int, base: int = 0):
    """run_seashore's body for one read, hash.c:502-834.

    data: the chunk's data_array (0-terminated position lists, hash.c:396-461);
    offs: data_array_offsets; counter: index of the read's first list in offs;
    tmp_read: the read as the strand sees it (reverse complement for strand 4).
    Adds the read's candidates to `strand` and returns the S of every perfect hit
    the reference prints (hash.c:564-583), in order.
    """
    L, k, q, rest = P.L, P.k, P.q, P.rest
    half = P.n // 2
    dr = do_reverse
    printed = []

    def D(j):
        # reads past the end of the array see zeros (calloc'd / memset tail, hash.c:384-461)
        return data[j] if 0 <= j < len(data) else 0

    def Db(list_start, idx):
        # first-pass read of list b at index idx (Q12); counts reads that land
        # outside list b
        v = D(list_start + idx)
        end = list_start
        while end < len(data) and data[end] != 0:
            end += 1
        if list_start + idx > end:
            STATS["spill"] += 1
        return v

    for a in range(half):
        fqm = 1 if a == 0 else 0
        in_read = a * q + 1
        off = 0
        Pa = D(offs[counter + a] + off)
        while Pa > 0:
            gs = Pa - in_read
            if gs + 1 - base > 0:
                tg = genome[gs + 1 - base:gs + 1 - base + L]
                tg = tg.split(b"\x00")[0]
                if tg == tmp_read:
                    if strand.add_start_position(read_id, -(gs + 1), 0, P.exact_tresh,
                                                 P.exact_tresh, k) != 1:
                        printed.append(gs + 1)
                possible = 1
                # first pass, hash.c:606-692
                for b in range(a + 1, half):
                    rstart = b * q + 1
                    for err in range(-k, k + 1):
                        Pb = Db(offs[counter + b], off)
                        nxt = 0
                        while Pb > 0:
                            if Pb == gs + rstart + err:
                                if b == half - 1 and fqm:
                                    flag, pos = 1 + dr, gs + 1
                                elif a == half - 1:
                                    if Pa + q > L:
                                        flag, pos = 3 + dr, gs - k + 1
                                    else:
                                        flag, pos = 2 + dr, gs - k + 1
                                elif fqm:
                                    flag, pos = 1 + dr, gs + 1
                                else:
                                    flag, pos = 3 + dr, gs - k + 1
                                possible += 1
                                strand.add_start_position(read_id, pos, flag, possible,
                                                          P.exact_tresh, k)
                            nxt += 1
                            Pb = D(offs[counter + b] + nxt)
                # second pass, hash.c:695-794
                for j in range(half):
                    lim = abs(half - 1 - j - a)
                    for err in range(-lim, lim + 1):
                        aerr = -err if j - a < 0 else err
                        nin = (half - j - 1) * q + 1 + rest
                        so = 0
                        Pb = D(offs[counter + half + j] + so)
                        while Pb > 0:
                            if Pb == gs + nin + aerr:
                                if j == 0 and fqm:
                                    flag, pos = 1 + dr, gs + 1
                                elif j == 0:
                                    if nin + q > L:
                                        flag, pos = 3 + dr, gs - k + 1
                                    else:
                                        flag, pos = 2 + dr, gs - k + 1
                                elif fqm:
                                    flag, pos = 1 + dr, gs + 1
                                else:
                                    flag, pos = 3 + dr, gs - k + 1
                                possible += 1
                                strand.add_start_position(read_id, pos, flag, possible,
                                                          P.exact_tresh, k)
                            so += 1
                            Pb = D(offs[counter + half + j] + so)
            off += 1
            Pa = D(offs[counter + a] + off)
    return printed


def run_strand(genome: bytes, reads: list, names: list, L: int, k: int, do_reverse: int,
               chunk_value: int, index: dict | None = None, base: int = 0,
               limit: int | None = None, window_genome: bytes | None = None):
    """qgram_to_read_mapping + run_seashore for one strand.

    Returns (batches, perfect_records) where batches is a list of candidate lists
    (one per cuda_needleman_wunsch call, in order) and perfect_records the record
    lines printed by the filter (hash.c:580).
    """
    P = derive_params(L, k)
    q, n = P.q, P.n
    if index is None:
        index = build_index(genome, q, base)
    total_bases = len(genome) if limit is None else limit
    strand = Strand()
    batches = []
    perfect = []
    read_number = len(reads)
    chunksize = chunk_value
    max_runs = int(math.ceil(read_number / float(chunksize)))

    for run in range(max_runs):
        # ---- data_array construction, hash.c:377-461 ----
        data = []
        offs = [0] * (chunksize * n)
        for i in range(run * chunksize, (run + 1) * chunksize):
            pidx = (i - run * chunksize) * n
            if i < read_number:
                rd = reads[i]
                if do_reverse == 4:
                    rd = reverse_complement(rd)
                for x in range(n):
                    offs[pidx + x] = len(data)
                    qg = rd[P.positions[x]:P.positions[x] + q]
                    data.extend(index.get(qg, []))
                    data.append(0)

        # ---- run_seashore, hash.c:473-839 ----
        limit = read_number
        if (run + 1) * chunksize > limit:
            limit = read_number % chunksize
        read_id = run * chunksize
        counter = 0
        counter2 = 0
        while counter2 < limit and read_id < (run + 1) * chunksize:
            fwd_read = reads[read_id]
            tmp_read = reverse_complement(fwd_read) if do_reverse == 4 else fwd_read
            for S in seashore_read(strand, data, offs, counter, read_id, tmp_read, genome, P,
                                   do_reverse, base):
                perfect.append(
                    "%s\t%d\t%d\t%s\t%s\t%s\t%d\n" % (
                        names[read_id], S, S + L - 1, "<<" if do_reverse else ">>",
                        tmp_read.decode(), tmp_read.decode(), 0))
            # cuda_queue_check, hash.c:835, 841-851
            if strand.queue > chunk_value:
                b = prepare_batch(strand, genome, reads, L, k, do_reverse, total_bases, base,
                                  window_genome)
                if b:
                    batches.append(b)
                strand.queue = 0
            read_id += 1
            counter2 += 1
            counter += n
    b = prepare_batch(strand, genome, reads, L, k, do_reverse, total_bases, base,
                      window_genome)   # hash.c:468
    if b:
        batches.append(b)
    return batches, perfect
