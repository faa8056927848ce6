"""The reference entry point cuda_needleman_wunsch (src/cuda-functions.h:55-63)
through libsaruman_hip.so: a batch in the prepare_cuda_alignments layout
(src/hash.c:1055-1059), records printed to stdout, globals updated.  Read names
come from read_names either as a plain insertion-ordered list or as a uthash
1.9.4 table (the lookup the reference makes per hit, HASH_FIND_INT); the two
strand threads may call at once (hash.c:295-345, saruman.c:297-305)."""
import ctypes as C
import os
import sys
import tempfile
import threading

import numpy as np
import pytest

from saruman_amd import _lib
from saruman_amd.testing import random_batch

pytestmark = pytest.mark.gpu


class UTHandle(C.Structure):
    _fields_ = [("tbl", C.c_void_p), ("prev", C.c_void_p), ("next", C.c_void_p),
                ("hh_prev", C.c_void_p), ("hh_next", C.c_void_p), ("key", C.c_void_p),
                ("keylen", C.c_uint), ("hashv", C.c_uint)]


class ReadName(C.Structure):   # util.h:62-66
    _fields_ = [("read_number", C.c_int), ("name", C.c_char * 60), ("hh", UTHandle)]


class UTBucket(C.Structure):   # uthash.h:836-854
    _fields_ = [("hh_head", C.c_void_p), ("count", C.c_uint), ("expand_mult", C.c_uint)]


class UTTable(C.Structure):    # uthash.h:860-890
    _fields_ = [("buckets", C.c_void_p), ("num_buckets", C.c_uint),
                ("log2_num_buckets", C.c_uint), ("num_items", C.c_uint), ("tail", C.c_void_p),
                ("hho", C.c_ssize_t), ("ideal_chain_maxlen", C.c_uint),
                ("nonideal_items", C.c_uint), ("ineff_expands", C.c_uint),
                ("noexpand", C.c_uint), ("signature", C.c_uint32)]


def fnv_int(key: int) -> int:
    """FNV-1 of an int key's 4 bytes, each XORed in as a signed char: uthash's
    HASH_FNV (uthash.h:350-358), the hash SARUMAN's Makefile selects."""
    h = 2166136261
    for b in int(key).to_bytes(4, "little", signed=True):
        h = ((h * 16777619) & 0xFFFFFFFF) ^ ((b - 256 if b >= 128 else b) & 0xFFFFFFFF)
    return h


def read_names_table(names, buckets: bool):
    """read_names for the library.  buckets=False: the insertion-ordered list only
    (hh.next).  buckets=True: a uthash table with 64 buckets and the list links
    cut, so only a bucket lookup finds a name.  Returns the objects to keep alive."""
    n = len(names)
    arr = (ReadName * n)()
    for i in range(n):
        arr[i].read_number = i
        arr[i].name = names[i].encode()
    keep = [arr]
    if not buckets:
        for i in range(n - 1):
            arr[i].hh.next = C.addressof(arr[i + 1])
        return arr, keep
    nb = 64
    bk = (UTBucket * nb)()
    tbl = UTTable()
    tbl.buckets = C.addressof(bk)
    tbl.num_buckets = nb
    tbl.log2_num_buckets = 6
    tbl.num_items = n
    tbl.hho = ReadName.hh.offset
    tbl.signature = 0xA0111FE1
    for i in range(n):
        hh = arr[i].hh
        hh.tbl = C.addressof(tbl)
        hh.key = C.addressof(arr[i]) + ReadName.read_number.offset
        hh.keylen = 4
        hh.hashv = fnv_int(i)
        b = bk[hh.hashv & (nb - 1)]
        hh.hh_next = b.hh_head      # uthash adds at the head of the bucket
        if b.hh_head:
            UTHandle.from_address(b.hh_head).hh_prev = C.addressof(hh)
        b.hh_head = C.addressof(hh)
        b.count += 1
    keep += [bk, tbl]
    return arr, keep


def capture_stdout(fn):
    libc = C.CDLL(None)
    sys.stdout.flush()
    libc.fflush(None)
    saved = os.dup(1)
    with tempfile.TemporaryFile() as tmp:
        os.dup2(tmp.fileno(), 1)
        try:
            fn()
        finally:
            libc.fflush(None)
            os.dup2(saved, 1)
            os.close(saved)
        tmp.seek(0)
        return tmp.read().decode()


class StrandCall:
    """One strand's batch in the prepare_cuda_alignments layout (hash.c:1055-1059)."""

    def __init__(self, rng, do_reverse, n, L, k, nreads, names):
        self.do_reverse, self.n, self.L, self.k = do_reverse, n, L, k
        flags = (5, 7) if do_reverse else (1, 3)
        self.batch = random_batch(rng, n, L=L, k=k, flags=flags)
        self.ids = rng.integers(0, nreads, n).astype(np.int32)
        self.batch.names = [names[i] for i in self.ids]
        wl = np.array([len(x) for x in self.batch.windows], np.int64)
        self.wl = wl
        self.gbuf = C.create_string_buffer(b"".join(self.batch.windows) + b"\0" * (n * k),
                                           int(wl.sum()) + n * k)
        self.rbuf = C.create_string_buffer(b"".join(self.batch.reads), n * L)
        self.goff = np.cumsum(wl).astype(np.int32)
        self.roff = (np.arange(1, n + 1) * L).astype(np.int32)
        self.eoff = np.concatenate([[0], np.cumsum((wl + 1) * (L + 1))[:-1]]).astype(np.int32)
        self.pos = self.batch.pos.astype(np.int32)
        self.dirs = np.asarray(self.batch.flags, np.int8)

    def call(self, lib, genome_size=None):
        n, L, k = self.n, self.L, self.k
        lib.cuda_needleman_wunsch(
            0, 1, C.addressof(self.gbuf), C.addressof(self.rbuf), self.goff.ctypes.data,
            self.roff.ctypes.data, self.eoff.ctypes.data, n, self.ids.ctypes.data,
            self.pos.ctypes.data, self.dirs.ctypes.data, k,
            int(self.wl.sum() + n * k) if genome_size is None else genome_size, n * L,
            int(((self.wl + 1) * (L + 1)).sum()), self.do_reverse)


def counters(lib, do_reverse):
    done = "reverse_alignments_done" if do_reverse else "forward_alignments_done"
    cand = "reverse_alignment_candidates" if do_reverse else "forward_alignment_candidates"
    return C.c_uint.in_dll(lib, done).value, C.c_uint.in_dll(lib, cand).value


def output_control(lib, do_reverse, nreads):
    """A fresh *_output_control array for the library to write; the caller keeps
    it alive until detach_globals()."""
    ctl = (C.c_char * nreads)()
    name = "reverse_output_control" if do_reverse else "forward_output_control"
    C.c_void_p.in_dll(lib, name).value = C.addressof(ctl)
    return ctl


def detach_globals(lib):
    """Points the globals the library writes or reads back at nothing, so that no
    later call touches a test's freed arrays."""
    for g in ("read_names", "forward_output_control", "reverse_output_control"):
        C.c_void_p.in_dll(lib, g).value = None


@pytest.mark.parametrize("do_reverse,kernel,buckets", [(0, "auto", False), (4, "auto", True),
                                                       (0, "plain", True)])
def test_cuda_needleman_wunsch(orc, do_reverse, kernel, buckets, monkeypatch):
    print(f"[gpu] dropin {do_reverse} {kernel} buckets={buckets}", flush=True)
    if kernel == "plain":
        monkeypatch.setenv("SARUMAN_KERNEL", "plain")
    lib = _lib.load()
    rng = np.random.default_rng(21 + do_reverse)
    nreads = 700
    names = [f"read_{i}" for i in range(nreads)]
    sc = StrandCall(rng, do_reverse, 2000, 100, 5, nreads, names)
    arr, keep = read_names_table(names, buckets)
    C.c_void_p.in_dll(lib, "read_names").value = C.addressof(arr[0])
    ctl = output_control(lib, do_reverse, nreads)
    done0, cand0 = counters(lib, do_reverse)
    lib.cuda_allocate_memory(do_reverse)
    try:
        out = capture_stdout(lambda: sc.call(lib))
    finally:
        lib.cuda_free_memory(do_reverse)
        detach_globals(lib)
    want = orc.align_batch(sc.batch)
    want_lines = sorted(r for r in want.records if r)
    got_lines = sorted(line + "\n" for line in out.splitlines())
    assert got_lines == want_lines
    hit_ids = {int(sc.ids[t]) for t in range(sc.n) if want.records[t]}
    assert {i for i in range(nreads) if ctl[i] == b"\x01"} == hit_ids
    done1, cand1 = counters(lib, do_reverse)
    assert done1 - done0 == len(want_lines)
    assert cand1 - cand0 == sc.n
    assert C.c_float.in_dll(lib, "alignment_time").value >= 0


def test_two_strand_threads_at_once(orc):
    """The reference's two worker threads (do_reverse 0 and 4) call the entry
    point concurrently, each with its own pool (hash.c:295-345): both calls run
    at the same time here, and together print exactly both batches' records and
    update each strand's globals."""
    print("[gpu] dropin two threads", flush=True)
    lib = _lib.load()
    rng = np.random.default_rng(77)
    nreads = 900
    names = [f"r{i}" for i in range(nreads)]
    calls = {dr: StrandCall(rng, dr, 3000, 150, 8, nreads, names) for dr in (0, 4)}
    arr, keep = read_names_table(names, True)
    C.c_void_p.in_dll(lib, "read_names").value = C.addressof(arr[0])
    ctls = {dr: output_control(lib, dr, nreads) for dr in (0, 4)}
    before = {dr: counters(lib, dr) for dr in (0, 4)}
    for dr in (0, 4):
        lib.cuda_allocate_memory(dr)
    gate = threading.Barrier(2)
    errors = []

    def worker(dr):
        try:
            gate.wait()
            for _ in range(3):      # repeated calls, as successive flushes of one strand
                calls[dr].call(lib)
        except Exception as e:      # pragma: no cover - reported below
            errors.append(e)

    def both():
        th = [threading.Thread(target=worker, args=(dr,)) for dr in (0, 4)]
        for t in th:
            t.start()
        for t in th:
            t.join()

    try:
        out = capture_stdout(both)
    finally:
        for dr in (0, 4):
            lib.cuda_free_memory(dr)
        detach_globals(lib)
    assert not errors
    got = sorted(line + "\n" for line in out.splitlines())
    want = []
    for dr in (0, 4):
        w = orc.align_batch(calls[dr].batch)
        lines = [r for r in w.records if r]
        want += lines * 3
        hit_ids = {int(calls[dr].ids[t]) for t in range(calls[dr].n) if w.records[t]}
        assert {i for i in range(nreads) if ctls[dr][i] == b"\x01"} == hit_ids
        done, cand = counters(lib, dr)
        assert done - before[dr][0] == 3 * len(lines)
        assert cand - before[dr][1] == 3 * calls[dr].n
    assert got == sorted(want)


def test_malformed_batches_are_rejected(orc):
    """A batch the caller laid out wrongly is refused with a message and no
    output, where the reference read past its arrays: decreasing read offsets
    (L < 1) and a genome array shorter than its k slot bytes per candidate.  The
    candidate counter still counts the call, as cuda-functions.cu:604-608 does
    before any check; the next well-formed call works."""
    print("[gpu] dropin malformed", flush=True)
    lib = _lib.load()
    rng = np.random.default_rng(5)
    names = [f"m{i}" for i in range(300)]
    sc = StrandCall(rng, 0, 400, 36, 2, 300, names)
    arr, keep = read_names_table(names, False)
    C.c_void_p.in_dll(lib, "read_names").value = C.addressof(arr[0])
    ctl = output_control(lib, 0, 300)
    done0, cand0 = counters(lib, 0)
    lib.cuda_allocate_memory(0)
    try:
        good_roff = sc.roff.copy()
        sc.roff[10] = sc.roff[9]        # candidate 10 has L = 0
        out = capture_stdout(lambda: sc.call(lib))
        assert out == ""
        sc.roff[:] = good_roff
        out = capture_stdout(lambda: sc.call(lib, genome_size=sc.n * sc.k - 1))
        assert out == ""
        assert counters(lib, 0) == (done0, cand0 + 2 * sc.n)
        out = capture_stdout(lambda: sc.call(lib))
        hits = bytes(ctl)
    finally:
        lib.cuda_free_memory(0)
        detach_globals(lib)
    want = orc.align_batch(sc.batch)
    assert sorted(line + "\n" for line in out.splitlines()) == sorted(r for r in want.records if r)
    assert hits.count(b"\x01") == len({int(sc.ids[t]) for t in range(sc.n) if want.records[t]})


def test_get_cuda_infos():
    """get_cuda_infos (cuda-functions.cu:1051-1120): with use_gpu_value at its
    default 5000 (util.c:192-194) it picks the GPU with the most free memory,
    writes its index to use_gpu_value and returns that GPU's free bytes; with a
    GPU forced it returns the forced GPU's free bytes (Q22) and leaves
    use_gpu_value alone."""
    print("[gpu] get_cuda_infos", flush=True)
    lib = _lib.load()
    ugv = C.c_ushort.in_dll(lib, "use_gpu_value")
    saved = ugv.value
    try:
        ugv.value = 5000
        auto = int(lib.get_cuda_infos())
        assert ugv.value == 0             # the only GPU of the box
        assert 128_000_000 <= auto <= 300 * 2**30
        ugv.value = 0
        forced = int(lib.get_cuda_infos())
        assert ugv.value == 0
        assert abs(forced - auto) < 2**30   # the same device, read twice
    finally:
        ugv.value = saved
