"""The reference entry point cuda_needleman_wunsch (src/cuda-functions.h:55-63)
through libsaruman_hip.so: a batch in the prepare_cuda_alignments layout
(src/hash.c:1055-1059), records printed to stdout, globals updated."""
import ctypes as C
import os
import sys
import tempfile

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


@pytest.mark.parametrize("do_reverse,kernel", [(0, "auto"), (4, "auto"), (0, "plain")])
def test_cuda_needleman_wunsch(orc, do_reverse, kernel, monkeypatch):
    print(f"[gpu] dropin {do_reverse} {kernel}", flush=True)
    if kernel == "plain":
        monkeypatch.setenv("SARUMAN_KERNEL", "plain")
    lib = _lib.load()
    rng = np.random.default_rng(21 + do_reverse)
    k, L, n = 5, 100, 2000
    flags = (5, 7) if do_reverse else (1, 3)
    batch = random_batch(rng, n, L=L, k=k, flags=flags)
    nreads = 700
    ids = rng.integers(0, nreads, n).astype(np.int32)
    names = [f"read_{i}" for i in range(nreads)]
    batch.names = [names[i] for i in ids]
    # read_names: insertion-ordered list (uthash hh.next)
    arr = (ReadName * nreads)()
    for i in range(nreads):
        arr[i].read_number = i
        arr[i].name = names[i].encode()
    for i in range(nreads - 1):
        arr[i].hh.next = C.addressof(arr[i + 1])
    C.c_void_p.in_dll(lib, "read_names").value = C.addressof(arr[0])
    ctl = (C.c_char * nreads)()
    ctl_name = "reverse_output_control" if do_reverse else "forward_output_control"
    C.c_void_p.in_dll(lib, ctl_name).value = C.addressof(ctl)
    done_name = "reverse_alignments_done" if do_reverse else "forward_alignments_done"
    cand_name = "reverse_alignment_candidates" if do_reverse else "forward_alignment_candidates"
    done0 = C.c_uint.in_dll(lib, done_name).value
    cand0 = C.c_uint.in_dll(lib, cand_name).value

    wl = np.array([len(x) for x in batch.windows], np.int64)
    genome = b"".join(batch.windows) + b"\0" * (n * k)
    reads = b"".join(batch.reads)
    goff = np.cumsum(wl).astype(np.int32)
    roff = (np.arange(1, n + 1) * L).astype(np.int32)
    eoff = np.concatenate([[0], np.cumsum((wl + 1) * (L + 1))[:-1]]).astype(np.int32)
    pos = batch.pos.astype(np.int32)
    dirs = np.asarray(batch.flags, np.int8)
    gbuf = C.create_string_buffer(genome, len(genome))
    rbuf = C.create_string_buffer(reads, len(reads))
    lib.cuda_allocate_memory(do_reverse)
    try:
        out = capture_stdout(lambda: lib.cuda_needleman_wunsch(
            0, 1, C.addressof(gbuf), C.addressof(rbuf), goff.ctypes.data, roff.ctypes.data,
            eoff.ctypes.data, n, ids.ctypes.data, pos.ctypes.data, dirs.ctypes.data, k,
            int(wl.sum() + n * k), n * L, int(((wl + 1) * (L + 1)).sum()), do_reverse))
    finally:
        lib.cuda_free_memory(do_reverse)
    want = orc.align_batch(batch)
    want_lines = sorted(r for r in want.records if r)
    got_lines = sorted(line + "\n" for line in out.splitlines())
    assert got_lines == want_lines
    hit_ids = {int(ids[t]) for t in range(n) if want.records[t]}
    assert {i for i in range(nreads) if ctl[i] == b"\x01"} == hit_ids
    assert C.c_uint.in_dll(lib, done_name).value - done0 == len(want_lines)
    assert C.c_uint.in_dll(lib, cand_name).value - cand0 == n
    assert C.c_float.in_dll(lib, "alignment_time").value >= 0
