"""Host post-processing of the product (srm_slot_rows / srm_format_record in
saruman_amd/csrc/records.cpp) against the oracle's record of the same alignment.
CPU only: the transcript is derived from the oracle's slot rows here, so these
tests check the record path independently of the kernels."""
import ctypes as C

import numpy as np

from saruman_amd import _lib
from saruman_amd.testing import random_batch

OP = {"diag": 0, "top": 1, "left": 2}


def rows_to_transcript(R: bytes, G: bytes, n: int):
    """Slot rows (alignment order) -> transcript in traceback order."""
    ops = []
    for p in range(n - 1, -1, -1):
        if R[p] == ord("_"):
            ops.append(OP["top"])
        elif G[p] == ord("_"):
            ops.append(OP["left"])
        else:
            ops.append(OP["diag"])
    words = np.zeros((len(ops) + 15) // 16 + 1, np.uint32)
    for j, o in enumerate(ops):
        words[j >> 4] |= np.uint32(o << ((j & 15) * 2))
    return words


def product_record(lib, name, pos, flag, k, read, win, score, n, er, ec):
    res = np.zeros(1, _lib.RES_DTYPE)
    res["score"], res["n_ops"], res["end_row"], res["end_col"] = score, n, er, ec
    return res


def test_records_match_oracle(orc):
    lib = _lib.load()
    rng = np.random.default_rng(11)
    buf = C.create_string_buffer(8192)
    checked = 0
    for ma, mm in ((0, 1), (0, 2), (1, 2)):
        batch = random_batch(rng, 600, L=(20, 60), k=3, flags=(0, 1, 2, 3, 4, 5, 6, 7),
                             match=ma, mismatch=mm, per_cand_k=lambda L: 1 + L // 15)
        for t in range(batch.n):
            win, read, flag = batch.windows[t], batch.reads[t], int(batch.flags[t])
            k = int(batch.ks[t])
            q1 = bool(rng.random() < 0.1)
            r = orc.align_one(ma, mm, k, win, read, flag, stop_at_row0=q1)
            if r.score > k:
                continue
            want = orc.record(batch.names[t], int(batch.pos[t]), flag, k, len(read), len(win),
                              r.score, r.R, r.G)
            ops = rows_to_transcript(r.R, r.G, r.n)
            res = product_record(lib, batch.names[t], int(batch.pos[t]), flag, k, read, win,
                                 r.score, r.n, r.end_row, r.end_col)
            rb = C.create_string_buffer(read, len(read))
            wb = C.create_string_buffer(win, len(win))
            n = lib.srm_format_record(batch.names[t].encode(), int(batch.pos[t]), flag, k,
                                      len(read), len(win), res.ctypes.data, ops.ctypes.data,
                                      C.addressof(rb), C.addressof(wb), buf, len(buf))
            assert n > 0
            assert buf.value.decode() == want, (t, want, buf.value)
            # slot rows rebuilt from the transcript equal the oracle's slot
            cap = len(win) + k
            Rs = C.create_string_buffer(cap)
            Gs = C.create_string_buffer(cap)
            assert lib.srm_slot_rows(res.ctypes.data, ops.ctypes.data, C.addressof(rb),
                                     C.addressof(wb), C.addressof(Rs), C.addressof(Gs), cap) == r.n
            assert Rs.raw == r.R and Gs.raw == r.G
            checked += 1
    assert checked > 400


def test_pack_codes_roundtrip():
    from saruman_amd.batch import pack_codes
    rng = np.random.default_rng(5)
    seq = bytes(b"ACGT"[i] for i in rng.integers(0, 4, 1000))
    words = pack_codes(seq)
    codes = ((words[:, None] >> (2 * np.arange(16, dtype=np.uint32))) & 3).reshape(-1)
    got = bytes(b"ACGT"[c] for c in codes[256:256 + 1000])
    assert got == seq
    assert not codes[:256].any() and not codes[256 + 1000:].any()
    # the C packer agrees
    lib = _lib.load()
    w2 = np.zeros_like(words)
    assert lib.srm_pack_codes(seq, len(seq), w2.ctypes.data, 256) == 0
    assert (w2 == words).all()


def test_pack_codes_offsets_and_other_bytes():
    """srm_pack_codes at every alignment: the head before a word boundary, whole
    words and the tail; bytes other than ACGT pack as 0 and are counted; codes
    outside [code_off, code_off + n) keep their value."""
    lib = _lib.load()
    rng = np.random.default_rng(9)
    alphabet = np.frombuffer(b"ACGTNacgtRX", np.uint8)
    table = {ord("A"): 0, ord("C"): 1, ord("G"): 2, ord("T"): 3}
    for _ in range(200):
        n = int(rng.integers(0, 90))
        off = int(rng.integers(0, 40))
        p_other = float(rng.choice([0.0, 0.2]))
        pick = np.where(rng.random(n) < p_other, rng.integers(4, len(alphabet), n),
                        rng.integers(0, 4, n))
        seq = alphabet[pick].tobytes()
        nw = (off + n + 15) // 16 + 1
        before = rng.integers(0, 2**32, nw, dtype=np.uint64).astype(np.uint32)
        want = [int(x) for x in before]
        bad = 0
        for i, ch in enumerate(seq):
            c = table.get(ch, 0)
            bad += ch not in table
            p = off + i
            sh = 2 * (p & 15)
            want[p >> 4] = (want[p >> 4] & ~(3 << sh) & 0xFFFFFFFF) | (c << sh)
        got = before.copy()
        assert lib.srm_pack_codes(seq, n, got.ctypes.data, off) == bad
        assert [int(x) for x in got] == want, (n, off)
