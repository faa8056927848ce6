"""Pure-Python restatement of SARUMAN's -G genome reader.

TEST INFRASTRUCTURE ONLY: the checker for srm_genome_chunks_load
(saruman_amd/csrc/ingest.cpp).

Restated reference code (read as text): the genome reading loop of
src/saruman.c:139-271 and its end test at :415.  The file is an io.BytesIO whose
read/seek/tell stand for the reference's fread/fseek/ftell.  `buf` is its
calloc'd buffer of G + L bytes: fread(buffer, 1, n) overwrites only the bytes it
returns, and a single-byte fread at end of file leaves buffer[0] as it was
(saruman.c:237-246 and 258-266 then still test and append that byte).

Each chunk is (text, base, limit): the bases of the pass (its G part, then the
overhang of up to L bases), the genome position of text[0] (chunk number * G,
util.c:481 and hash.c:998), and total_bases_read after the pass (hash.c:982).
"""
from __future__ import annotations

import io


def _up(b: int) -> int:
    # toupper in the C locale
    return b - 32 if 97 <= b <= 122 else b


def read_chunks(data: bytes, G: int, L: int) -> list:
    f = io.BytesIO(data)
    chars_to_read = G + L
    buf = bytearray(chars_to_read)
    buffer_chars_read = 0
    buffer_character_position = 0
    total_bases_read = 0
    first_run = True
    file_position = 0
    chunks = []
    run = 0
    while True:
        seq = bytearray()
        if not first_run:
            f.seek(file_position)
            if buffer_character_position < chars_to_read:     # saruman.c:187-203
                p = buffer_character_position
                while len(seq) < G and p < buffer_chars_read:
                    if buf[p] != 10:
                        seq.append(_up(buf[p]))
                    p += 1
        got = f.read(chars_to_read)                              # saruman.c:208
        buf[:len(got)] = got
        buffer_chars_read = len(got)
        sequence_start = 0
        if first_run:                                            # strcspn(buffer, "\n")
            while (sequence_start < chars_to_read and buf[sequence_start] != 10
                   and buf[sequence_start] != 0):
                sequence_start += 1
        buffer_character_position = sequence_start
        while len(seq) < G and buffer_character_position < buffer_chars_read:   # :225-232
            if buf[buffer_character_position] != 10:
                seq.append(_up(buf[buffer_character_position]))
            buffer_character_position += 1
        first_run = False
        result = 1
        while len(seq) < G and result == 1:                      # :237-246
            one = f.read(1)
            result = len(one)
            if result:
                buf[0] = one[0]
            if buf[0] != 10:
                seq.append(_up(buf[0]))
        total_bases_read += len(seq)                             # :248-249
        file_position = f.tell()
        overhang = 0
        while overhang < L and result == 1:                      # :258-266
            one = f.read(1)
            result = len(one)
            if result:
                buf[0] = one[0]
            if buf[0] != 10:
                seq.append(_up(buf[0]))
                overhang += 1
        chunks.append((bytes(seq), run * G, total_bases_read))
        run += 1
        if len(seq) < G:                                         # :415
            return chunks
        if run > len(data) // G + 3:
            raise ValueError("the reader does not end (G too small for this file)")


def base_stream(chunks: list) -> bytes:
    """The genome in genome coordinates: every chunk's bases up to its limit
    (the G part; windows never reach further)."""
    return b"".join(t[:lim - base] for t, base, lim in chunks)
