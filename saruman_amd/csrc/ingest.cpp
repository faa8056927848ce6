// ingest.cpp -- FASTA / FASTQ input.
//
// Reference behaviour kept (src/util.c:293-454, src/saruman.c:139-271):
//   - sequences are upper-cased;
//   - a read name is its header line after '>' / '@', at most 59 characters
//     (the reference reads it with fgets into 60 bytes);
//   - with a fixed read length, a read shorter than that length is skipped and a
//     longer one is cut to it; a read with more 'N' than its error bound is
//     dropped (util.c:385-434).
// Extended: multi-line FASTA sequences, every genome record is a contig (the
// reference reads one record and treats later headers as sequence, Q17), and
// reads of any length (each with its own error bound floor/ceil(rate * L)).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "saruman_hip.h"

struct srm_fastx {
    std::string seq;                 // all records concatenated
    std::vector<uint64_t> off;
    std::vector<uint32_t> len;
    std::vector<std::string> name;
    size_t skipped_short = 0, skipped_n = 0;
};

namespace {

bool read_line(FILE *f, std::string &s)
{
    s.clear();
    int c;
    bool any = false;
    while ((c = fgetc(f)) != EOF) {
        any = true;
        if (c == '\n') break;
        s.push_back((char)c);
    }
    if (!s.empty() && s.back() == '\r') s.pop_back();
    return any;
}

inline char upper(char c) { return (c >= 'a' && c <= 'z') ? (char)(c - 32) : c; }

}  // namespace

extern "C" {

/*
 * kind: 0 auto (first byte '>' FASTA, '@' FASTQ), 1 FASTA, 2 FASTQ.
 * fixed_len > 0: reads shorter are skipped, longer are cut (the reference's -l).
 * n_rate >= 0: drop a record with more 'N' than k = ceil(n_rate * L) (n_mode 1) or
 * floor(n_rate * L) (n_mode 0, the reference's k); n_rate < 0 keeps every record.
 * FASTA records may span several lines.
 */
int srm_fastx_load(const char *path, int kind, int fixed_len, double n_rate, int n_mode,
                   srm_fastx **out)
{
    *out = nullptr;
    FILE *f = fopen(path, "rb");
    if (!f) return SRM_EINVAL;
    int first = fgetc(f);
    if (first == EOF) {
        fclose(f);
        *out = new srm_fastx();
        return SRM_OK;
    }
    ungetc(first, f);
    if (kind == 0) kind = first == '@' ? 2 : 1;
    auto *X = new srm_fastx();
    std::string line, hdr, seq, qual;
    auto commit = [&](const std::string &h, std::string &s) {
        for (auto &c : s) c = upper(c);
        if (fixed_len > 0) {
            if ((int)s.size() < fixed_len) {
                X->skipped_short++;
                return;
            }
            s.resize((size_t)fixed_len);
        }
        if (n_rate >= 0) {
            const double kk = n_rate * (double)s.size();
            const long k = n_mode ? (long)std::ceil(kk - 1e-9) : (long)std::floor(kk);
            long nn = 0;
            for (char c : s) nn += (c == 'N');
            if (nn > k) {
                X->skipped_n++;
                return;
            }
        }
        X->off.push_back(X->seq.size());
        X->len.push_back((uint32_t)s.size());
        X->seq += s;
        X->name.push_back(h.substr(0, 59));
    };
    if (kind == 2) {
        while (read_line(f, hdr)) {
            if (hdr.empty()) continue;
            if (hdr[0] != '@') {
                delete X;
                fclose(f);
                return SRM_EINVAL;
            }
            if (!read_line(f, seq) || !read_line(f, line) || !read_line(f, qual)) {
                delete X;
                fclose(f);
                return SRM_EINVAL;
            }
            commit(hdr.substr(1), seq);
        }
    } else {
        bool have = false;
        while (read_line(f, line)) {
            if (!line.empty() && line[0] == '>') {
                if (have) commit(hdr, seq);
                hdr = line.substr(1);
                seq.clear();
                have = true;
            } else if (have) {
                seq += line;
            } else if (!line.empty()) {
                delete X;
                fclose(f);
                return SRM_EINVAL;
            }
        }
        if (have) commit(hdr, seq);
    }
    fclose(f);
    *out = X;
    return SRM_OK;
}

int srm_fastx_sizes(const srm_fastx *X, uint64_t *n_records, uint64_t *n_bases,
                    uint64_t *skipped_short, uint64_t *skipped_n)
{
    *n_records = X->off.size();
    *n_bases = X->seq.size();
    if (skipped_short) *skipped_short = X->skipped_short;
    if (skipped_n) *skipped_n = X->skipped_n;
    return SRM_OK;
}

/* names: n_records x 60 bytes, NUL-terminated */
int srm_fastx_copy(const srm_fastx *X, char *seq, uint64_t *off, uint32_t *len, char *names)
{
    if (seq) std::memcpy(seq, X->seq.data(), X->seq.size());
    for (size_t i = 0; i < X->off.size(); ++i) {
        if (off) off[i] = X->off[i];
        if (len) len[i] = X->len[i];
        if (names) {
            std::memset(names + 60 * i, 0, 60);
            std::memcpy(names + 60 * i, X->name[i].data(), X->name[i].size());
        }
    }
    return SRM_OK;
}

void srm_fastx_destroy(srm_fastx *X) { delete X; }

}  // extern "C"

// ---------------------------------------------------------------------------
// Genome chunks of the reference's -G mode (src/saruman.c:139-271, :415).
//
// The reference reads the genome file in passes.  Each pass yields one chunk:
// up to G bases (newlines dropped, upper-cased) plus an overhang of up to L more
// bases, where L is the read length.  The chunk's q-gram positions are offset by
// chunk * G (util.c:481).  Its windows are bounded by total_bases_read, the G-part
// bases of every pass so far (hash.c:982).  Kept here byte for byte, because they
// decide which windows exist:
//   - the first pass drops the bytes before the first '\n' (or NUL) of its buffer,
//     whatever they are (saruman.c:217-223);
//   - a pass first re-uses the bytes of the previous bulk read after the point
//     where that pass's G part ended.  The previous overhang, though, was read from
//     the file position after the bulk read.  When one bulk read held more than G
//     bases, that overhang therefore skips those bytes (saruman.c:187-203, 250-268);
//   - a single-byte read at end of file leaves the buffer's first byte as it was,
//     and that byte is still appended when it is not '\n' (saruman.c:237-246,
//     258-266).  A genome that ends inside a bulk read thus gains its first byte:
//     the header's '>' on the first pass;
//   - passes continue while a pass filled its G part (saruman.c:415).
// The reference's strcspn may read past its buffer when the first G + L bytes hold
// neither '\n' nor NUL; here the scan stops at the buffer's end.
// ---------------------------------------------------------------------------
struct srm_gchunks {
    std::string text;             // chunk texts, concatenated
    std::vector<uint64_t> off;    // chunk i is text[off[i], off[i+1])
    std::vector<uint64_t> limit;  // total_bases_read after chunk i
};

namespace {

// the file as the reference's stdio calls see it: bulk reads of the buffer's size
// and single-byte reads into its first byte; the buffer keeps what it last held
struct PassReader {
    std::vector<char> file;
    size_t at = 0;                // file position
    std::vector<char> buf;        // G + L bytes, zero at the start (calloc)
    size_t filled = 0;            // bytes of the last bulk read
    bool got = true;              // the last single-byte read returned a byte

    void bulk()
    {
        filled = std::min(buf.size(), file.size() - at);
        std::memcpy(buf.data(), file.data() + at, filled);
        at += filled;
    }
    char one()
    {
        got = at < file.size();
        if (got) buf[0] = file[at++];
        return buf[0];
    }
};

}  // namespace

extern "C" {

int srm_genome_chunks_load(const char *path, uint32_t chunk, int read_len, srm_gchunks **out)
{
    *out = nullptr;
    if (chunk < 1 || read_len < 0) return SRM_EINVAL;
    FILE *f = fopen(path, "rb");
    if (!f) return SRM_EINVAL;
    PassReader R;
    char tmp[1 << 16];
    size_t got;
    while ((got = fread(tmp, 1, sizeof tmp, f)) > 0) R.file.insert(R.file.end(), tmp, tmp + got);
    fclose(f);
    const size_t G = chunk, cap = (size_t)chunk + (size_t)read_len;
    R.buf.assign(cap, 0);
    auto *X = new srm_gchunks();
    std::string t;
    size_t scan = 0, resume = 0;
    uint64_t total = 0;
    const size_t max_passes = R.file.size() / G + 3;
    for (size_t pass = 0;; ++pass) {
        if (pass > max_passes) {       // a G of 1 on a file that ends in a base
            delete X;
            return SRM_EINVAL;
        }
        t.clear();
        if (pass > 0) {
            R.at = resume;
            for (size_t i = scan; t.size() < G && i < R.filled; ++i)
                if (R.buf[i] != '\n') t.push_back(upper(R.buf[i]));
        }
        R.bulk();
        size_t i = 0;
        if (pass == 0)
            while (i < cap && R.buf[i] != '\n' && R.buf[i] != 0) ++i;
        for (; t.size() < G && i < R.filled; ++i)
            if (R.buf[i] != '\n') t.push_back(upper(R.buf[i]));
        scan = i;
        R.got = true;
        while (t.size() < G && R.got) {
            const char c = R.one();
            if (c != '\n') t.push_back(upper(c));
        }
        total += t.size();
        resume = R.at;
        for (int ov = 0; ov < read_len && R.got;) {
            const char c = R.one();
            if (c != '\n') {
                t.push_back(upper(c));
                ++ov;
            }
        }
        X->off.push_back(X->text.size());
        X->text += t;
        X->limit.push_back(total);
        if (t.size() < G) break;
    }
    X->off.push_back(X->text.size());
    *out = X;
    return SRM_OK;
}

int srm_genome_chunks_sizes(const srm_gchunks *X, uint64_t *n_chunks, uint64_t *n_bytes)
{
    *n_chunks = X->limit.size();
    *n_bytes = X->text.size();
    return SRM_OK;
}

/* text: n_bytes; off: n_chunks + 1; limit: n_chunks */
int srm_genome_chunks_copy(const srm_gchunks *X, char *text, uint64_t *off, uint64_t *limit)
{
    if (text) std::memcpy(text, X->text.data(), X->text.size());
    if (off) std::memcpy(off, X->off.data(), X->off.size() * 8);
    if (limit) std::memcpy(limit, X->limit.data(), X->limit.size() * 8);
    return SRM_OK;
}

void srm_genome_chunks_destroy(srm_gchunks *X) { delete X; }

}  // extern "C"
