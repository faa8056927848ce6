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
