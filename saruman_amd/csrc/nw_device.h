// nw_device.h -- argument blocks shared by the kernels (nw_kernels.hip) and the
// host runtime (nw_runtime.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "saruman_hip.h"

// value seen by the traceback for matrix cells outside the candidate's own matrix
// (row -1).  The reference reads the previous candidate's matrix there
// (src/cuda-functions.cu:140,152); this never equals a cell value <= k.
#define SRM_FOREIGN 30000
// diag/left at column 0, src/cuda-functions.cu:145-146
#define SRM_COL0 999

namespace srm {

struct PlainArgs {
    int ma, mm;
    int n;                          // candidates in this launch
    const uint32_t *idx;            // [n] candidate indices
    const srm_cand *cand;
    const uint32_t *read_words;     // 2-bit codes (CODES variant)
    const uint32_t *win_words;
    const char *read_bytes;         // raw bytes (bytes variant)
    const char *win_bytes;
    int16_t *mat;                   // matrix scratch
    const uint64_t *mat_off;        // [n] element offsets into mat
    srm_result *res;
    uint32_t *ops;
};

struct BandArgs {
    int ma, mm;
    const srm_cand *cand;
    const uint32_t *perm;           // candidate indices, tile-major
    const srm_cand *cand_tile;      // the candidates in the same order (unit-cost kernels)
    int32_t *tile_score;            // scores in the same order (bits_band_kernel -> walk)
    const uint4 *tiles;             // x perm start, y count (<= 64), z max L, w max k
    uint32_t ntiles;
    uint32_t *queue;                // tile counter (zeroed before the launch)
    const uint32_t *read_words;
    const uint32_t *win_words;
    uint32_t *scratch;              // band nibbles of every tile of the class
    const uint64_t *tile_scratch;   // [ntiles] word offset of each tile's band in scratch
    srm_result *res;
    uint32_t *ops;
    int code_words;                 // LDS words per lane staging the read's codes
                                    // (bits_band_kernel, trace_kernel; 0: read_words)
    int block_waves = 4;            // waves per block of the unit-cost kernels (1, 2 or 4)
    int walk_ck = 0;                // unit costs: checkpoint layout (the walk recomputes
                                    // each 32-column segment) instead of stored D bits
    int stage_words = 0;            // unit-cost walk: transcript words per lane staged in
                                    // LDS (0: written straight to the slot)
};

struct RecArgs {
    int n;
    const srm_cand *cand;
    const srm_result *res;
    const uint32_t *ops;
    const uint32_t *read_words;     // unpadded origin (2-bit path)
    const uint32_t *win_words;
    const char *read_bytes;         // raw bytes (SRM_CAND_BYTES candidates)
    const char *win_bytes;
    const int64_t *pos;             // window genome position per candidate
    const int32_t *read_id;         // name index per candidate (null: candidate index)
    const char *names;              // name table, name_stride bytes per name (or null)
    int name_stride;
    const char *prefix;             // name prefix when there is no table
    int prefix_len;
    uint64_t *len;                  // [n + 1] line lengths (0: no record)
    uint4 *info;                    // [n] start, stop, lead | chars << 16
    uint32_t *hits;                 // [2] records counted, lines longer than rec_write_lines takes
    const uint64_t *off;            // [n + 1] line offsets
    char *out;
};

// band height for the banded kernels: d = row - col in [-k, max(2s+k, w-L)]
inline int band_cells(const srm_cand &c)
{
    const int k = c.k;
    const int s = (c.flag <= 7 && (c.flag & 2)) ? k : 0;
    const int hi = 2 * s + k > (int)c.w - (int)c.L ? 2 * s + k : (int)c.w - (int)c.L;
    return hi + k + 1;
}

// words of one tile's nibble band scratch (other costs): 8*NW diagonals x 8-column
// blocks x 64 lanes, then the tail (columns L - 1 and L packed by diagonal,
// equal-character bits)
inline size_t tile_scratch_words(int nw, int maxL)
{
    return ((size_t)nw * 8 * ((maxL + 8) / 8) + 2 * (size_t)nw + 2) * 64;
}

hipError_t launch_plain(const PlainArgs &a, bool codes, hipStream_t st);
hipError_t launch_rec_measure(const RecArgs &a, hipStream_t st);
hipError_t launch_rec_write(const RecArgs &a, hipStream_t st);
hipError_t rec_scan(void *tmp, size_t &tmp_bytes, const uint64_t *len, uint64_t *off, int n,
                    hipStream_t st);
hipError_t launch_band(const BandArgs &a, int nw, bool unit, int blocks, int maxL, hipStream_t st);
hipError_t launch_trace(const BandArgs &a, int nw, bool unit, int maxL, hipStream_t st);
// unit costs (band_bits.hip): the bit-parallel forward pass and its walk
hipError_t launch_bits_band(const BandArgs &a, int nw, int blocks, int maxL, hipStream_t st);
hipError_t launch_bits_trace(const BandArgs &a, int nw, int maxL, hipStream_t st);
int bits_trace_stage_words(int maxL);   // LDS transcript words per lane for a class's walk
int bits_band_occupancy(int nw);
size_t bits_tile_words(int nw, int maxL, bool ck);   // one tile's scratch (words)
hipError_t launch_reset(uint32_t *counters, int n, hipStream_t st);
int band_occupancy(int nw, bool unit);
hipError_t launch_copy(const void *src, void *dst, size_t bytes, hipStream_t st);

// runtime services for the other device modules (defined in nw_runtime.hip)
int set_error(int code, const char *fmt, ...);   // sets srm_last_error(), returns code
int ctx_device(const srm_ctx *c);
hipStream_t ctx_stream(const srm_ctx *c);

}  // namespace srm
