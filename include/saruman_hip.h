/*
 * saruman_hip.h -- C ABI of libsaruman_hip.so, the MI355X (gfx950) replacement of
 * SARUMAN's GPU alignment module.
 *
 * Part 1 keeps the reference's four entry points and the globals they read and
 * write, with the reference's signatures (src/cuda-functions.h:15-32, 55-78), so
 * the reference's host C (hash.c, saruman.c, util.c) links against this library
 * unchanged.  Part 2 is the extended batch API the batched, graph and serving
 * paths use: per-candidate read length and error bound, results in arrays
 * instead of printf, explicit error codes, an explicit stream per context.
 *
 * No torch types, no HIP types: plain pointers and sizes only.
 */
#ifndef SARUMAN_HIP_H_
#define SARUMAN_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ===================================================================== */
/* Part 1. Drop-in entry points (reference: src/cuda-functions.h:55-78)   */
/* ===================================================================== */

/*
 * Replaces cuda_needleman_wunsch, src/cuda-functions.h:55-63 (definition
 * src/cuda-functions.cu:584-1049).  Same arguments, same synchronous behaviour:
 * aligns sample_size candidates laid out as prepare_cuda_alignments builds them
 * (src/hash.c:898-1111), prints one TSV record per hit (score <= error) to stdout
 * (src/cuda-functions.cu:987-994), sets forward_/reverse_output_control[read_id],
 * and increments *_alignments_done / *_alignment_candidates.  Every input array
 * is borrowed.  edit_offsets and total_edit_matrix_size are accepted and ignored
 * (the library sizes its own scratch).  On an internal error it prints a message
 * to stderr and returns without output, as the reference does when a host
 * allocation fails (src/cuda-functions.cu:626-638).
 */
void cuda_needleman_wunsch(int match, int mismatch, char *genome_sequences_on_host,
                           char *read_sequences_on_host, int *genome_offsets, int *read_offsets,
                           int *edit_offsets, int sample_size, int *read_id_array,
                           int *genome_pos_array, char *direction_array, int error,
                           size_t total_genome_array_size, size_t total_read_sequence_array_size,
                           size_t total_edit_matrix_size, int do_reverse);

/*
 * Replaces cuda_allocate_memory, src/cuda-functions.h:65-67 (definition
 * src/cuda-functions.cu:541-562).  Creates the per-strand context (HIP stream and
 * device pool) for do_reverse == 4 (reverse) or any other value (forward) on GPU
 * use_gpu_value.  The pool grows on demand and is capped by cuda_memory_per_thread
 * when that global is non-zero; larger batches are split (the reference never
 * checked, Q19).
 */
void cuda_allocate_memory(int do_reverse);

/*
 * Replaces cuda_free_memory, src/cuda-functions.h:69-71 (definition
 * src/cuda-functions.cu:564-575): do_reverse != 0 frees the reverse context,
 * 0 the forward one.
 */
void cuda_free_memory(int do_reverse);

/*
 * Replaces get_cuda_infos, src/cuda-functions.h:77-78 (definition
 * src/cuda-functions.cu:1051-1120).  Lists the GPUs on stderr, picks the one with
 * the most free memory unless use_gpu_value names a valid device, and returns the
 * free bytes of the device that will be used (the reference returned the best
 * device's even when another was forced, Q22).  Exits with status 1 below
 * 128,000,000 free bytes, as the reference does.
 */
unsigned long int get_cuda_infos(void);

/* ---- globals of the drop-in contract ---------------------------------- */
/* Defined by the reference's host program (src/util.h, src/saruman.h,
 * src/hash.h tentative definitions).  libsaruman_hip.so also defines them so it
 * loads stand-alone; in a process whose executable defines them, the dynamic
 * linker binds the library's references to the executable's copies. */

/* uthash 1.9.4 handle layout (src/uthash.h), needed only to walk read_names */
typedef struct srm_ut_hash_handle {
    void *tbl;
    void *prev;
    void *next;      /* next element in insertion order */
    void *hh_prev;
    void *hh_next;
    void *key;
    unsigned keylen;
    unsigned hashv;
} srm_ut_hash_handle;

/* uthash 1.9.4 bucket and the head of its table (src/uthash.h:836-865), as far as
 * the bucket lookup of a read id reads them (HASH_FIND_INT, src/uthash.h:84-95,
 * 568-579): the bucket array, its size and the handle's offset in an element */
typedef struct srm_ut_hash_bucket {
    srm_ut_hash_handle *hh_head;
    unsigned count;
    unsigned expand_mult;
} srm_ut_hash_bucket;

typedef struct srm_ut_hash_table {
    srm_ut_hash_bucket *buckets;
    unsigned num_buckets, log2_num_buckets;
    unsigned num_items;
    srm_ut_hash_handle *tail;
    ptrdiff_t hho;
} srm_ut_hash_table;

/* src/util.h:62-66 */
struct read_name_hash {
    int read_number;
    char name[60];
    srm_ut_hash_handle hh;
};

extern struct read_name_hash *read_names;          /* util.h:55, read by name */
extern unsigned short int read_length_value;       /* util.h:25 */
extern unsigned short int be_verbose_value;        /* util.h:26 */
extern unsigned short int use_gpu_value;           /* util.h:24, written by get_cuda_infos */
extern unsigned long int cuda_memory_per_thread;   /* saruman.h:60 */
extern char *forward_output_control;               /* saruman.h:55 */
extern char *reverse_output_control;               /* saruman.h:56 */
extern unsigned int forward_alignments_done;       /* saruman.h:62 */
extern unsigned int reverse_alignments_done;       /* saruman.h:63 */
extern unsigned int forward_alignment_candidates;  /* saruman.h:65 */
extern unsigned int reverse_alignment_candidates;  /* saruman.h:66 */
extern char *main_alloc_ptr_for;                   /* cuda-functions.h:31 */
extern char *main_alloc_ptr_rev;                   /* cuda-functions.h:32 */
extern float alignment_time[2];                    /* cuda-functions.cu:61 (ms, fwd/rev) */

/* ===================================================================== */
/* Part 2. Extended batch API (no counterpart in the reference)           */
/* ===================================================================== */

/* error codes */
#define SRM_OK 0
#define SRM_EINVAL -1      /* malformed argument or candidate */
#define SRM_ENOMEM -2      /* host or device allocation failed / over budget */
#define SRM_EDEVICE -3     /* HIP runtime error */
#define SRM_ENODEV -4      /* no usable GPU */
#define SRM_ESTATE -5      /* call out of order */

/* kernel selection for srm_job_run */
#define SRM_KERNEL_AUTO 0    /* banded where eligible, plain otherwise */
#define SRM_KERNEL_PLAIN 1   /* one alignment per thread, full int16 matrix */
#define SRM_KERNEL_BANDED 2  /* persistent banded kernel (ineligible candidates -> plain) */

/* srm_cand.bits */
#define SRM_CAND_STOP_ROW0 1u   /* reproduce the first-candidate traceback stop (Q1) */
#define SRM_CAND_BYTES 2u       /* offsets are byte offsets into the byte arrays */

/* limits: read and window lengths, the banded kernels' band height and k, the plain
 * matrices' chunk under a pool cap */
#define SRM_MAX_L 1000
#define SRM_MAX_W 900
#define SRM_BAND_MAX 64      /* widest band the banded kernel holds in registers */
#define SRM_BAND_KMAX 14     /* saturated band values must fit a nibble: k + 1 <= 15 */
#define SRM_PLAIN_CHUNK ((size_t)64 << 20)   /* plain-matrix chunk under a pool cap (bytes) */

/*
 * One candidate alignment.  The read is given in the orientation it is aligned
 * in (reverse-complemented for flags 4..7, as prepare_cuda_alignments does,
 * src/hash.c:1013-1034).  The window holds rows 0..w-1 of the DP matrix.
 * 2-bit codes: A=0 C=1 G=2 T=3, 16 codes per uint32, code j of a word in bits
 * 2j..2j+1.  Code arrays must carry at least SRM_CODE_PAD codes of readable
 * padding before offset 0 and after the last code referenced.
 */
typedef struct srm_cand {
    uint32_t read_off;   /* code index (or byte offset with SRM_CAND_BYTES) of read[0] */
    uint32_t win_off;    /* code index (or byte offset) of window row 0 */
    uint32_t ops_off;    /* uint32 index of the transcript slot, >= srm_ops_words(w, k) words */
    uint16_t L;          /* read length, 1..SRM_MAX_L */
    uint16_t w;          /* window length, 0..SRM_MAX_W */
    uint8_t flag;        /* direction_array value; selects s, e (cuda-functions.cu:383-410) */
    uint8_t k;           /* error bound */
    uint8_t bits;        /* SRM_CAND_* */
    uint8_t pad;
} srm_cand;

#define SRM_CODE_PAD 256

/*
 * Per-candidate result.  score: M[w][L] for the plain kernel; the banded kernel
 * saturates scores above k at k+1.  Transcript: n_ops 2-bit operations in
 * traceback order (first op = the move out of cell (w, L)), SRM_OP_DIAG /
 * SRM_OP_TOP / SRM_OP_LEFT; the walk stopped at (end_row, end_col).  n_ops == 0
 * when score > k.  status != 0 flags a candidate the kernel rejected.
 */
typedef struct srm_result {
    int16_t score;
    uint16_t n_ops;
    uint16_t end_row;
    uint16_t end_col;
} srm_result;

#define SRM_OP_DIAG 0u   /* emits (read[col-1], window[row-1]) */
#define SRM_OP_TOP 1u    /* emits ('_', window[row-1]) */
#define SRM_OP_LEFT 2u   /* emits (read[col-1], '_') */

typedef struct srm_ctx srm_ctx;
typedef struct srm_job srm_job;

/* library/version */
const char *srm_version(void);
int srm_device_count(void);
int srm_device_sync(int device);
const char *srm_last_error(void);

/* transcript slot size in uint32 words for a window of w rows and bound k */
static inline uint32_t srm_ops_words(uint32_t w, uint32_t k) { return (w + k + 15u) / 16u; }

/* context = one HIP stream + device pool on one GPU */
int srm_ctx_create(int device, size_t pool_cap_bytes, srm_ctx **out);
void srm_ctx_destroy(srm_ctx *ctx);
/* device bytes the context's buffers (its genome and its jobs') hold now */
size_t srm_ctx_pool_used(const srm_ctx *ctx);
/* bound on those bytes for later allocations (0 = none): a buffer that would
 * take the pool past it fails with SRM_ENOMEM */
int srm_ctx_set_pool_cap(srm_ctx *ctx, size_t pool_cap_bytes);
/* testing hook for the failure paths: the next `count` device allocations of
 * any context fail with SRM_ENOMEM (0 clears) */
void srm_debug_fail_allocs(int count);

/*
 * Keep a packed genome resident on the device (uploaded once).  words/n_words
 * include the caller's padding; jobs created with win_words == NULL take their
 * windows from it (srm_cand.win_off = code index into it).
 */
int srm_ctx_set_genome(srm_ctx *ctx, const uint32_t *words, size_t n_words);
/* use the resident genome of another context on the same device (not copied; the
 * source must outlive dst's jobs) */
int srm_ctx_borrow_genome(srm_ctx *dst, const srm_ctx *src);

/*
 * Launch shape of the banded kernels, chosen by the autotuner
 * (saruman_amd/tune.py) and applied to jobs uploaded afterwards.  Every field 0
 * is the default launch.
 * blocks_per_cu[nw]: persistent 256-thread blocks per CU for band class nw
 * (8*nw band cells per lane) of the nibble kernels, 0 = as many as stay resident;
 * merge_below: band classes narrower than this run on the kernel of this class
 * (fewer launches).  The tile shape of the unit-cost kernels: tile_cands
 * candidates per wave's tile (32 or 64), block_waves waves per block (1, 2 or 4),
 * split_len: each band class splits into reads shorter than split_len and the
 * rest (two launches, each with its own LDS staging size).  walk_layout: what
 * the unit-cost forward pass leaves for the walk (1 cuts its HBM traffic; the
 * walk then does the segment's bit-parallel steps again).
 */
typedef struct srm_tuning {
    int blocks_per_cu[9];   /* by band class (8 diagonals per unit); 0 = as many as fit */
    int merge_below;        /* classes below this one run with it (0/1 = none) */
    int class_streams;      /* streams the band classes run on side by side (up to 8);
                               0 = default (4), 1 = one after another */
    int tile_cands;         /* candidates per tile: 0 = 64, or 32 */
    int block_waves;        /* waves per block of the unit-cost kernels: 0 = 4, or 1, 2 */
    int split_len;          /* 0 = no split; else the read length that splits each class */
    int walk_layout;        /* unit costs: 0 = the forward pass stores the D bits the walk
                               reads; 1 = it stores the state of every 32nd column and the
                               walk recomputes each 32-column segment from it */
} srm_tuning;
int srm_ctx_set_tuning(srm_ctx *ctx, const srm_tuning *tuning);   /* NULL = defaults */
/* after an eager srm_job_run: grid, tiles and event time (ms) of band class nw */
int srm_job_band_info(srm_job *job, int nw, int *blocks, int *ntiles, float *ms);
/* resident 256-thread blocks per CU of the banded kernel for class nw */
int srm_band_occupancy(int nw, int unit);

/*
 * A job: one batch of candidates, uploaded once, runnable many times.
 * win_words == NULL -> windows come from the context's resident genome.
 * read_bytes / win_bytes serve the SRM_CAND_BYTES candidates (may be NULL).
 */
int srm_job_create(srm_ctx *ctx, srm_job **out);
void srm_job_destroy(srm_job *job);
/* frees the job's device buffers (after waiting for its work) and the host arrays
 * an upload keeps for the next one; the next upload allocates them again */
int srm_job_release(srm_job *job);
/* Upper bound on the device bytes a job of ctx (NULL: default tuning) needs for
 * these candidates: its uploads, band scratch, results and transcripts, a chunk of
 * plain matrices (SRM_PLAIN_CHUNK, or one matrix if larger) when any candidate runs
 * on the plain kernel, and with_records the record stage (srm_job_set_records /
 * srm_job_records with a name table or prefix of names_bytes bytes, names printed
 * with at most name_max characters).  A context capped at what it already holds
 * plus this bound runs the job. */
size_t srm_job_bytes_bound(const srm_ctx *ctx, int match, int mismatch, const srm_cand *cand, int n,
                           size_t n_read_words, size_t n_win_words, size_t n_read_bytes,
                           size_t n_win_bytes, int kernel, int with_records, size_t names_bytes,
                           int name_max);
int srm_job_upload(srm_job *job, int match, int mismatch, const srm_cand *cand, int n,
                   const uint32_t *read_words, size_t n_read_words, const uint32_t *win_words,
                   size_t n_win_words, const char *read_bytes, size_t n_read_bytes,
                   const char *win_bytes, size_t n_win_bytes);
/* launch the alignment of every candidate on the context's stream and wait for it
 * (its device time is then srm_job_last_ms) */
int srm_job_run(srm_job *job, int kernel);
/* wait for the job; copy results (n) and transcript words (ops, may be NULL) */
int srm_job_download(srm_job *job, srm_result *res, uint32_t *ops, size_t n_ops_words);
/* device time (ms) of the last srm_job_run, measured with HIP events on the job's stream */
float srm_job_last_ms(srm_job *job);
/* device time (ms) of the last eager run by stage: kind 0 = banded, 1 = plain,
 * 2 = records (srm_job_set_records) */
float srm_job_kernel_ms(srm_job *job, int kind);
/* capture srm_job_run(job, kernel) into a hipGraph and replay it; 0 = plain launches */
int srm_job_use_graph(srm_job *job, int enable);
/* count of candidates per path after the last upload: 0 banded, 1 plain(codes), 2 plain(bytes) */
int srm_job_path_count(srm_job *job, int path);

/*
 * TSV records of every hit of the job's last run, formatted on the device in
 * candidate order: the lines srm_format_records writes on the host, i.e. the
 * post-processing and printf of src/cuda-functions.cu:828-994 for each candidate
 * with score <= k.  pos[n]: the window's genome position (record start before
 * leading gaps).  Names: with a table, names[read_id[t] * name_stride ...] (NUL
 * terminated within the stride, name_stride <= 256); without, name_prefix (at
 * most 255 bytes) + decimal read_id[t]
 * (read_id NULL: the candidate index).  Returns the text's byte count; the text
 * is written to out only when outcap >= that count (a size query otherwise), or
 * an SRM_E* code (< 0).  n_hits (may be NULL) receives the number of records.
 * Synchronous on the job's stream.
 */
long long srm_job_records(srm_job *job, const int64_t *pos, const int32_t *read_id,
                          const char *names, int name_stride, size_t n_names,
                          const char *name_prefix, char *out, size_t outcap, long long *n_hits);

/*
 * The same records as a stage of the job's launch sequence: after this call every
 * srm_job_run (eager or graph-replayed) also formats the records of every hit on
 * the device, into an output sized to the longest text the candidates can give,
 * with no host round trip between the alignment and the text.  The arguments are
 * those of srm_job_records.  A new upload (or srm_job_records) removes the stage.
 */
int srm_job_set_records(srm_job *job, const int64_t *pos, const int32_t *read_id,
                        const char *names, int name_stride, size_t n_names,
                        const char *name_prefix);
/* byte count of the last run's record text (and its number of records); copies
 * the text to out when outcap >= that count.  Synchronous on the job's stream. */
long long srm_job_records_text(srm_job *job, char *out, size_t outcap, long long *n_hits);

/*
 * Record text for one hit: the post-processing of src/cuda-functions.cu:828-994
 * applied to a candidate whose slot rows are reconstructed from its transcript.
 * read/win are the candidate's bytes (read oriented as aligned).  Writes one
 * '\n'-terminated line, returns its length (snprintf style) or < 0 on error.
 */
int srm_format_record(const char *name, unsigned int pos, int flag, int k, int L, int w,
                      const srm_result *res, const uint32_t *ops, const char *read,
                      const char *win, char *out, int outcap);

/* slot rows (as the reference's kernel leaves them) from a transcript; cap = w + k */
int srm_slot_rows(const srm_result *res, const uint32_t *ops, const char *read, const char *win,
                  char *R, char *G, int cap);

/*
 * srm_job_upload + srm_job_run(SRM_KERNEL_BANDED) + the results, allocating
 * nothing.  No counterpart in the reference: this is for callers that cannot
 * allocate -- code inside a hipGraph capture, tight serving loops, harnesses that
 * forbid allocation.  The caller passes the candidate descriptors on the host
 * (srm_cand, as for srm_job_upload) and device pointers to the read and window
 * codes (SRM_CODE_PAD codes of padding at both ends, as srm_job_upload takes
 * them), the results (n srm_result) and the transcripts (n_ops_words words, each
 * candidate's slot at its ops_off).  Every candidate must fit the exact band
 * (match >= 0, mismatch >= 1, k <= SRM_BAND_KMAX, band <= SRM_BAND_MAX, w >= L - k,
 * 2-bit codes): the plain kernel's matrices are not provided for.  Everything else
 * the kernels use lives in one static device array of SRM_STATIC_SCRATCH_BYTES per
 * GPU, held until the process ends; at most SRM_STATIC_MAX_CAND candidates, and a
 * batch whose band scratch does not fit is refused with SRM_ENOMEM (split it).
 * Calls are serialised and synchronous on `stream` (a hipStream_t, NULL for the
 * default stream): they never overlap, and each finds the scratch as the previous
 * call left it.
 */
#define SRM_STATIC_SCRATCH_BYTES ((size_t)128 << 20)
#define SRM_STATIC_MAX_CAND 65536
int srm_align_static(int device, void *stream, int match, int mismatch, const srm_cand *cand,
                     int n, const uint32_t *d_read_words, size_t n_read_words,
                     const uint32_t *d_win_words, size_t n_win_words, srm_result *d_res,
                     uint32_t *d_ops, size_t n_ops_words);

/* page-locked host memory (fast transfers to and from the device); NULL on failure */
void *srm_host_alloc(size_t bytes);
void srm_host_free(void *p);

/* one streaming copy of `bytes` bytes on the device (reads `bytes`, writes
 * `bytes`); used to calibrate hardware traffic counters.  ms = kernel time. */
int srm_calibrate_copy(int device, size_t bytes, float *ms);

/* pack ASCII ACGT into 2-bit codes; returns the number of non-ACGT bytes seen
 * (those are packed as A) */
long long srm_pack_codes(const char *seq, size_t n, uint32_t *words, size_t code_off);

/*
 * Records of every hit of n candidates in candidate order (bulk form of
 * srm_format_record): windows from the ASCII genome at pos[t], reads from
 * reads_fwd / reads_rev (flags 0..3 / 4..7) at read_off[t], names name_prefix +
 * read_id[t].  Returns the bytes needed (text complete only if <= outcap) or -1.
 */
long long srm_format_records(int n, const srm_result *res, const uint32_t *ops,
                             const srm_cand *cand, const char *genome, const int64_t *pos,
                             const char *reads_fwd, const char *reads_rev,
                             const uint64_t *read_off, const int32_t *read_id,
                             const char *name_prefix, char *out, size_t outcap,
                             long long *n_hits, int threads);

/* ===================================================================== */
/* Part 3. Host candidate generation (reference semantics: src/hash.c,    */
/*         src/saruman.c, src/util.c)                                     */
/* ===================================================================== */

typedef struct srm_index srm_index;
typedef struct srm_filter srm_filter;

/* q, rest, number of q-grams and exact_tresh for a read length and error bound
 * (src/saruman.c:42-75, :119) */
int srm_filter_params(int L, int k, int *q, int *rest, int *n, int *exact_tresh);

/* q-gram index over every genome offset 0..len-q for each q in qs
 * (replaces construct_qgram_index, src/util.c:456-499) */
int srm_index_build(const char *genome, size_t len, const int *qs, int nq, int threads,
                    srm_index **out);
void srm_index_destroy(srm_index *index);

/* the index of one genome chunk (construct_qgram_index with run > 0,
 * src/util.c:456-499): offsets 0..len-q of the chunk text, stored as
 * base + offset (base = chunk number * -G); offset 0 of chunk 0 stays the list
 * terminator.  The device filter takes only base 0. */
int srm_index_build_at(const char *chunk_text, size_t len, const int *qs, int nq, int threads,
                       uint32_t base, srm_index **out);

/*
 * Candidates of one strand (do_reverse 0 or 4) in the reference's batch order
 * (replaces qgram_to_read_mapping/run_seashore/add_start_position/
 * cuda_queue_check/prepare_cuda_alignments, src/hash.c:40-1111).  Each read has
 * its own length and error bound.  chunk_value is the reference's -s.  blocked
 * (may be NULL) marks separator bases between contigs.
 */
int srm_filter_run(const srm_index *index, const char *genome, size_t glen,
                   const uint8_t *blocked, const char *reads, const uint64_t *read_off,
                   const uint16_t *read_len, const uint8_t *read_k, size_t n_reads,
                   int do_reverse, int chunk_value, int threads, srm_filter **out);
/*
 * srm_filter_run over one genome chunk (the reference's pass genome_chunk_run of
 * -G mode, src/hash.c:545-552, 982, 998): positions are genome positions
 * (chunk offset plus the index's base), a read start must lie after the chunk's
 * first base, the perfect check reads chunk_text (its overhang included), and a
 * window is queued only if it ends at or before `limit` (total_bases_read).
 */
int srm_filter_run_chunk(const srm_index *index, const char *chunk_text, size_t text_len,
                         uint64_t limit, const char *reads, const uint64_t *read_off,
                         const uint16_t *read_len, const uint8_t *read_k, size_t n_reads,
                         int do_reverse, int chunk_value, int threads, srm_filter **out);
int srm_filter_sizes(const srm_filter *f, uint64_t *n_cand, uint64_t *n_batches,
                     uint64_t *n_perfect);
int srm_filter_copy(const srm_filter *f, int32_t *cand_read, int64_t *cand_pos,
                    uint8_t *cand_flag, uint16_t *cand_w, uint64_t *batch_end,
                    int32_t *perfect_read, int64_t *perfect_pos);
void srm_filter_destroy(srm_filter *f);

/*
 * The same filter with the per-read work on the GPU (same replaced reference
 * functions as srm_filter_run, src/hash.c:40-839): a device copy of the index
 * and the ASCII genome is made once on ctx's device; every run uploads the
 * reads, builds the position lists, pairs them on the device and returns the
 * same srm_filter (candidates, batch boundaries, perfect hits) as
 * srm_filter_run on the same arguments.  Runs on ctx's stream; not thread-safe
 * per object.
 */
typedef struct srm_dfilter srm_dfilter;
int srm_dfilter_create(srm_ctx *ctx, const srm_index *index, const char *genome, size_t glen,
                       srm_dfilter **out);
int srm_dfilter_run(srm_dfilter *df, const uint8_t *blocked, const char *reads,
                    const uint64_t *read_off, const uint16_t *read_len, const uint8_t *read_k,
                    size_t n_reads, int do_reverse, int chunk_value, srm_filter **out);
/* both strands over one upload of the reads: the results of do_reverse 0 and 4 */
int srm_dfilter_run2(srm_dfilter *df, const uint8_t *blocked, const char *reads,
                     const uint64_t *read_off, const uint16_t *read_len, const uint8_t *read_k,
                     size_t n_reads, int chunk_value, srm_filter **out_fwd, srm_filter **out_rev);
/* last run: device time incl. transfers, lists, list entries, reads that needed HBM slots */
int srm_dfilter_stats(const srm_dfilter *df, float *device_ms, uint64_t *n_lists,
                      uint64_t *n_list_entries, uint64_t *n_overflow);
void srm_dfilter_destroy(srm_dfilter *df);

/* seeded synthetic workloads (xoshiro256**) */
int srm_synth_genome(uint64_t seed, size_t len, double gc, char *out);
long long srm_synth_repeats(uint64_t seed, char *genome, size_t len, int n_families, int fam_min,
                            int fam_max, int n_copies, double cover_frac, double divergence);
int srm_synth_reads(uint64_t seed, const char *genome, const uint64_t *contig_lo,
                    const uint64_t *contig_hi, int n_contigs, size_t n_reads, int Lmin, int Lmax,
                    int mode, int max_err, double p_sub, double p_ins, int indel_max,
                    double p_random, char *out, uint64_t *off, uint16_t *len, int64_t *true_pos,
                    uint8_t *strand);

/* FASTA / FASTQ input (src/util.c:293-454, src/saruman.c:139-271; extended to
 * multi-line FASTA, every record a contig, reads of any length) */
typedef struct srm_fastx srm_fastx;
int srm_fastx_load(const char *path, int kind, int fixed_len, double n_rate, int n_mode,
                   srm_fastx **out);
int srm_fastx_sizes(const srm_fastx *x, uint64_t *n_records, uint64_t *n_bases,
                    uint64_t *skipped_short, uint64_t *skipped_n);
int srm_fastx_copy(const srm_fastx *x, char *seq, uint64_t *off, uint32_t *len, char *names);
void srm_fastx_destroy(srm_fastx *x);

/*
 * Genome chunks of the reference's -G mode (replaces the genome reading loop of
 * src/saruman.c:139-271 and its end test at :415): chunk i holds up to `chunk`
 * bases plus an overhang of up to read_len bases, exactly as the reference's
 * reader forms them (its header skip, buffer re-use and end-of-file byte
 * included).  Its q-gram positions start at i * chunk; limit[i] is the
 * reference's total_bases_read after chunk i, which bounds its windows.
 */
typedef struct srm_gchunks srm_gchunks;
int srm_genome_chunks_load(const char *path, uint32_t chunk, int read_len, srm_gchunks **out);
int srm_genome_chunks_sizes(const srm_gchunks *x, uint64_t *n_chunks, uint64_t *n_bytes);
/* text: n_bytes; off: n_chunks + 1 (chunk i is text[off[i], off[i+1])); limit: n_chunks */
int srm_genome_chunks_copy(const srm_gchunks *x, char *text, uint64_t *off, uint64_t *limit);
void srm_genome_chunks_destroy(srm_gchunks *x);

/* ===================================================================== */
/* Part 4. Serving: continuous batching over a resident genome            */
/* ===================================================================== */

typedef struct srm_server srm_server;

typedef struct srm_server_config {
    int device;
    int n_slots;          /* requests in flight on the device at once (1..16) */
    size_t hbm_budget;    /* device bytes the alignment slots may hold, resident genome
                             and kept buffers included (0 = 32 GiB) */
    int filter_threads;   /* host threads of the q-gram filter (0 = all) */
    int format_threads;   /* host threads of record formatting (0 = all) */
    int chunk_value;      /* the reference's -s for every request */
    int kernel;           /* SRM_KERNEL_* */
    int gpu_filter;       /* 1: q-gram filter on the device (srm_dfilter_*), 0: on the host */
    int gpu_records;      /* 1: records formatted on the device (srm_job_records) */
} srm_server_config;

typedef struct srm_request_stats {
    double t_submit, t_filtered, t_admitted, t_aligned, t_done;   /* s since server start */
    double device_ms;
    long long n_reads, n_candidates, n_hits, n_perfect, n_mapped;
    long long dev_bytes;   /* device footprint charged against the HBM budget */
    int slot, status;
    char message[160];     /* what failed, when status != SRM_OK */
} srm_request_stats;

int srm_server_create(const srm_server_config *cfg, const char *genome, size_t glen,
                      const int *qs, int nq, srm_server **out);
/* returns the request id (> 0) or an SRM_E* code */
long long srm_server_submit(srm_server *s, const char *reads, const uint64_t *read_off,
                            const uint16_t *read_len, const uint8_t *read_k, size_t n_reads);
int srm_server_wait(srm_server *s, long long id, int timeout_ms);
long long srm_server_result(srm_server *s, long long id, char *out, size_t outcap,
                            srm_request_stats *stats, int release);
/* on != 0 pauses admission (filtered requests wait; running ones finish); 0 resumes */
void srm_server_hold(srm_server *s, int on);
/* queue depths: not yet filtered, waiting for admission, on a slot */
void srm_server_queue(srm_server *s, uint64_t *incoming, uint64_t *ready, uint64_t *running);
/* HBM of the alignment slots: what they account for now (idle slots: the buffers
 * they keep; busy slots: their reservation, enforced as their pool cap) and the
 * most accounted at once.  Both stay within hbm_budget. */
void srm_server_memory(srm_server *s, size_t *now_bytes, size_t *peak_bytes);
void srm_server_destroy(srm_server *s);

#ifdef __cplusplus
}
#endif
#endif /* SARUMAN_HIP_H_ */
