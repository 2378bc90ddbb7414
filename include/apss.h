/*
 * apss.h -- C-ABI of the MI355X (gfx950) all-pairs similarity search library
 * (libapss.so).  Plain C types only; no HIP or torch types in any signature.
 *
 * The reference exposes exactly one entry point,
 *   void similarity_join_device(int* dim_A, float* feature_A, ... 70 args)
 * (similarity_search.h:3-14, defined at similarity_search.cu:934-3015), with a
 * compile-time threshold (similarity_search.cu:18), a literal head/tail split
 * of 4096 and no output.  This header keeps that symbol unchanged (bottom of
 * the file) and adds the runtime-parameterised interface it implies: a CSR
 * matrix, a threshold t and a split p in, (i, j, score) pairs out.
 *
 * Semantics shared by every entry point (reference hybrid, similarity_search.cu
 * :370-437, 441-506, 510-531, 642-735, 833-915):
 *   - rows are sparse vectors; a score is the raw dot product sum_d x_id*x_jd
 *     (the reference never normalises; inputs are expected L2-normalised);
 *   - the result is every pair i < j with score >= t (self-join), or every
 *     (query, corpus) pair with score >= t (index queries);
 *   - feature ids < p form the head (forward-list dot product), ids >= p the
 *     tail (inverted lists; the most frequent tail features go to a dense
 *     MFMA head); the emitted pair set does not depend on p;
 *   - the pair set equals the exact FP64 result (products exact in double,
 *     summed in ascending feature id): every pair whose fast-path score lies
 *     within its rigorous rounding bound of t is rescored in FP64 on the GPU;
 *   - APSS_PREC_FP32: every emitted score is within 1e-6 relative of the FP64
 *     score.  APSS_PREC_BF16_HEAD: the pair set is still exact; an emitted
 *     score is within c * |x_H| |y_H| + 2^-S + (FP32 tail bound) of the FP64
 *     score, c = 2^-7 + 2^-16 + 2*gamma_128 and |x_H| the norm of the row's
 *     features in the BF16 head, or, when no value of either matrix is
 *     negative, within c / (1 - c) * s (+ the same terms) for the score s
 *     (either way at most 7.93e-3 for unit rows);
 *   - errors are returned as negative APSS_E_* codes; nothing calls exit().
 */
#ifndef APSS_H
#define APSS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define APSS_VERSION_MAJOR 0
#define APSS_VERSION_MINOR 1
#define APSS_VERSION_PATCH 0

enum {
  APSS_OK = 0,
  APSS_E_INVALID = -1,     /* malformed CSR, bad t/p/option */
  APSS_E_NOMEM = -2,       /* allocation failed or memory budget exceeded */
  APSS_E_DEVICE = -3,      /* HIP runtime error */
  APSS_E_UNSUPPORTED = -4, /* input outside the supported envelope */
  APSS_E_CANCELLED = -5,   /* request cancelled before completion */
  APSS_E_BUSY = -6         /* serving admission refused (queue full / over budget) */
};

enum { APSS_PREC_FP32 = 0, APSS_PREC_BF16_HEAD = 1 };

/* CSR matrix: row_ptr[n_rows+1] (row_ptr[0]=0, nondecreasing), col[nnz]
 * strictly ascending inside each row and < n_cols, val[nnz] finite.
 * on_device=1 means the three arrays are device pointers on opts.device.
 * Replaces the SELL-32 arrays dim_X/feature_X/tile_start_X that main.cu
 * builds per partition (main.cu:286-339). */
typedef struct {
  int64_t n_rows;
  int64_t n_cols;
  int64_t nnz;
  const int64_t* row_ptr;
  const int32_t* col;
  const float* val;
  int32_t on_device;
  int32_t reserved_;
} apss_csr;

/* One result record (12 bytes).  For self-joins i < j are row indices; for
 * index queries i is the query row and j the corpus row. */
typedef struct {
  uint32_t i;
  uint32_t j;
  float score;
} apss_pair;

typedef struct {
  int32_t device;          /* HIP device ordinal (default 0) */
  int32_t precision;       /* APSS_PREC_FP32 (default) or APSS_PREC_BF16_HEAD */
  int32_t sort_output;     /* 1: pairs sorted by (i, j) */
  int32_t rescore_all;     /* 1: every emitted score is the FP64 rescore */
  int32_t output_on_device;/* 1: result.pairs is device memory (free with apss_result_free) */
  int32_t dense_features;  /* cap of the bitmap tier (0 = automatic) */
  int64_t mem_budget;      /* device bytes a call may allocate (0 = no limit).  Device
                              blocks a call frees are kept for later calls on the same
                              device, up to 8 GiB per device, for the life of the process
                              (a failed allocation releases them first). */
  void* stream;            /* hipStream_t to run on (NULL = library-owned stream) */
  int32_t mfma_head;       /* dense MFMA head of the most frequent tail features:
                              0 = on (64 features as exact int8 limbs in FP32 mode,
                              128 BF16 features in APSS_PREC_BF16_HEAD mode),
                              -1 = off (every feature on the sparse path) */
  int32_t band_rows;       /* height cap of the bands, 2..128 (0 = 128); tiles take half;
                              p < 0 lets the autotuner choose */
  int32_t count_work;      /* 1: count the bitmap-tier list visits and products
                              (apss_result.list_visits / products; -1 otherwise) with a
                              form of the band kernel that is about 2% slower */
  int32_t reserved32_;
  int64_t reserved_[2];
} apss_opts;

typedef struct {
  int64_t count;           /* number of result pairs */
  apss_pair* pairs;        /* library-allocated; release with apss_result_free */
  int32_t on_device;
  int32_t p_used;          /* split actually used (after autotune) */
  /* counters */
  int64_t cells;           /* pair cells evaluated on the GPU */
  int64_t band;            /* pairs sent to the FP64 rescore */
  int64_t products;        /* partial products accumulated (-1 unless opts.count_work) */
  int64_t head_pairs;      /* pairs that ran the head forward-list dot */
  int64_t list_visits;     /* tile-side inverted-list entries visited: the J tile's segment of
                              every feature the band holds, once per tile pair (-1 unless
                              opts.count_work) */
  int64_t cursor_products; /* products of the short-list (cursor) tier */
  int64_t mfma_features;   /* features in the dense MFMA head (0: none) */
  int64_t dense_rows;      /* 256-byte dense head records fetched by the scan */
  int64_t tile_pairs;      /* band x tile pairs the scan visited */
  int64_t bitmap_features; /* features in the bitmap tier of the tail */
  int64_t band_entries;    /* band-side bitmap-tier entries, loaded once per work item (they
                              stay in registers and LDS while the item's tiles pass by) */
  int64_t band_records;    /* band-side cursor records, loaded once per work item */
  int64_t walk_entries;    /* cursor-list positions read by the counting walks */
  int64_t walk_scatter_entries; /* cursor-list entries (position + value) read by the
                              scatter walks */
  int64_t seg_row_bytes;   /* bytes of one J tile's segment-start row, fetched per tile pair */
  int64_t tile_pairs_skipped; /* tile pairs of sparse bands (rows with rare features only)
                              skipped: the two blocks share no feature */
  int64_t tile_pairs_pruned;  /* tile pairs of sparse bands whose block bound (L1 x max of the
                              blocks, or the sum of their products) stays below t */
  /* timings in milliseconds (hipEvent on the run's stream) */
  double ms_upload, ms_index, ms_scan, ms_rescore, ms_download, ms_total;
  double ms_tune;          /* autotuner (p < 0): time spent timing candidate configurations */
  int32_t band_rows;       /* height cap of the bands (autotuned or 128) */
  int32_t tile_rows;       /* height cap of the tiles (autotuned or 64) */
  int32_t items_per_cu;    /* scan work items per CU (autotuned or 8) */
  int32_t graph_reused;    /* batched runs: 1 if the device job ran in the library's kept
                              graph with updated arguments (0: the graph was built) */
  double ms_queue;         /* server requests: submit -> start of the launch that ran it */
  double ms_handover;      /* server requests: end of the launch -> result ready (split of a
                              packed batch's pairs by request) */
  void* internal_;
} apss_result;

typedef void (*apss_emit_fn)(const apss_pair* pairs, size_t count, void* user);

const char* apss_version(void);
const char* apss_strerror(int code);
/* Detail message of the last failed call on this thread. */
const char* apss_last_error(void);
void apss_opts_init(apss_opts* o);

/* Self-join APSS of X at threshold t with head/tail split p.  p < 0 autotunes:
 * the split is 0 and the tail's tiering (dense MFMA head on / off, bitmap tier
 * size) is the fastest of the candidates timed on a sampled count-only scan.
 * The pair set does not depend on the choice.
 * Replaces similarity_join_device (similarity_search.h:3-14) for one matrix. */
int apss_run(const apss_csr* X, double t, int32_t p, const apss_opts* o, apss_result* out);

/* Streaming variant: pairs are handed to emit() in chunks (up to 2^20 pairs)
 * on the calling thread instead of being returned; out (optional) receives the
 * counters.  The pairs stay on the device until their chunk is copied, so the
 * host never holds more than one chunk; opts.output_on_device is ignored. */
int apss_run_cb(const apss_csr* X, double t, int32_t p, const apss_opts* o, apss_emit_fn emit,
                void* user, apss_result* out);

/* Many independent instances in one batched launch sequence (one stream,
 * one tile kernel over all instances' tiles).  out[k] receives instance k. */
int apss_run_batched(const apss_csr* const* X, int32_t count, const double* t, int32_t p,
                     const apss_opts* o, apss_result* out);

void apss_result_free(apss_result* r);

/* Inverted index of X: the CSR -> CSC transpose of the features >= p, built on
 * the device (column histogram, scan, scatter, row sort of every list).
 * col_ptr[n_cols + 1] (features < p get empty lists), row[m] and val[m] are
 * caller-owned host arrays, m = the nonzeros with id >= p; rows ascend in each
 * list.  Replaces generate_index_ptr4 + the host scan + construct_Index2
 * (similarity_search.cu:105-191, 1091-1263, 221-274), whose lists come out in
 * arbitrary order. */
int apss_transpose(const apss_csr* X, int32_t p, const apss_opts* o, int64_t* col_ptr, int32_t* row,
                   float* val);

/* Device scratch of apss_run_noalloc, per GPU (MiB). */
#define APSS_NOALLOC_SCRATCH_MIB 512

/* apss_run without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  X must be a
 * device CSR (on_device = 1); the pairs go to `pairs`, a device buffer of
 * `capacity` records the caller owns; *count receives the number of result
 * pairs.  Every other device structure lives in one scratch array of
 * APSS_NOALLOC_SCRATCH_MIB MiB per GPU, held by the library until the process
 * ends; an input whose structures do not fit returns APSS_E_NOMEM (the limit is
 * about 1e5 rows of 40 nonzeros).  A result larger than `capacity` returns
 * APSS_E_NOMEM with *count set to the room it needs (the buffer then holds an
 * arbitrary `capacity` of its pairs).  Calls on one GPU share its array and are
 * serialised by the library.  p >= 0 (no autotuning), no sorting; out
 * (optional) receives the counters, with pairs = NULL.  Same pair set and
 * score bounds as apss_run. */
int apss_run_noalloc(const apss_csr* X, double t, int32_t p, const apss_opts* o, apss_pair* pairs,
                     int64_t capacity, int64_t* count, apss_result* out);

/* Serving: a corpus index resident in device memory. */
typedef struct apss_index apss_index;
int apss_index_create(const apss_csr* corpus, int32_t p, const apss_opts* o, apss_index** out);
/* Append rows to the corpus; they get ids n_rows_before .. n_rows_before+k-1. */
int apss_index_insert(apss_index* ix, const apss_csr* rows);
/* All (query, corpus) pairs with score >= t; safe to call from several threads. */
int apss_index_query(apss_index* ix, const apss_csr* queries, double t, apss_result* out);
int64_t apss_index_rows(const apss_index* ix);
int apss_index_destroy(apss_index* ix);

/* Continuous-batching server over an index: requests are admitted by memory
 * budget and run concurrently on `slots` streams. */
typedef struct apss_server apss_server;
int apss_server_create(apss_index* ix, int32_t slots, int64_t mem_budget, int32_t max_queue,
                       apss_server** out);
/* Enqueue a query batch (copied); returns a ticket >= 0 or a negative code. */
int64_t apss_server_submit(apss_server* s, const apss_csr* queries, double t);
/* Wait for a ticket; fills out and the request's latency (submit->done, ms). */
int apss_server_wait(apss_server* s, int64_t ticket, apss_result* out, double* latency_ms);
int apss_server_cancel(apss_server* s, int64_t ticket);
int apss_server_destroy(apss_server* s);

/* Drop-in for the reference's only entry point (similarity_search.h:3-14):
 * the same 70 arguments (8 SELL-32 partitions, their head norms, sizes and
 * the three lookup tables).  The SELL-32 slices are converted to one CSR
 * matrix (padding -1 dropped, rows in partition order A..H), then apss_run
 * runs at the reference's threshold t = 0.3f (similarity_search.cu:18) with
 * p = 0 (the pair set does not depend on p; the reference's literal 4096 only
 * moves work into its slow pair-parallel head dot).  Like the reference it
 * prints a summary and returns nothing; the pair count of the last call is
 * available from apss_compat_last_count(). */
void similarity_join_device(int* dim_A, float* feature_A, int* tile_start_A, int* dim_B,
                            float* feature_B, int* tile_start_B, int* dim_C, float* feature_C,
                            int* tile_start_C, int* dim_D, float* feature_D, int* tile_start_D,
                            int* dim_E, float* feature_E, int* tile_start_E, int* dim_F,
                            float* feature_F, int* tile_start_F, int* dim_G, float* feature_G,
                            int* tile_start_G, int* dim_H, float* feature_H, int* tile_start_H,
                            float* pre_norm_A1, float* pre_norm_B1, float* pre_norm_C1,
                            float* pre_norm_D1, float* pre_norm_E1, float* pre_norm_F1,
                            float* pre_norm_G1, float* pre_norm_H1, int sum_vec_A, int sum_feat_A,
                            int sum_tile_A, int sum_vec_B, int sum_feat_B, int sum_tile_B,
                            int sum_vec_C, int sum_feat_C, int sum_tile_C, int sum_vec_D,
                            int sum_feat_D, int sum_tile_D, int sum_vec_E, int sum_feat_E,
                            int sum_tile_E, int sum_vec_F, int sum_feat_F, int sum_tile_F,
                            int sum_vec_G, int sum_feat_G, int sum_tile_G, int sum_vec_H,
                            int sum_feat_H, int sum_tile_H, int* b_id_a, int* apss_block_seg,
                            int* apss_seg_start);
int64_t apss_compat_last_count(void);

/* ---- Files (host only; no device calls) ----------------------------------
 * Text rows in the reference's input format: one row per line, a leading
 * separator character, then "d v" pairs up to the line end.  Replaces the
 * per-partition reading loops of main.cu:286-733 (which fill SELL-32 arrays;
 * convert with any CSR->SELL step the caller already has, or call apss_run).
 * Rows keep file order; ids are sorted within a row, a repeated id is
 * APSS_E_INVALID.  n_cols <= 0 takes max id + 1.  *out owns malloc'd arrays:
 * release with apss_csr_free. */
int apss_load_text(const char* path, int64_t n_cols, apss_csr* out);
/* Writes a host CSR in the same text format (' ' leading separator). */
int apss_save_text(const char* path, const apss_csr* X);
/* Binary CSR: "APSSCSR1", int64 n_rows, n_cols, nnz, row_ptr[n_rows+1]
 * (int64), col[nnz] (int32), val[nnz] (float32), little-endian.  No
 * counterpart in the reference (its only input is the text file). */
int apss_save_binary(const char* path, const apss_csr* X);
int apss_load_binary(const char* path, apss_csr* out);
void apss_csr_free(apss_csr* X);
/* Pair file: one "i j score" line per pair, in the given order. */
int apss_write_pairs(const char* path, const apss_pair* pairs, int64_t count);
/* Compares the pair sets of two pair files (first two columns; (i, j) and
 * (j, i) are the same pair).  Returns the number of pairs found in only one
 * of them (0: same set) or a negative error code; only_a / only_b (optional)
 * receive the two one-sided counts.  Replaces compare() (main.cu:143-171),
 * which compares the first 22000 lines of two files line by line. */
int64_t apss_compare_pair_files(const char* a, const char* b, int64_t* only_a, int64_t* only_b);

#ifdef __cplusplus
}
#endif
#endif
