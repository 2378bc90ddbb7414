/* APSS oracle -- test infrastructure only; see oracle.c for the rules. */
#ifndef APSS_ORACLE_H
#define APSS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int64_t count;
  uint32_t* i;
  uint32_t* j;
  double* s;
} oracle_pairs;

typedef struct {
  int64_t candidates;   /* pairs the prefix-norm filter looked at */
  int64_t survivors;    /* pairs that passed it */
  int64_t head_gathers; /* head entries walked by the forward-list dot */
  double search_seconds; /* wall time of the search loop (index build excluded) */
} oracle_hybrid_stats;

int oracle_transpose(int64_t n, int64_t d, const int64_t* rp, const int32_t* col, const float* val,
                     int32_t p, int64_t* cptr, int32_t* crow, float* cval);
int oracle_apss_fp64(int64_t n, int64_t d, const int64_t* rp, const int32_t* col, const float* val,
                     double t, oracle_pairs* out);
int oracle_query_fp64(int64_t nq, const int64_t* qrp, const int32_t* qcol, const float* qval,
                      int64_t nc, const int64_t* crp, const int32_t* ccol, const float* cval,
                      int64_t d, double t, oracle_pairs* out);
double oracle_pair_fp64(const int32_t* ac, const float* av, int64_t na, const int32_t* bc,
                        const float* bv, int64_t nb);
/* rows: optional subset of query rows (NULL = all); each row i is joined with all j > i. */
int oracle_hybrid_fp32(int64_t n, int64_t d, const int64_t* rp, const int32_t* col,
                       const float* val, double t, int32_t p, const int64_t* rows, int64_t nrows,
                       oracle_pairs* out, oracle_hybrid_stats* st);
/* Serial (one thread) inverted-index search: counts pairs (i, j > i) with
 * FP32 score >= t for the query rows (NULL = all), taken in order until
 * budget_s seconds have passed (budget_s <= 0: all); *done = rows searched;
 * seconds excludes the index build. */
int oracle_inverted_search_fp32(int64_t n, int64_t d, const int64_t* rp, const int32_t* col,
                                const float* val, double t, const int64_t* rows, int64_t nrows,
                                double budget_s, int64_t* done, int64_t* count, double* seconds);
void oracle_pairs_free(oracle_pairs* p);
int oracle_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
