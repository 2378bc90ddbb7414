/*
 * APSS oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * may load this library, and only as the checker or the timed CPU baseline.
 * The product (libapss.so) never links or calls it.
 *
 * Functions and what pins each one (see tests/test_oracle.py):
 *
 *  oracle_transpose   textbook stable counting-sort CSR->CSC of the features
 *                     d >= p (the reference's inverted index of the tail,
 *                     similarity_search.cu:105-191, 221-274, 1168-1186).
 *                     Pinned against scipy.sparse csr.tocsc() and the golden
 *                     fixtures in tests/golden/.
 *  oracle_apss_fp64   exact all-pairs search: pairs (i<j) with
 *                     s = sum_d x_id*x_jd >= t, products exact in double and
 *                     summed from 0.0 in ascending feature order.  Pinned
 *                     against a dense float64 Gram matrix (numpy matmul) on
 *                     small inputs, closed forms (identical rows, disjoint
 *                     supports, one shared feature) and brute force.
 *  oracle_query_fp64  the same for query rows against corpus rows (every
 *                     (q,c) pair, no i<j restriction).  Pinned like
 *                     oracle_apss_fp64 (dense float64 Q*C^T).
 *  oracle_hybrid_fp32 CPU restatement of the reference hybrid method
 *                     (tail inverted index d>=p, FP32 tail partials,
 *                     prefix-norm filter, FP32 head forward-list dot with a
 *                     dense head vector, threshold).  Follows
 *                     similarity_search.cu:370-437/642-735 (tail partials),
 *                     :510-531/:833-852 (filter), :441-506/:854-915 (head dot)
 *                     and main.cu:308-325 (head norm), without the defects the
 *                     reference carries (uninitialised accumulators, staging
 *                     stride 1024 vs 512 threads, half-covered Y range, prune
 *                     on `<= t`).  Pinned against oracle_apss_fp64: same pair
 *                     set except pairs within FP32 rounding of t.
 *  oracle_inverted_search_fp32
 *                     serial inverted-index search (main.cu:104-141): for
 *                     each query row, every entry of the inverted list of
 *                     each of its features (all rows, no split, no filter)
 *                     adds vec_fea * x to an FP32 accumulator, the product
 *                     and sum formed in double and rounded once (Term::val
 *                     is float, the row's value and the sum are double,
 *                     the accumulator std::map<int,float>); pairs >= t are
 *                     counted (rows in the given order until budget_s
 *                     seconds have passed, if budget_s > 0).  A dense accumulator with a touched list
 *                     replaces std::map (same sums, same order).  Pinned
 *                     against oracle_apss_fp64: same count except pairs
 *                     within FP32 rounding of t.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#include "oracle.h"

/* ------------------------------------------------------------------------ */
/* transpose                                                                 */

int oracle_transpose(int64_t n, int64_t d, const int64_t* rp, const int32_t* col, const float* val,
                     int32_t p, int64_t* cptr, int32_t* crow, float* cval) {
  if (n < 0 || d < 0 || p < 0) return -1;
  memset(cptr, 0, sizeof(int64_t) * (size_t)(d + 1));
  for (int64_t i = 0; i < n; ++i)
    for (int64_t e = rp[i]; e < rp[i + 1]; ++e)
      if (col[e] >= p) cptr[col[e] + 1]++;
  for (int64_t c = 0; c < d; ++c) cptr[c + 1] += cptr[c];
  int64_t* cur = (int64_t*)malloc(sizeof(int64_t) * (size_t)(d + 1));
  if (!cur) return -1;
  memcpy(cur, cptr, sizeof(int64_t) * (size_t)(d + 1));
  for (int64_t i = 0; i < n; ++i)
    for (int64_t e = rp[i]; e < rp[i + 1]; ++e)
      if (col[e] >= p) {
        int64_t q = cur[col[e]]++;
        crow[q] = (int32_t)i;
        cval[q] = val[e];
      }
  free(cur);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* growable per-thread pair buffers                                          */

typedef struct {
  int64_t n, cap;
  uint32_t *i, *j;
  double* s;
} pbuf;

static int pbuf_push(pbuf* b, uint32_t i, uint32_t j, double s) {
  if (b->n == b->cap) {
    int64_t nc = b->cap ? 2 * b->cap : 1024;
    uint32_t* ni = (uint32_t*)realloc(b->i, sizeof(uint32_t) * (size_t)nc);
    if (ni) b->i = ni;
    uint32_t* nj = (uint32_t*)realloc(b->j, sizeof(uint32_t) * (size_t)nc);
    if (nj) b->j = nj;
    double* ns = (double*)realloc(b->s, sizeof(double) * (size_t)nc);
    if (ns) b->s = ns;
    if (!ni || !nj || !ns) return -1;
    b->cap = nc;
  }
  b->i[b->n] = i;
  b->j[b->n] = j;
  b->s[b->n] = s;
  b->n++;
  return 0;
}

/* Concatenate per-row-block buffers (already in row order) into out. */
static int gather(pbuf* bufs, int nb, oracle_pairs* out) {
  int64_t tot = 0;
  for (int b = 0; b < nb; ++b) tot += bufs[b].n;
  out->count = tot;
  out->i = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(tot ? tot : 1));
  out->j = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(tot ? tot : 1));
  out->s = (double*)malloc(sizeof(double) * (size_t)(tot ? tot : 1));
  int rc = (out->i && out->j && out->s) ? 0 : -1;
  int64_t o = 0;
  for (int b = 0; b < nb; ++b) {
    if (rc == 0 && bufs[b].n) {
      memcpy(out->i + o, bufs[b].i, sizeof(uint32_t) * (size_t)bufs[b].n);
      memcpy(out->j + o, bufs[b].j, sizeof(uint32_t) * (size_t)bufs[b].n);
      memcpy(out->s + o, bufs[b].s, sizeof(double) * (size_t)bufs[b].n);
    }
    o += bufs[b].n;
    free(bufs[b].i); free(bufs[b].j); free(bufs[b].s);
  }
  free(bufs);
  return rc;
}

void oracle_pairs_free(oracle_pairs* p) {
  if (!p) return;
  free(p->i); free(p->j); free(p->s);
  p->i = p->j = NULL;
  p->s = NULL;
  p->count = 0;
}

static int cmp_u32(const void* a, const void* b) {
  uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
  return (x > y) - (x < y);
}

/* ------------------------------------------------------------------------ */
/* exact FP64 search                                                         */

/*
 * Scores of one query row against rows of a CSC (crow/cval by column) whose
 * row index is > jmin.  Each acc[j] receives its terms in the order of the
 * query's features, i.e. ascending feature id, starting from 0.0.
 */
static int row_search(const int32_t* qcol, const float* qval, int64_t qlen, const int64_t* cptr,
                      const int32_t* crow, const float* cval, int64_t jmin, double t, double* acc,
                      uint8_t* seen, uint32_t* touched, uint32_t qid, pbuf* out) {
  int64_t nt = 0;
  for (int64_t e = 0; e < qlen; ++e) {
    int32_t c = qcol[e];
    double x = (double)qval[e];
    int64_t lo = cptr[c], hi = cptr[c + 1];
    if (jmin >= 0) { /* skip rows <= jmin: lists are ascending in row */
      int64_t a = lo, b = hi;
      while (a < b) {
        int64_t m = (a + b) >> 1;
        if (crow[m] <= jmin) a = m + 1; else b = m;
      }
      lo = a;
    }
    for (int64_t q = lo; q < hi; ++q) {
      uint32_t j = (uint32_t)crow[q];
      if (!seen[j]) { seen[j] = 1; acc[j] = 0.0; touched[nt++] = j; }
      acc[j] += x * (double)cval[q];
    }
  }
  qsort(touched, (size_t)nt, sizeof(uint32_t), cmp_u32);
  int rc = 0;
  for (int64_t k = 0; k < nt; ++k) {
    uint32_t j = touched[k];
    if (acc[j] >= t && rc == 0) rc = pbuf_push(out, qid, j, acc[j]);
    seen[j] = 0;
  }
  return rc;
}

static int search_fp64(int64_t nq, const int64_t* qrp, const int32_t* qcol, const float* qval,
                       int64_t nc, const int64_t* crp, const int32_t* ccol, const float* cval_rows,
                       int64_t d, double t, int self, oracle_pairs* out) {
  int64_t nnz = crp[nc];
  int64_t* cptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(d + 1));
  int32_t* crow = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
  float* cval = (float*)malloc(sizeof(float) * (size_t)(nnz ? nnz : 1));
  if (!cptr || !crow || !cval) { free(cptr); free(crow); free(cval); return -1; }
  oracle_transpose(nc, d, crp, ccol, cval_rows, 0, cptr, crow, cval);
  const int64_t blk = 64;
  int nb = (int)((nq + blk - 1) / blk);
  pbuf* bufs = (pbuf*)calloc((size_t)(nb ? nb : 1), sizeof(pbuf));
  int err = bufs ? 0 : -1;
#pragma omp parallel
  {
    double* acc = (double*)malloc(sizeof(double) * (size_t)(nc ? nc : 1));
    uint8_t* seen = (uint8_t*)calloc((size_t)(nc ? nc : 1), 1);
    uint32_t* touched = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(nc ? nc : 1));
    if (!acc || !seen || !touched) err = -1;
#pragma omp for schedule(dynamic, 1)
    for (int b = 0; b < nb; ++b) {
      if (err) continue;
      for (int64_t i = b * blk; i < nq && i < (b + 1) * blk; ++i)
        if (row_search(qcol + qrp[i], qval + qrp[i], qrp[i + 1] - qrp[i], cptr, crow, cval,
                       self ? i : -1, t, acc, seen, touched, (uint32_t)i, &bufs[b]))
          err = -1;
    }
    free(acc); free(seen); free(touched);
  }
  free(cptr); free(crow); free(cval);
  if (!bufs) return -1;
  int rc = gather(bufs, nb, out);
  return err ? -1 : rc;
}

int oracle_apss_fp64(int64_t n, int64_t d, const int64_t* rp, const int32_t* col, const float* val,
                     double t, oracle_pairs* out) {
  return search_fp64(n, rp, col, val, n, rp, col, val, d, t, 1, out);
}

int oracle_query_fp64(int64_t nq, const int64_t* qrp, const int32_t* qcol, const float* qval,
                      int64_t nc, const int64_t* crp, const int32_t* ccol, const float* cval,
                      int64_t d, double t, oracle_pairs* out) {
  return search_fp64(nq, qrp, qcol, qval, nc, crp, ccol, cval, d, t, 0, out);
}

/* Exact FP64 score of one pair, ascending feature order (the rescore rule). */
double oracle_pair_fp64(const int32_t* ac, const float* av, int64_t na, const int32_t* bc,
                        const float* bv, int64_t nb) {
  double s = 0.0;
  int64_t x = 0, y = 0;
  while (x < na && y < nb) {
    if (ac[x] < bc[y]) ++x;
    else if (ac[x] > bc[y]) ++y;
    else { s += (double)av[x] * (double)bv[y]; ++x; ++y; }
  }
  return s;
}

/* ------------------------------------------------------------------------ */
/* reference hybrid (FP32), the CPU baseline                                 */

int oracle_hybrid_fp32(int64_t n, int64_t d, const int64_t* rp, const int32_t* col,
                       const float* val, double t, int32_t p, const int64_t* rows, int64_t nrows,
                       oracle_pairs* out, oracle_hybrid_stats* st) {
  if (p < 0 || p > d) return -1;
  int64_t nnz = rp[n];
  /* tail inverted index (d >= p) -- similarity_search.cu:105-191, 221-274 */
  int64_t* cptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(d + 1));
  int32_t* crow = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
  float* cval = (float*)malloc(sizeof(float) * (size_t)(nnz ? nnz : 1));
  /* prefix (head) norm in double stored as float -- main.cu:308-312, 324-325 */
  float* hn = (float*)malloc(sizeof(float) * (size_t)(n ? n : 1));
  if (!cptr || !crow || !cval || !hn) { free(cptr); free(crow); free(cval); free(hn); return -1; }
  oracle_transpose(n, d, rp, col, val, p, cptr, crow, cval);
  for (int64_t i = 0; i < n; ++i) {
    double s = 0.0;
    for (int64_t e = rp[i]; e < rp[i + 1] && col[e] < p; ++e) s += pow((double)val[e], 2);
    hn[i] = (float)sqrt(s);
  }
  const float tf = (float)t;
  int64_t nq = rows ? nrows : n;
  const int64_t blk = 16;
  int nb = (int)((nq + blk - 1) / blk);
  pbuf* bufs = (pbuf*)calloc((size_t)(nb ? nb : 1), sizeof(pbuf));
  int64_t cand = 0, surv = 0, gathers = 0;
  int err = bufs ? 0 : -1;
  const double t_search0 = omp_get_wtime();
#pragma omp parallel reduction(+ : cand, surv, gathers)
  {
    float* acc = (float*)calloc((size_t)(n ? n : 1), sizeof(float));
    float* head = (float*)calloc((size_t)(p ? p : 1), sizeof(float));
    if (!acc || !head) err = -1;
#pragma omp for schedule(dynamic, 1)
    for (int b = 0; b < nb; ++b) {
      if (err) continue;
      for (int64_t q = b * blk; q < nq && q < (b + 1) * blk; ++q) {
        int64_t i = rows ? rows[q] : q;
        /* tail partials (FP32) for partners j > i -- similarity_search.cu:642-735 */
        for (int64_t e = rp[i]; e < rp[i + 1]; ++e) {
          int32_t c = col[e];
          if (c < p) continue;
          float x = val[e];
          for (int64_t k = cptr[c + 1] - 1; k >= cptr[c] && crow[k] > i; --k)
            acc[crow[k]] += x * cval[k];
        }
        /* dense head vector of row i -- similarity_search.cu:862-878 */
        for (int64_t e = rp[i]; e < rp[i + 1] && col[e] < p; ++e) head[col[e]] = val[e];
        for (int64_t j = i + 1; j < n; ++j) {
          /* prefix-norm filter, prune when the bound is below t
           * (the reference prunes on <= t, similarity_search.cu:844) */
          float part = acc[j];
          acc[j] = 0.0f;
          ++cand;
          if (part + hn[i] * hn[j] < tf) continue;
          ++surv;
          /* head forward-list dot -- similarity_search.cu:895-905 */
          float rest = 0.0f;
          for (int64_t e = rp[j]; e < rp[j + 1]; ++e) {
            int32_t c = col[e];
            if (c >= p) break;
            rest += val[e] * head[c];
            ++gathers;
          }
          if (rest + part >= tf && pbuf_push(&bufs[b], (uint32_t)i, (uint32_t)j, (double)(rest + part)))
            err = -1;
        }
        for (int64_t e = rp[i]; e < rp[i + 1] && col[e] < p; ++e) head[col[e]] = 0.0f;
      }
    }
    free(acc); free(head);
  }
  const double t_search = omp_get_wtime() - t_search0;
  free(cptr); free(crow); free(cval); free(hn);
  if (st) {
    st->candidates = cand; st->survivors = surv; st->head_gathers = gathers;
    st->search_seconds = t_search;
  }
  if (!bufs) return -1;
  int rc = gather(bufs, nb, out);
  return err ? -1 : rc;
}

/* ------------------------------------------------------------------------ */
/* serial inverted-index search (FP32), the second CPU data point            */

int oracle_inverted_search_fp32(int64_t n, int64_t d, const int64_t* rp, const int32_t* col,
                                const float* val, double t, const int64_t* rows, int64_t nrows,
                                double budget_s, int64_t* done, int64_t* count, double* seconds) {
  int64_t nnz = rp[n];
  int64_t* cptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(d + 1));
  int32_t* crow = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nnz ? nnz : 1));
  float* cval = (float*)malloc(sizeof(float) * (size_t)(nnz ? nnz : 1));
  float* acc = (float*)calloc((size_t)(n ? n : 1), sizeof(float));
  uint8_t* seen = (uint8_t*)calloc((size_t)(n ? n : 1), 1);
  uint32_t* touched = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
  int rc = (cptr && crow && cval && acc && seen && touched) ? 0 : -1;
  int64_t cnt = 0;
  double t0 = 0.0, t1 = 0.0;
  if (rc == 0) {
    oracle_transpose(n, d, rp, col, val, 0, cptr, crow, cval);
    const float tf = (float)t;
    t0 = omp_get_wtime();
    int64_t nq = rows ? nrows : n, q = 0;
    for (; q < nq; ++q) {
      if (budget_s > 0.0 && omp_get_wtime() - t0 > budget_s) break;
      int64_t i = rows ? rows[q] : q;
      int64_t nt = 0;
      for (int64_t e = rp[i]; e < rp[i + 1]; ++e) {
        const double x = (double)val[e];
        for (int64_t k = cptr[col[e]]; k < cptr[col[e] + 1]; ++k) {
          uint32_t j = (uint32_t)crow[k];
          if (!seen[j]) { seen[j] = 1; touched[nt++] = j; }
          acc[j] = (float)((double)acc[j] + (double)cval[k] * x);
        }
      }
      for (int64_t k = 0; k < nt; ++k) {
        uint32_t j = touched[k];
        cnt += (j > (uint64_t)i && acc[j] >= tf);
        acc[j] = 0.0f;
        seen[j] = 0;
      }
    }
    t1 = omp_get_wtime();
    if (done) *done = q;
  }
  free(cptr); free(crow); free(cval); free(acc); free(seen); free(touched);
  if (count) *count = cnt;
  if (seconds) *seconds = t1 - t0;
  return rc;
}

int oracle_num_threads(void) { return omp_get_max_threads(); }
