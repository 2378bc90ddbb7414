/*
 * Deterministic synthetic sparse-matrix generator for APSS workloads.
 *
 * The reference reads one fixed tf-idf text file (main.cu:270) whose rows are
 * L2-normalised sparse vectors with feature ids ascending inside a row
 * (main.cu:301-323).  This generator produces CSR matrices of the same kind:
 *
 *   - row lengths: truncated discrete power law P(k) ~ k^-alpha on [kmin,kmax]
 *     (alpha can be solved from a target mean with apss_gen_alpha);
 *   - feature ids: Zipf(s) over [0,d), id 0 the most frequent, drawn without
 *     repetition inside a row, then sorted ascending;
 *   - values: uniform(0,1] (1-U[0,1)) or lognormal(0,1) (exp(N(0,1))),
 *     L2-normalised in double precision and stored as float.
 *
 * Every random number is a pure function of (seed, row, purpose, counter)
 * (SplitMix64 finaliser over a per-row stream key), so the matrix does not
 * depend on the number of threads or on which rows are generated together.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "apss_gen.h"

static inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static inline uint64_t stream_key(uint64_t seed, int64_t row, uint64_t purpose) {
  return mix64(seed ^ mix64((uint64_t)row * 4u + purpose + 0x5851F42D4C957F2DULL));
}

static inline uint64_t draw(uint64_t key, uint64_t ctr) {
  return mix64(key + ctr * 0xD1B54A32D192ED03ULL);
}

static inline double u01(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

double apss_gen_alpha(int32_t kmin, int32_t kmax, double mean) {
  if (kmin < 1 || kmax < kmin) return -1.0;
  if (mean <= kmin) return 64.0;
  if (mean >= kmax) return 0.0;
  double lo = 0.0, hi = 8.0;
  for (int it = 0; it < 200; ++it) {
    double a = 0.5 * (lo + hi), sw = 0.0, swk = 0.0;
    for (int32_t k = kmin; k <= kmax; ++k) {
      double w = pow((double)k, -a);
      sw += w;
      swk += w * k;
    }
    if (swk / sw > mean) lo = a; else hi = a;
  }
  return 0.5 * (lo + hi);
}

/* Cumulative table of the truncated power law, normalised to 1. */
static double* length_cdf(const apss_gen_spec* s) {
  int32_t m = s->kmax - s->kmin + 1;
  double* cdf = (double*)malloc(sizeof(double) * (size_t)m);
  if (!cdf) return NULL;
  double acc = 0.0;
  for (int32_t i = 0; i < m; ++i) {
    acc += pow((double)(s->kmin + i), -s->alpha);
    cdf[i] = acc;
  }
  for (int32_t i = 0; i < m; ++i) cdf[i] /= acc;
  cdf[m - 1] = 1.0;
  return cdf;
}

static int32_t row_length(const apss_gen_spec* s, const double* cdf, int64_t row) {
  double u = u01(draw(stream_key(s->seed, row, 0), 0));
  int32_t lo = 0, hi = s->kmax - s->kmin;
  while (lo < hi) { /* first index with cdf >= u */
    int32_t mid = (lo + hi) >> 1;
    if (cdf[mid] >= u) hi = mid; else lo = mid + 1;
  }
  int64_t k = (int64_t)s->kmin + lo;
  if (k > s->d) k = s->d;
  return (int32_t)k;
}

int64_t apss_gen_rowptr(const apss_gen_spec* s, int64_t* row_ptr) {
  if (!s || !row_ptr || s->n < 0 || s->d < 1 || s->kmin < 1 || s->kmax < s->kmin) return -1;
  double* cdf = length_cdf(s);
  if (!cdf) return -1;
  row_ptr[0] = 0;
  /* lengths in parallel, prefix sum serially */
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < s->n; ++i) row_ptr[i + 1] = row_length(s, cdf, s->row_offset + i);
  for (int64_t i = 0; i < s->n; ++i) row_ptr[i + 1] += row_ptr[i];
  free(cdf);
  return row_ptr[s->n];
}

/* Vose alias table for Zipf(s) over [0,d): weight of id r is (r+1)^-s. */
typedef struct {
  double* prob;
  int32_t* alias;
  int32_t d;
} alias_table;

static int alias_build(alias_table* t, int32_t d, double zs) {
  t->d = d;
  t->prob = (double*)malloc(sizeof(double) * (size_t)d);
  t->alias = (int32_t*)malloc(sizeof(int32_t) * (size_t)d);
  double* p = (double*)malloc(sizeof(double) * (size_t)d);
  int32_t* small = (int32_t*)malloc(sizeof(int32_t) * (size_t)d);
  int32_t* large = (int32_t*)malloc(sizeof(int32_t) * (size_t)d);
  if (!t->prob || !t->alias || !p || !small || !large) {
    free(t->prob); free(t->alias); free(p); free(small); free(large);
    return -1;
  }
  double sum = 0.0;
  for (int32_t r = 0; r < d; ++r) { p[r] = pow((double)r + 1.0, -zs); sum += p[r]; }
  int32_t ns = 0, nl = 0;
  for (int32_t r = 0; r < d; ++r) {
    p[r] = p[r] * (double)d / sum;
    if (p[r] < 1.0) small[ns++] = r; else large[nl++] = r;
  }
  while (ns > 0 && nl > 0) {
    int32_t a = small[--ns], b = large[--nl];
    t->prob[a] = p[a];
    t->alias[a] = b;
    p[b] = (p[b] + p[a]) - 1.0;
    if (p[b] < 1.0) small[ns++] = b; else large[nl++] = b;
  }
  while (nl > 0) { int32_t b = large[--nl]; t->prob[b] = 1.0; t->alias[b] = b; }
  while (ns > 0) { int32_t a = small[--ns]; t->prob[a] = 1.0; t->alias[a] = a; }
  free(p); free(small); free(large);
  return 0;
}

static inline int32_t alias_draw(const alias_table* t, uint64_t r1, uint64_t r2) {
  int32_t c = (int32_t)(((r1 >> 32) * (uint64_t)t->d) >> 32);
  return u01(r2) < t->prob[c] ? c : t->alias[c];
}

static int cmp_i32(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

/* Open-addressing set of int32 ids; capacity is a power of two > 2k. */
static int set_insert(int32_t* tab, uint32_t mask, int32_t v) {
  uint32_t h = (uint32_t)v * 0x9E3779B1u;
  for (;;) {
    h &= mask;
    if (tab[h] == v) return 0;
    if (tab[h] < 0) { tab[h] = v; return 1; }
    ++h;
  }
}

static void fill_row(const apss_gen_spec* s, const alias_table* at, int64_t row, int32_t k,
                     int32_t* col, float* val, int32_t* tab, uint32_t cap) {
  memset(tab, 0xff, sizeof(int32_t) * cap);
  uint64_t kid = stream_key(s->seed, row, 1), kval = stream_key(s->seed, row, 2);
  int32_t got = 0;
  uint64_t ctr = 0, limit = 100u * (uint64_t)k + 1000u;
  while (got < k && ctr < limit) {
    int32_t c = alias_draw(at, draw(kid, 2 * ctr), draw(kid, 2 * ctr + 1));
    ++ctr;
    if (set_insert(tab, cap - 1, c)) col[got++] = c;
  }
  /* deterministic completion when rejection stalls (k close to d) */
  for (int32_t c = 0; got < k && c < s->d; ++c)
    if (set_insert(tab, cap - 1, c)) col[got++] = c;
  qsort(col, (size_t)k, sizeof(int32_t), cmp_i32);
  double ss = 0.0;
  double* tmp = (double*)tab; /* reuse: cap ints >= 2k ints >= k doubles */
  for (int32_t e = 0; e < k; ++e) {
    double v;
    if (s->values == APSS_GEN_LOGNORMAL) {
      double u1 = 1.0 - u01(draw(kval, 2 * (uint64_t)e));
      double u2 = u01(draw(kval, 2 * (uint64_t)e + 1));
      v = exp(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
    } else {
      v = 1.0 - u01(draw(kval, 2 * (uint64_t)e));
    }
    tmp[e] = v;
    ss += v * v;
  }
  double inv = 1.0 / sqrt(ss);
  for (int32_t e = 0; e < k; ++e) val[e] = (float)(tmp[e] * inv);
}

int apss_gen_fill(const apss_gen_spec* s, const int64_t* row_ptr, int32_t* col, float* val) {
  if (!s || !row_ptr || !col || !val) return -1;
  alias_table at;
  if (alias_build(&at, s->d, s->zipf_s) != 0) return -1;
  int err = 0;
#pragma omp parallel
  {
    int32_t* tab = NULL;
    uint32_t tab_cap = 0;
#pragma omp for schedule(dynamic, 256)
    for (int64_t i = 0; i < s->n; ++i) {
      int32_t k = (int32_t)(row_ptr[i + 1] - row_ptr[i]);
      if (k == 0) continue;
      uint32_t cap = 16;
      while (cap < 4u * (uint32_t)k) cap <<= 1;
      if (cap > tab_cap) {
        free(tab);
        tab = (int32_t*)malloc(sizeof(int32_t) * cap);
        tab_cap = tab ? cap : 0;
        if (!tab) { err = -1; continue; }
      }
      fill_row(s, &at, s->row_offset + i, k, col + row_ptr[i], val + row_ptr[i], tab, tab_cap);
    }
    free(tab);
  }
  free(at.prob);
  free(at.alias);
  return err;
}
