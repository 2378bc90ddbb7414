/* Deterministic synthetic CSR generator (host library libapss_host.so). */
#ifndef APSS_GEN_H
#define APSS_GEN_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { APSS_GEN_UNIFORM = 0, APSS_GEN_LOGNORMAL = 1 };

typedef struct {
  int64_t n;          /* rows to generate */
  int64_t row_offset; /* global index of the first generated row (stream key) */
  int32_t d;          /* feature-id range [0,d) */
  int32_t kmin, kmax; /* row-length range */
  double alpha;       /* power-law exponent of the row length */
  double zipf_s;      /* Zipf exponent of feature ids (id 0 most frequent) */
  int32_t values;     /* APSS_GEN_UNIFORM or APSS_GEN_LOGNORMAL */
  int32_t pad_;
  uint64_t seed;
} apss_gen_spec;

/* Exponent of the truncated power law on [kmin,kmax] whose mean is `mean`. */
double apss_gen_alpha(int32_t kmin, int32_t kmax, double mean);
/* Fills row_ptr[0..n]; returns nnz (or -1 on bad spec). */
int64_t apss_gen_rowptr(const apss_gen_spec* spec, int64_t* row_ptr);
/* Fills col[nnz] (ascending per row, distinct) and val[nnz] (unit L2 per row). */
int apss_gen_fill(const apss_gen_spec* spec, const int64_t* row_ptr, int32_t* col, float* val);

#ifdef __cplusplus
}
#endif
#endif
