// apss_transpose: the inverted index of a matrix as a standalone operator.
//
// The reference builds its tail index with a per-partition column histogram
// (generate_index_ptr4, similarity_search.cu:105-191), a host exclusive scan
// with a pointer upload (:1091-1108, 1139-1226) and an atomic scatter
// (construct_Index2, :221-274), which leaves each list in arbitrary order.
// Here the same steps run on the device as the search's cursor-tier build
// (build_tiers with every feature >= p in the lists): LDS-privatised column
// histogram, host scan of the histogram (one device-to-host copy), scatter,
// and a row sort of every list (bitonic sorts in LDS, merge passes for lists
// longer than one LDS piece), so the result is the deterministic CSC with
// rows ascending in each list.
#include <cstring>

#include "engine.h"

namespace apss {

int run_transpose(const apss_csr* X, int32_t p, const apss_opts* o, int64_t* col_ptr, int32_t* row,
                  float* val) {
  apss_opts def;
  apss_opts_init(&def);
  if (!o) o = &def;
  if (!X || !col_ptr) throw Error{APSS_E_INVALID, "null argument"};
  check_csr_host(*X, p);
  StreamScope ss(o->stream, o->device);
  hipStream_t st = ss.s;
  Arena ar(o->device, o->mem_budget, &run_pool(o->device));
  ar.set_stream(st);
  Side s = prepare_side(ar, st, *X, p);
  Tiers tr = build_tiers(ar, st, s, p, 0, 0, DENSE_OFF, true);
  const int64_t d = s.d, m = tr.n_cursor_entries;
  if (m > 0 && (!row || !val)) throw Error{APSS_E_INVALID, "null output array"};
  std::vector<int2> ent((size_t)std::max<int64_t>(m, 1));
  APSS_HIP(hipMemcpyAsync(col_ptr, tr.sptr, sizeof(int64_t) * (size_t)(d + 1), hipMemcpyDeviceToHost, st));
  if (m > 0)
    APSS_HIP(hipMemcpyAsync(ent.data(), tr.sent, sizeof(int2) * (size_t)m, hipMemcpyDeviceToHost, st));
  APSS_HIP(hipStreamSynchronize(st));
  for (int64_t k = 0; k < m; ++k) {
    row[k] = ent[(size_t)k].x;
    std::memcpy(&val[k], &ent[(size_t)k].y, sizeof(float));
  }
  return APSS_OK;
}

}  // namespace apss
