// Drop-in for the reference's only entry point, similarity_join_device
// (similarity_search.h:3-14; definition similarity_search.cu:934-3015).
//
// The reference caller (main.cu:754-765) hands over eight row partitions A..H
// in SELL-32 layout: rows grouped in slices of 32, each slice stored
// column-major with width = the nnz of its first row, padding id -1
// (main.cu:286-339).  This shim decodes the slices into one CSR matrix (rows in
// partition order), then runs apss_run with the reference's compile-time
// threshold 0.3f (similarity_search.cu:18).  The reference's literal split 4096
// only decides which features go through its pair-parallel head dot; the pair
// set is the same for every split, so the shim runs the split-0 path (dense
// MFMA head, bitmap and cursor tiers): on 50,000 rows of the large synthetic
// workload its scan takes 4.9 ms against 770 ms with the split at 4096.  The
// prefix norms,
// the triangular lookup table b_id_a and the head-grid tables are not needed
// (norms are recomputed on the device; indices are closed form) and are only
// accepted for signature compatibility.  Like the reference it prints a
// summary; unlike it, the pairs are really found: their count is kept for
// apss_compat_last_count().
#include <atomic>
#include <cstdio>
#include <vector>

#include "compat.h"
#include "engine.h"

namespace apss {

static std::atomic<int64_t> g_compat_count{-1};

constexpr int REF_MAX_DIM = 198456;  // feature range of the reference dataset (main.cu:17)
constexpr float REF_THRESHOLD = 0.3f;  // similarity_search.cu:18
// The reference's split is 4096; the pair set does not depend on it (see above).
constexpr int32_t SHIM_SPLIT = 0;

// Append the rows of one SELL-32 partition to a CSR under construction.
static void decode_sell32(const int* dim, const float* feat, const int* tile_start, int n_vec,
                          std::vector<int64_t>& rp, std::vector<int32_t>& col,
                          std::vector<float>& val, int& max_d) {
  for (int r = 0; r < n_vec; ++r) {
    const int tile = r / 32;
    const int width = (tile_start[tile + 1] - tile_start[tile]) / 32;
    const int base = tile_start[tile] + r % 32;
    for (int j = 0; j < width; ++j) {
      const int d = dim[base + j * 32];
      if (d < 0) continue;  // padding
      col.push_back(d);
      val.push_back(feat[base + j * 32]);
      if (d > max_d) max_d = d;
    }
    rp.push_back((int64_t)col.size());
  }
}

// The shim itself.  Two exported symbols forward here: the C-linkage
// similarity_join_device declared in apss.h and the C++-linkage one a caller
// compiled against the reference header binds (compat_cxx.hip).
void similarity_join_impl(int* dim_A, float* feature_A, int* tile_start_A, int* dim_B,
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
                          int* apss_seg_start) {
  (void)pre_norm_A1; (void)pre_norm_B1; (void)pre_norm_C1; (void)pre_norm_D1;
  (void)pre_norm_E1; (void)pre_norm_F1; (void)pre_norm_G1; (void)pre_norm_H1;
  (void)sum_feat_A; (void)sum_feat_B; (void)sum_feat_C; (void)sum_feat_D;
  (void)sum_feat_E; (void)sum_feat_F; (void)sum_feat_G; (void)sum_feat_H;
  (void)sum_tile_A; (void)sum_tile_B; (void)sum_tile_C; (void)sum_tile_D;
  (void)sum_tile_E; (void)sum_tile_F; (void)sum_tile_G; (void)sum_tile_H;
  (void)b_id_a; (void)apss_block_seg; (void)apss_seg_start;
  g_compat_count = -1;
  const int* dims[8] = {dim_A, dim_B, dim_C, dim_D, dim_E, dim_F, dim_G, dim_H};
  const float* feats[8] = {feature_A, feature_B, feature_C, feature_D,
                           feature_E, feature_F, feature_G, feature_H};
  const int* starts[8] = {tile_start_A, tile_start_B, tile_start_C, tile_start_D,
                          tile_start_E, tile_start_F, tile_start_G, tile_start_H};
  const int nvec[8] = {sum_vec_A, sum_vec_B, sum_vec_C, sum_vec_D,
                       sum_vec_E, sum_vec_F, sum_vec_G, sum_vec_H};
  std::vector<int64_t> rp(1, 0);
  std::vector<int32_t> col;
  std::vector<float> val;
  int max_d = -1;
  for (int k = 0; k < 8; ++k)
    if (nvec[k] > 0 && dims[k] && feats[k] && starts[k])
      decode_sell32(dims[k], feats[k], starts[k], nvec[k], rp, col, val, max_d);
  apss_csr X{(int64_t)rp.size() - 1, std::max<int64_t>(REF_MAX_DIM, (int64_t)max_d + 1),
             (int64_t)col.size(), rp.data(), col.data(), val.data(), 0, 0};
  apss_opts o;
  apss_opts_init(&o);
  o.output_on_device = 1;  // the caller never reads pairs back; keep them off the host
  apss_result r;
  const int rc = apss_run(&X, (double)REF_THRESHOLD, SHIM_SPLIT, &o, &r);
  if (rc != APSS_OK) {
    std::printf("similarity_join_device: %s (%s)\n", apss_strerror(rc), apss_last_error());
    return;
  }
  g_compat_count = r.count;
  std::printf("rows=%lld sum pair=%lld  (scan %.2f ms, total %.2f ms)\n", (long long)X.n_rows,
              (long long)r.count, r.ms_scan, r.ms_total);
  apss_result_free(&r);
}

}  // namespace apss

extern "C" {

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
                            int* apss_seg_start) {
  apss::similarity_join_impl(dim_A, feature_A, tile_start_A, dim_B, feature_B, tile_start_B, dim_C,
                             feature_C, tile_start_C, dim_D, feature_D, tile_start_D, dim_E, feature_E,
                             tile_start_E, dim_F, feature_F, tile_start_F, dim_G, feature_G,
                             tile_start_G, dim_H, feature_H, tile_start_H, pre_norm_A1, pre_norm_B1,
                             pre_norm_C1, pre_norm_D1, pre_norm_E1, pre_norm_F1, pre_norm_G1,
                             pre_norm_H1, sum_vec_A, sum_feat_A, sum_tile_A, sum_vec_B, sum_feat_B,
                             sum_tile_B, sum_vec_C, sum_feat_C, sum_tile_C, sum_vec_D, sum_feat_D,
                             sum_tile_D, sum_vec_E, sum_feat_E, sum_tile_E, sum_vec_F, sum_feat_F,
                             sum_tile_F, sum_vec_G, sum_feat_G, sum_tile_G, sum_vec_H, sum_feat_H,
                             sum_tile_H, b_id_a, apss_block_seg, apss_seg_start);
}

int64_t apss_compat_last_count(void) { return apss::g_compat_count.load(); }

}  // extern "C"
