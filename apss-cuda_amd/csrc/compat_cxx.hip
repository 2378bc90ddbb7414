// similarity_join_device with C++ linkage.
//
// The reference declares its entry point without extern "C"
// (similarity_search.h:3-14), so a caller compiled against that header
// (main.cu:11 includes it, main.cu:754-765 calls it) references the mangled
// C++ symbol.  This translation unit defines exactly that function and
// forwards to the shim (compat.hip); it must not see apss.h, whose
// extern "C" declaration of the same name would conflict.
#include "compat.h"

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
                             feature_C, tile_start_C, dim_D, feature_D, tile_start_D, dim_E,
                             feature_E, tile_start_E, dim_F, feature_F, tile_start_F, dim_G,
                             feature_G, tile_start_G, dim_H, feature_H, tile_start_H, pre_norm_A1,
                             pre_norm_B1, pre_norm_C1, pre_norm_D1, pre_norm_E1, pre_norm_F1,
                             pre_norm_G1, pre_norm_H1, sum_vec_A, sum_feat_A, sum_tile_A,
                             sum_vec_B, sum_feat_B, sum_tile_B, sum_vec_C, sum_feat_C, sum_tile_C,
                             sum_vec_D, sum_feat_D, sum_tile_D, sum_vec_E, sum_feat_E, sum_tile_E,
                             sum_vec_F, sum_feat_F, sum_tile_F, sum_vec_G, sum_feat_G, sum_tile_G,
                             sum_vec_H, sum_feat_H, sum_tile_H, b_id_a, apss_block_seg,
                             apss_seg_start);
}
