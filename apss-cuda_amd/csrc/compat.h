// The drop-in shim's body, shared by the two exported forms of the
// reference's entry point: similarity_join_device with C linkage (apss.h,
// for ctypes / cgo-style callers) and with C++ linkage (compat_cxx.hip, the
// symbol a caller compiled against similarity_search.h:3-14 binds).  This
// header declares no exported name, so both translation units may include it.
#pragma once

namespace apss {

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
                          int* apss_seg_start);

}  // namespace apss
