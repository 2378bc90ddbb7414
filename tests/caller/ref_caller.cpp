// A caller written the way the reference's driver is (main.cu): it declares
// the entry point exactly as similarity_search.h:3-14 does -- plain C++
// linkage, no extern "C" -- builds eight SELL-32 partitions (slices of 32
// rows, column-major inside a slice, width = the slice's longest row,
// padding -1; main.cu:286-339) and calls the function once (main.cu:754-765).
// Linking this file against libapss.so shows that such a caller needs no
// change.  Input: a matrix in the reference's text format (one row per line,
// a leading separator character, then "id value" pairs).  Output: the pair
// count of the call at the reference threshold, as "count N".
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

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

extern "C" int64_t apss_compat_last_count(void);

struct Part {
  std::vector<int> dim, tile_start;
  std::vector<float> feat, pre_norm;
  int n = 0;
};

static Part make_part(const std::vector<std::vector<std::pair<int, float>>>& rows, size_t a, size_t b) {
  Part p;
  p.n = (int)(b - a);
  const int slices = (p.n + 31) / 32;
  p.tile_start.assign(slices + 1, 0);
  for (int s = 0; s < slices; ++s) {
    size_t w = 0;
    for (size_t r = a + 32 * s; r < b && r < a + 32 * (s + 1); ++r) w = std::max(w, rows[r].size());
    p.tile_start[s + 1] = p.tile_start[s] + (int)(32 * w);
  }
  p.dim.assign(p.tile_start[slices], -1);
  p.feat.assign(p.tile_start[slices], 0.f);
  p.pre_norm.assign(p.n, 0.f);
  for (int r = 0; r < p.n; ++r) {
    const int s = r / 32, base = p.tile_start[s] + r % 32;
    const auto& row = rows[a + r];
    for (size_t j = 0; j < row.size(); ++j) {
      p.dim[base + 32 * j] = row[j].first;
      p.feat[base + 32 * j] = row[j].second;
    }
  }
  return p;
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s matrix.txt\n", argv[0]);
    return 2;
  }
  std::ifstream in(argv[1]);
  std::vector<std::vector<std::pair<int, float>>> rows;
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream ls(line.substr(line.empty() ? 0 : 1));  // the leading separator
    std::vector<std::pair<int, float>> row;
    int d;
    float v;
    while (ls >> d >> v) row.push_back({d, v});
    rows.push_back(row);
  }
  Part P[8];
  for (int k = 0; k < 8; ++k) P[k] = make_part(rows, rows.size() * k / 8, rows.size() * (k + 1) / 8);
  similarity_join_device(
      P[0].dim.data(), P[0].feat.data(), P[0].tile_start.data(), P[1].dim.data(), P[1].feat.data(),
      P[1].tile_start.data(), P[2].dim.data(), P[2].feat.data(), P[2].tile_start.data(),
      P[3].dim.data(), P[3].feat.data(), P[3].tile_start.data(), P[4].dim.data(), P[4].feat.data(),
      P[4].tile_start.data(), P[5].dim.data(), P[5].feat.data(), P[5].tile_start.data(),
      P[6].dim.data(), P[6].feat.data(), P[6].tile_start.data(), P[7].dim.data(), P[7].feat.data(),
      P[7].tile_start.data(), P[0].pre_norm.data(), P[1].pre_norm.data(), P[2].pre_norm.data(),
      P[3].pre_norm.data(), P[4].pre_norm.data(), P[5].pre_norm.data(), P[6].pre_norm.data(),
      P[7].pre_norm.data(), P[0].n, (int)P[0].dim.size(), (int)P[0].tile_start.size() - 1, P[1].n,
      (int)P[1].dim.size(), (int)P[1].tile_start.size() - 1, P[2].n, (int)P[2].dim.size(),
      (int)P[2].tile_start.size() - 1, P[3].n, (int)P[3].dim.size(), (int)P[3].tile_start.size() - 1,
      P[4].n, (int)P[4].dim.size(), (int)P[4].tile_start.size() - 1, P[5].n, (int)P[5].dim.size(),
      (int)P[5].tile_start.size() - 1, P[6].n, (int)P[6].dim.size(), (int)P[6].tile_start.size() - 1,
      P[7].n, (int)P[7].dim.size(), (int)P[7].tile_start.size() - 1, nullptr, nullptr, nullptr);
  std::printf("count %lld\n", (long long)apss_compat_last_count());
  return apss_compat_last_count() < 0 ? 1 : 0;
}
