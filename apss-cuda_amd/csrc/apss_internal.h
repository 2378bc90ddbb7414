// Internal definitions shared by the libapss.so sources (HIP, gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <string>
#include <vector>

#include "../../include/apss.h"

namespace apss {

// ---------------------------------------------------------------------------
// Layout constants
//
// Rows are cut into tiles of TILE consecutive rows.  A tile pair (I, J) owns a
// TILE x TILE FP32 score tile in LDS (row stride TILE+1 so that the cells one
// wave updates for a fixed column fall in different banks).
constexpr int TILE = 128;
constexpr int TSTRIDE = TILE + 1;
constexpr int SCAN_THREADS = 1024;           // tile kernel workgroup (16 waves)
constexpr int KD_MAX = 32768;                // features in the bitmap tier
constexpr int KW_MAX = KD_MAX / 64;          // bitmap words per tile
constexpr int AND_CHUNK_WORDS = 64;          // bitmap words intersected per pass
constexpr int COMMON_CAP = AND_CHUNK_WORDS * 64;
constexpr int S_LDS_CAP = 6144;              // cursor heads kept in LDS per band
constexpr int S_LIST_MAX = 4096;             // longest list of the cursor tier
constexpr int BIG_PRODUCTS = 32;             // common features with >= this many
                                             // products get a whole wave

// Feature classes in fmap[] (values >= 0 are bitmap-tier indices).
constexpr int32_t F_HEAD = -1;    // id < p: head (forward list / MFMA)
constexpr int32_t F_CURSOR = -2;  // short list: cursor tier
constexpr int32_t F_IGNORE = -3;  // cannot form a pair (list shorter than needed)

// Error bits raised by the validation kernel.
enum : uint32_t {
  VERR_ROWPTR = 1u,   // row_ptr not monotone / wrong ends
  VERR_COL = 2u,      // column out of range
  VERR_ORDER = 4u,    // columns not strictly ascending in a row
  VERR_VALUE = 8u,    // non-finite value
  VERR_NORM = 16u,    // row norm too large for the FP32 path
};

// ---------------------------------------------------------------------------
// Error handling

struct Error {
  int code;
  std::string what;
};

#define APSS_HIP(expr)                                                              \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess)                                                           \
      throw ::apss::Error{APSS_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)}; \
  } while (0)

// ---------------------------------------------------------------------------
// Device views

// Row-tile structures of one matrix side (corpus or queries).
struct TileSide {
  int64_t n = 0;               // rows
  int32_t T = 0;               // tiles
  const int64_t* rp = nullptr; // CSR (device)
  const int32_t* col = nullptr;
  const float* val = nullptr;
  const float* norm = nullptr;   // ||x_i||_2 rounded up
  const float* hnorm = nullptr;  // ||x_i restricted to ids < p||_2 rounded up
  // bitmap tier
  const uint64_t* bmp = nullptr; // T * KW
  const int32_t* ptr = nullptr;  // T * (KD+1), offsets inside the tile block
  const int64_t* base = nullptr; // T+1, start of each tile block
  const uint8_t* drow = nullptr; // local row of each bitmap-tier entry
  const float* dval = nullptr;
  // cursor-tier entries of this side (one record per entry, grouped by tile)
  const int64_t* sbase = nullptr;  // T+1
  const int4* srec = nullptr;      // {local row, float bits of x, cursor, end}
};

// Cursor-tier inverted lists of the corpus side (ascending row per list).
struct CursorLists {
  const int64_t* ptr = nullptr;  // d+1
  const int32_t* row = nullptr;
  const float* val = nullptr;
};

struct ScanArgs {
  TileSide I, J;
  CursorLists S;
  int32_t KD, KW, p;
  int32_t self;              // 1: J >= I and only i < j
  int32_t head_mode;         // 0: no head (p == 0), 1: forward-list dot
  double t;
  float t_lo;                // float lower bound of t - max error (fast reject)
  float pad_;
  int64_t i_off, j_off;      // id offsets added to emitted i / j
  // outputs
  apss_pair* out;
  unsigned long long* out_cnt;
  int64_t out_cap;
  unsigned long long* band;  // packed (i << 32 | j) pairs to rescore (local ids)
  unsigned long long* band_cnt;
  int64_t band_cap;
  int32_t rescore_all;
  int32_t band_first;        // first band (tile row of I) of this launch
  int32_t band_count;
  int32_t pad2_;
  unsigned long long* stats; // [0] products, [1] head pairs, [2] cells
  int32_t* snext_spill;      // per band: S_next beyond the LDS capacity
  int64_t spill_stride;
};

// Arguments of the per-tile bitmap-tier build kernels.
struct BuildSide {
  int64_t n;
  int32_t T;
  int32_t KD;
  int32_t p;
  int32_t self;
  int64_t row_off;            // global id of local row 0 (for cursor starts)
  const int64_t* rp;
  const int32_t* col;
  const float* val;
  const int32_t* fmap;
  int32_t* ptr;               // T*(KD+1)
  uint64_t* bmp;              // T*KW
  int64_t* dtot;              // T+1 (counts, then exclusive scan)
  int64_t* stot;              // T+1
  uint8_t* drow;
  float* dval;
  int4* srec;
  const int64_t* sptr;        // corpus cursor lists
  const int32_t* srow;
};

// ---------------------------------------------------------------------------
// Wave64 helpers

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Number of set bits of `mask` below this lane (the lane's compaction rank).
__device__ __forceinline__ int lane_rank(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0));
}

// Wave-aggregated reservation of `want` (0/1) slots in a global counter.
// Returns the slot of this lane (valid only if want) and the wave total.
__device__ __forceinline__ long long wave_reserve(unsigned long long* counter, bool want,
                                                  int* total_out) {
  unsigned long long m = __ballot(want);
  int total = __popcll(m);
  *total_out = total;
  if (total == 0) return 0;
  int leader = __ffsll((long long)m) - 1;
  unsigned long long base = 0;
  if (lane_id() == leader) base = atomicAdd(counter, (unsigned long long)total);
  base = __shfl(base, leader);
  return (long long)base + lane_rank(m);
}

// Unit roundoff bound gamma_k = k*u/(1-k*u) for FP32 (u = 2^-24).
__device__ __host__ __forceinline__ double gamma32(int64_t k) {
  const double u = 5.9604644775390625e-08;
  double ku = (double)k * u;
  return ku / (1.0 - ku);
}

}  // namespace apss
