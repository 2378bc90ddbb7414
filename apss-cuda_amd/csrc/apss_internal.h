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
// Rows are cut into bands (I side, at most BAND_ROWS rows) and tiles (J side,
// at most TILE_ROWS rows) of consecutive rows.  A band x tile pair owns a
// BAND_ROWS x TILE_ROWS FP32 score tile in LDS (row stride TILE_ROWS+1 so the
// cells one wave updates for a fixed I row fall in different banks).  Tile and
// band heights are variable: a band/tile also ends before its bitmap-tier
// entries exceed CAP_I / CAP_J, so its whole block always fits in LDS; since a
// row holds at most KD_MAX <= CAP_J bitmap-tier entries, a single row fits.
constexpr int BAND_ROWS = 128;
constexpr int TILE_ROWS = 64;
// The score tile is stored transposed: cell (il, jl) at jl * TSTRIDE + il.
// A lane's four cells of an MFMA output block (four consecutive band rows of
// one tile row) are then one 16-byte word, and the 16 tile rows one wave
// touches (stride 132 words = 4 banks apart) cover all 64 banks once; the
// sparse products of one feature (distinct band rows, same tile row) fall in
// distinct banks.
constexpr int TSTRIDE = BAND_ROWS + 4;
constexpr int SCAN_THREADS = 1024;           // band kernel workgroup (16 waves)
constexpr int KD_MAX = 4032;                 // features in the bitmap tier (KD + 1 <= 4 x 1024)
constexpr int KW_MAX = KD_MAX / 64;          // bitmap words per tile
constexpr int CAP_I = 8192;                  // bitmap-tier entries of a band
constexpr int CAP_J = 4096;                  // bitmap-tier entries of a tile
constexpr int CAP_S = 8192;                  // cursor-tier records of a band (band cut)
constexpr int S_LIST_MAX = 8192;             // longest list sorted in LDS in one piece
constexpr int MERGE_TILE = 2048;             // outputs per workgroup of a merge pass
constexpr int CH_TILES = 1024;               // J tiles per cursor-bucket chunk
constexpr int WALK_PAD = 16;                 // padding entries after the cursor-tier lists

// Dense MFMA head: the most frequent tail features, stored per row as a
// 256-byte record and contracted on the matrix cores.
//   DENSE_I8   64 features, each value as a 28-bit fixed-point integer split in
//              four signed 8-bit limbs (record = limb-major, 4 x 64 bytes);
//              integer MFMA makes the head exact up to the fixed-point step.
//   DENSE_BF16 128 features as BF16 (round to nearest even), FP32 MFMA
//              accumulation; pairs within the BF16 error bound are rescored.
enum : int32_t { DENSE_OFF = 0, DENSE_I8 = 1, DENSE_BF16 = 2 };
constexpr int DENSE_ROW_BYTES = 256;
constexpr int DENSE_I8_FEATURES = 64;
constexpr int DENSE_BF16_FEATURES = 128;
// With a dense head the sparse blocks are smaller, and the LDS they free holds
// the head operands: caps of the bitmap tier per band / tile / row.
// The tile cap (and with it the bitmap tier, since a row must fit one tile
// block) takes the LDS the kernels have left (the BF16 kernel's per-row head
// norms take some).  For the int8 head a larger tier measured slower: the
// features beyond ~3800 cost more as lookups than as cursor products.
constexpr int CAP_I_DENSE = 6144;
constexpr int CAP_J_I8 = 3808, CAP_J_BF16 = 3680;
__host__ __device__ constexpr int cap_i_for(int dm) { return dm ? CAP_I_DENSE : CAP_I; }
__host__ __device__ constexpr int cap_j_for(int dm) {
  return dm == DENSE_I8 ? CAP_J_I8 : dm == DENSE_BF16 ? CAP_J_BF16 : CAP_J;
}
// a row holds at most KD bitmap-tier entries and must fit one tile block
__host__ __device__ constexpr int kd_max_for(int dm) { return dm ? cap_j_for(dm) : KD_MAX; }

// Feature classes in fmap[] (values >= 0 are bitmap-tier indices).
constexpr int32_t F_HEAD = -1;    // id < p: head (forward list / MFMA)
constexpr int32_t F_CURSOR = -2;  // short list: cursor tier
constexpr int32_t F_IGNORE = -3;  // cannot form a pair (list shorter than needed)
constexpr int32_t F_DENSE0 = -16; // F_DENSE0 - k: k-th feature of the dense MFMA head

// Error bits raised by the validation kernel.
enum : uint32_t {
  VERR_ROWPTR = 1u,   // row_ptr not monotone / wrong ends
  VERR_COL = 2u,      // column out of range
  VERR_ORDER = 4u,    // columns not strictly ascending in a row
  VERR_VALUE = 8u,    // non-finite value
  VERR_NORM = 16u,    // row norm too large for the FP32 path
  VFLAG_NEG = 1u << 31,  // not an error: the matrix holds a negative value
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

// Tiled structures of one matrix side (corpus or queries) for one tiling.
struct TileSide {
  int64_t n = 0;                  // rows
  int32_t T = 0;                  // tiles (or bands)
  const int64_t* tstart = nullptr;  // T+1 first row of each tile
  const int32_t* row2tile = nullptr;  // n, tile of each row
  const int64_t* rp = nullptr;    // CSR (device)
  const int32_t* col = nullptr;
  const float* val = nullptr;
  const float* norm = nullptr;    // ||x_i||_2 rounded up
  const float* hnorm = nullptr;   // ||x_i restricted to ids < p||_2 rounded up
  // bitmap tier, per tile: presence bitmap, per-word prefix of present
  // features, compact segment starts (present+1, relative to the tile block)
  const uint64_t* bmp = nullptr;  // T * KW
  const uint16_t* wpre = nullptr; // T * KW
  const int64_t* cbase = nullptr; // T+1
  const int32_t* cst = nullptr;
  const int64_t* dbase = nullptr; // T+1, start of each tile block
  const uint8_t* drow = nullptr;  // local row of each bitmap-tier entry
  const float* dval = nullptr;
  const uint16_t* dfeat = nullptr;  // bitmap index of each bitmap-tier entry
  // cursor-tier entries of this side (one record per entry, grouped by tile)
  const int64_t* sbase = nullptr; // T+1
  const int4* srec = nullptr;     // {local row, float bits of x, list begin, list end}
  // bitmap tier, per tile: start of every feature's segment in the tile
  // block, indexed by bitmap index (KD + 1 entries, row stride kds)
  const uint16_t* dseg = nullptr;
  int32_t kds = 0;
  // dense MFMA head: one 256-byte record per row (16 x 16-byte chunks)
  const uint4* dense = nullptr;
  const float* dnorm = nullptr;   // norm of the head part of each row, rounded up (BF16)
  const int2* rq = nullptr;       // settle bounds {E, Q} of each row for this launch
  const float2* l1m = nullptr;    // per tile: largest row L1 norm, largest |value| (rounded up)
};

// Arguments of k_row_bounds (settle bounds of the rows of one side).
struct RowBoundArgs {
  double other_norm;  // largest norm of the other side (rounded up)
  double qinv, qscale;
  double e_fixed;     // error terms independent of the rows (dense head)
  double rel_k;       // relative score tolerance as 1 / tol (<= 0: none)
  int32_t head;       // split p > 0: add the FP32 head-dot term
  int32_t pad_;
};

// Cursor-tier inverted lists of the corpus side, ascending row inside each
// list, as two arrays per entry (value; J tile and row).  val and pos hold
// WALK_PAD entries past the last list, so the band kernel's walks can read
// aligned 16-byte words.
struct CursorLists {
  const int64_t* ptr = nullptr;   // d+1
  const float* val = nullptr;     // value of each entry
  const uint32_t* pos = nullptr;  // J tile << 8 | row within the tile, per entry
};

struct ScanArgs {
  TileSide I, J;
  CursorLists S;
  int32_t KD, KW, p;
  int32_t self;              // 1: J >= I and only i < j
  int32_t head_mode;         // 0: no head (p == 0), 1: forward-list dot (selects the kernel)
  double t;
  // Score tile cells are fixed point: a product x*y adds trunc(x * 2^S * y)
  // (int32).  Partial sums stay below 2^31 because 2^S * |x||y| <= 2^30.
  float qscale;              // 2^S
  float qinv_f;              // 2^-S (exact in FP32)
  float cbf_q;               // BF16 head: c_bf 2^S rounded up (per-cell part of the bound)
  uint32_t cabs32;           // BF16 head: c_bf / (1 - c_bf) x 2^32 rounded up
  int32_t bf_nonneg;         // BF16 head: no negative value on either side (the head
                             // error is bounded from the cell's own score)
  int32_t q_lo;              // cells below this cannot reach t (fast reject)
  double qinv;               // 2^-S
  double score_rel;          // emitted scores must be within this relative error
  int64_t i_off, j_off;      // id offsets added to emitted i / j
  // outputs
  apss_pair* out;
  unsigned long long* out_cnt;
  int64_t out_cap;
  unsigned long long* band;  // packed (i << 32 | j) pairs to rescore (local ids)
  unsigned long long* band_cnt;
  int64_t band_cap;
  int32_t rescore_all;
  int32_t n_items;           // work items of this launch
  const int4* items;         // {band I, first J tile, one past the last J tile, unused}
  unsigned long long* stats; // [0] products [1] head pairs [2] band bitmap-tier entries
                             // (per item) [3] tile-side list visits [4] cursor products
                             // [5] band cursor records (per item) [6] count-walk reads
                             // [8] scatter-walk reads ([7] is the work-item queue); cells,
                             // tile pairs and dense records follow from the work items and
                             // are counted on the host
  const double* band_t;      // per band: lowest threshold of its rows (NULL: t)
  const double* row_t;       // per I row: threshold (NULL: the band's)
  double eps_max;            // largest cell error bound (for per-band fast rejects)
  int4* scs;                 // per cursor record of the current band: {position, end,
                             // J tile at position (INT32_MAX: exhausted), unused};
                             // one region of scs_cap records per workgroup
  int64_t scs_cap;
  // cursor-tier products, bucketed by J tile for one chunk of CH_TILES tiles;
  // one region of cprod_cap records per workgroup of the persistent grid
  int2* cprod;               // {cell = jl * TSTRIDE + il, fixed-point product}
  int64_t cprod_cap;
  unsigned long long* item_next;  // work-item queue of the persistent grid
  // dense head: int8 limb groups combine as hi << dsh_hi + (lo >> dsh_lo);
  // eps_dense bounds the int8 head's error, c_bf * |x_D||y_D| the BF16 head's
  int32_t dsh_hi, dsh_lo;
  double eps_dense;
  double c_bf;
};
// stats[] slots beside the ScanArgs comment above: tile pairs of sparse bands
// skipped for want of a common feature, and skipped by a block bound
constexpr int STAT_SKIP_NO_COMMON = 9, STAT_SKIP_BOUND = 10;

// A piece of a cursor list sorted in LDS, and one merge-pass work unit: the
// sorted runs [a, a+la) and [a+la, a+la+lb) merge into the same positions of
// the other buffer; this unit writes outputs [k*MERGE_TILE, (k+1)*MERGE_TILE).
struct SortChunk {
  int64_t start;
  int32_t len;
  int32_t pad_;
};
struct MergeTask {
  int64_t a;
  int32_t la, lb;
  int32_t k;
  int32_t pad_;
};

// Arguments of the per-tile bitmap-tier build kernels.
struct BuildSide {
  int64_t n;
  int32_t T;
  int32_t KD;
  int32_t KW;
  int32_t p;
  int32_t self;
  int32_t pad_;
  const int64_t* tstart;      // T+1
  const int64_t* rp;
  const int32_t* col;
  const float* val;
  const int32_t* fmap;
  int32_t* full;              // T*(KD+1) scratch: counts, then pointers
  uint64_t* bmp;              // T*KW
  uint16_t* wpre;             // T*KW
  int64_t* ctot;              // T+1 present features, then exclusive scan
  int32_t* cst;
  int64_t* dtot;              // T+1 (counts, then exclusive scan)
  int64_t* stot;              // T+1
  uint8_t* drow;
  float* dval;
  uint16_t* dfeat;
  int4* srec;
  uint16_t* dseg;             // T*kds segment starts by bitmap index
  int32_t kds;
  const int64_t* sptr;        // corpus cursor lists
  const int2* sent;
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
  // leader is wave-uniform: read the two halves straight from its lane
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)base, leader);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(base >> 32), leader);
  base = ((unsigned long long)hi << 32) | lo;
  return (long long)base + lane_rank(m);
}

// Rigorous bound of |s_tile - s| for one cell: FP32 product rounding (u per
// product, summed <= u*|x||y|), fixed-point truncation (under one quantum
// 2^-S per product and one for the head), and the FP32 head dot when there is
// one.
__device__ __host__ __forceinline__ double cell_error_bound(int k, double norm_i, double norm_j,
                                                            double qinv, double head_bound) {
  const double u = 5.9604644775390625e-08;
  return u * norm_i * norm_j * (1.0 + 1e-6) + (double)(k + 2) * qinv + head_bound;
}

// Unit roundoff bound gamma_k = k*u/(1-k*u) for FP32 (u = 2^-24).
__device__ __host__ __forceinline__ double gamma32(int64_t k) {
  const double u = 5.9604644775390625e-08;
  double ku = (double)k * u;
  return ku / (1.0 - ku);
}

}  // namespace apss
