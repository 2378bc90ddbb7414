// Kernel (a): inverted-index construction (CSR -> tiled CSC) for gfx950.
//
// The reference builds its tail index with a global-atomic histogram per
// partition (generate_index_ptr4, similarity_search.cu:105-191), a host scan
// (:1168-1186) and an atomic-cursor scatter into one list per feature
// (construct_Index2, :221-274), then re-uploads the pointers to undo the
// cursor (:1256-1263).  Here the tail (ids >= p) is split by list length:
//
//   bitmap tier  features whose lists are long (present in a sizeable share of
//                the row tiles): stored per row tile as a CSC block with a
//                presence bitmap and segment pointers, so a tile pair finds
//                its common features by AND-ing two bitmaps;
//   cursor tier  short lists: one global CSC, ascending row per list, walked
//                by per-entry cursors while a band of tiles sweeps its
//                partner tiles.
//
// Histogram: LDS-privatised per workgroup for the most frequent ids, global
// atomics for the rest.  Scan: per tile in LDS.  Scatter: atomic cursors whose
// final values are the segment ends, so the pointer array is written shifted by
// one slot and comes out right without a restore pass.
#include "apss_internal.h"

namespace apss {

// ---------------------------------------------------------------------------
// fills (used instead of hipMemsetAsync so that the sequence can be captured)

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_u64(unsigned long long* p, int64_t n, unsigned long long v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_copy_i64(int64_t* dst, const int64_t* src, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// validation + norms: one wave per row.  err[0] collects the error bits (and
// VFLAG_NEG, which is no error), err[1] and err[2] (zeroed by the caller) the
// longest row and the largest norm.

__device__ __forceinline__ float round_up_f(double s) {
  float f = (float)s;
  if ((double)f < s) f = nextafterf(f, INFINITY);
  return f;
}

__global__ __launch_bounds__(256) void k_validate_norms(int64_t n, int64_t d, int64_t nnz,
                                                        const int64_t* __restrict__ rp,
                                                        const int32_t* __restrict__ col,
                                                        const float* __restrict__ val, int32_t p,
                                                        float* __restrict__ norm,
                                                        float* __restrict__ hnorm,
                                                        uint32_t* __restrict__ err) {
  const int lane = lane_id();
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
  uint32_t max_len = 0, max_norm = 0;  // of this wave's rows (norm as float bits, >= 0)
  for (int64_t r = blockIdx.x * (int64_t)(blockDim.x / 64) + threadIdx.x / 64; r < n; r += waves) {
    int64_t a = rp[r], b = rp[r + 1];
    uint32_t e_bits = 0;
    if (r == 0 && a != 0) e_bits |= VERR_ROWPTR;
    if (r == n - 1 && b != nnz) e_bits |= VERR_ROWPTR;
    if (a < 0 || b > nnz || a > b) {
      if (lane == 0) atomicOr(err, e_bits | VERR_ROWPTR);
      if (lane == 0) { norm[r] = 0.f; hnorm[r] = 0.f; }
      continue;
    }
    double ss = 0.0, hs = 0.0;
    for (int64_t e = a + lane; e < b; e += 64) {
      int32_t c = col[e];
      float v = val[e];
      if (c < 0 || c >= d) e_bits |= VERR_COL;
      if (e + 1 < b && col[e + 1] <= c) e_bits |= VERR_ORDER;
      if (!isfinite(v)) e_bits |= VERR_VALUE;
      if (v < 0.f) e_bits |= VFLAG_NEG;
      double v2 = (double)v * (double)v;
      ss += v2;
      if (c < p) hs += v2;
    }
    for (int o = 32; o > 0; o >>= 1) {
      ss += __shfl_xor(ss, o);
      hs += __shfl_xor(hs, o);
      e_bits |= __shfl_xor(e_bits, o);
    }
    if (lane == 0) {
      // rows must keep |score| <= ||x|| ||y|| far from FP32 overflow
      if (!(ss < 1.3e36)) e_bits |= VERR_NORM;
      const float nr = round_up_f(sqrt(ss));
      norm[r] = nr;
      hnorm[r] = round_up_f(sqrt(hs));
      if (e_bits) atomicOr(err, e_bits);
      const int64_t len = b - a;
      max_len = max(max_len, (uint32_t)(len < 0x7fffffff ? len : 0x7fffffff));
      max_norm = max(max_norm, __float_as_uint(nr));
    }
  }
  // the maxima over all rows: err[1] the longest row, err[2] the largest norm
  if (lane == 0 && (max_len || max_norm)) {
    atomicMax(err + 1, max_len);
    atomicMax(err + 2, max_norm);
  }
}

// ---------------------------------------------------------------------------
// feature histogram of ids >= p (LDS-privatised for ids < HIST_LDS)

constexpr int HIST_LDS = 16384;

__global__ __launch_bounds__(1024) void k_col_hist(int64_t nnz, const int32_t* __restrict__ col,
                                                   int32_t p, int64_t chunk,
                                                   int32_t* __restrict__ cnt) {
  __shared__ int32_t h[HIST_LDS];
  for (int i = threadIdx.x; i < HIST_LDS; i += blockDim.x) h[i] = 0;
  __syncthreads();
  int64_t lo = blockIdx.x * chunk, hi = lo + chunk < nnz ? lo + chunk : nnz;
  for (int64_t e = lo + threadIdx.x; e < hi; e += blockDim.x) {
    int32_t c = col[e];
    if (c < p) continue;
    if (c < HIST_LDS) atomicAdd(&h[c], 1);
    else atomicAdd(&cnt[c], 1);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < HIST_LDS; i += blockDim.x)
    if (h[i]) atomicAdd(&cnt[i], h[i]);
}

// ---------------------------------------------------------------------------
// cursor tier: scatter into one list per feature, then sort each list by row

__global__ __launch_bounds__(256) void k_cursor_scatter(int64_t n, const int64_t* __restrict__ rp,
                                                        const int32_t* __restrict__ col,
                                                        const float* __restrict__ val,
                                                        const int32_t* __restrict__ fmap,
                                                        unsigned long long* __restrict__ cur,
                                                        int2* __restrict__ sent) {
  const int lane = lane_id();
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
  for (int64_t r = blockIdx.x * (int64_t)(blockDim.x / 64) + threadIdx.x / 64; r < n; r += waves) {
    for (int64_t e = rp[r] + lane; e < rp[r + 1]; e += 64) {
      int32_t c = col[e];
      if (fmap[c] != F_CURSOR) continue;
      unsigned long long q = atomicAdd(&cur[c], 1ull);
      sent[q] = make_int2((int32_t)r, __float_as_int(val[e]));
    }
  }
}

// One wave sorts one list of length <= 64 (bitonic network over lanes).
__global__ __launch_bounds__(256) void k_sort_small(const int32_t* __restrict__ lists, int32_t nl,
                                                    const int64_t* __restrict__ sptr,
                                                    int2* __restrict__ sent) {
  int w = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
  if (w >= nl) return;
  const int lane = lane_id();
  int32_t f = lists[w];
  int64_t a = sptr[f];
  int len = (int)(sptr[f + 1] - a);
  int2 e = lane < len ? sent[a + lane] : make_int2(INT32_MAX, 0);
  for (int k = 2; k <= 64; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      int ox = __shfl_xor(e.x, j), oy = __shfl_xor(e.y, j);
      bool up = ((lane & k) == 0);
      bool lower = ((lane & j) == 0);
      bool swap = (lower == up) ? (ox < e.x) : (ox > e.x);
      if (swap) { e.x = ox; e.y = oy; }
    }
  }
  if (lane < len) sent[a + lane] = e;
}

// One workgroup sorts one piece (<= S_LIST_MAX entries) of a cursor list by
// row in LDS (bitonic network over the piece padded to a power of two).
__global__ __launch_bounds__(256) void k_sort_chunks(const SortChunk* __restrict__ chunks,
                                                     int2* __restrict__ sent) {
  __shared__ int2 buf[S_LIST_MAX];
  const SortChunk c = chunks[blockIdx.x];
  const int64_t a = c.start;
  const int len = c.len;
  int np = 64;
  while (np < len) np <<= 1;
  for (int i = threadIdx.x; i < np; i += blockDim.x)
    buf[i] = i < len ? sent[a + i] : make_int2(INT32_MAX, 0);
  __syncthreads();
  for (int k = 2; k <= np; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < np; i += blockDim.x) {
        int ixj = i ^ j;
        if (ixj > i) {
          bool up = ((i & k) == 0);
          int2 x = buf[i], y = buf[ixj];
          if ((x.x > y.x) == up) { buf[i] = y; buf[ixj] = x; }
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < len; i += blockDim.x) sent[a + i] = buf[i];
}

// One merge-pass unit: 256 threads x 8 consecutive outputs of the merge of
// two row-sorted runs.  Each thread finds where its first output splits the
// runs (merge path: the number i of A entries among the first d outputs is
// the first i with A[i] > B[d-i-1]), then merges sequentially.  A run pair
// with lb = 0 is a plain copy.
__global__ __launch_bounds__(256) void k_merge_runs(const MergeTask* __restrict__ tasks,
                                                    const int2* __restrict__ src,
                                                    int2* __restrict__ dst) {
  const MergeTask t = tasks[blockIdx.x];
  const int2* A = src + t.a;
  const int2* B = A + t.la;
  int2* O = dst + t.a;
  const int total = t.la + t.lb;
  const int d = t.k * MERGE_TILE + (int)threadIdx.x * 8;
  if (d >= total) return;
  int lo = d - t.lb > 0 ? d - t.lb : 0, hi = d < t.la ? d : t.la;
  while (lo < hi) {
    const int i = (lo + hi) >> 1;
    if (A[i].x <= B[d - i - 1].x) lo = i + 1; else hi = i;
  }
  int i = lo, j = d - lo;
  for (int q = 0; q < 8 && d + q < total; ++q) {
    const bool take_a = j >= t.lb || (i < t.la && A[i].x <= B[j].x);
    O[d + q] = take_a ? A[i++] : B[j++];
  }
}

// ---------------------------------------------------------------------------
// per-row tier counts (inputs of the host tiling)

__global__ __launch_bounds__(256) void k_row_counts(int64_t n, const int64_t* __restrict__ rp,
                                                    const int32_t* __restrict__ col, int32_t p,
                                                    const int32_t* __restrict__ fmap,
                                                    int32_t* __restrict__ dcnt,
                                                    int32_t* __restrict__ scnt,
                                                    int32_t* __restrict__ hcnt) {
  const int lane = lane_id();
  const int64_t waves = (int64_t)gridDim.x * (blockDim.x / 64);
  for (int64_t r = blockIdx.x * (int64_t)(blockDim.x / 64) + threadIdx.x / 64; r < n; r += waves) {
    int dc = 0, sc = 0, hc = 0;  // bitmap tier, cursor tier, head (ids < p) + dense head
    for (int64_t e = rp[r] + lane; e < rp[r + 1]; e += 64) {
      int32_t c = col[e];
      if (c < p) {
        ++hc;
        continue;
      }
      int32_t f = fmap[c];
      dc += f >= 0;
      sc += f == F_CURSOR;
      hc += f <= F_DENSE0;
    }
    for (int o = 32; o > 0; o >>= 1) {
      dc += __shfl_xor(dc, o);
      sc += __shfl_xor(sc, o);
      hc += __shfl_xor(hc, o);
    }
    if (lane == 0) { dcnt[r] = dc; scnt[r] = sc; hcnt[r] = hc; }
  }
}

// ---------------------------------------------------------------------------
// per-tile bitmap tier

// Count bitmap-tier entries per (tile, feature) into slot f+1 of the tile's
// scratch pointer row, and the tile totals of both tiers.
__global__ __launch_bounds__(256) void k_tile_count(BuildSide s) {
  __shared__ unsigned long long dcount, scount;
  const int tile = blockIdx.x;
  if (threadIdx.x == 0) { dcount = 0; scount = 0; }
  __syncthreads();
  int64_t r0 = s.tstart[tile], r1 = s.tstart[tile + 1];
  int32_t* prow = s.full + (int64_t)tile * (s.KD + 1);
  unsigned long long dc = 0, sc = 0;
  for (int64_t e = s.rp[r0] + threadIdx.x; e < s.rp[r1]; e += blockDim.x) {
    int32_t c = s.col[e];
    if (c < s.p) continue;
    int32_t f = s.fmap[c];
    if (f >= 0) { atomicAdd(&prow[f + 1], 1); ++dc; }
    else if (f == F_CURSOR) ++sc;
  }
  atomicAdd(&dcount, dc);
  atomicAdd(&scount, sc);
  __syncthreads();
  // blocks are padded to multiples of 4 entries (32/128-bit staging copies)
  if (threadIdx.x == 0) { s.dtot[tile] = (int64_t)((dcount + 3) & ~3ull); s.stot[tile] = (int64_t)scount; }
}

// Exclusive scan of a[0..n) in place, a[n] = total (single workgroup).
__global__ __launch_bounds__(1024) void k_scan_i64(int64_t* a, int64_t n) {
  __shared__ int64_t part[1024];
  __shared__ int64_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < n; base += 1024) {
    int64_t i = base + threadIdx.x;
    int64_t x = i < n ? a[i] : 0;
    part[threadIdx.x] = x;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      int64_t y = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0;
      __syncthreads();
      part[threadIdx.x] += y;
      __syncthreads();
    }
    if (i < n) a[i] = carry + part[threadIdx.x] - x;
    __syncthreads();
    if (threadIdx.x == 1023) carry += part[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) a[n] = carry;
}

// Per tile: counts c[f] (slot f+1) -> exclusive starts E[f] (slot f+1), slot 0
// = 0, plus the presence bitmap and the number of present features.  After
// the scatter, which advances slot f+1 from E[f] to E[f]+c[f] = E[f+1], the
// row is the standard pointer array.
__global__ __launch_bounds__(1024) void k_tile_ptr_scan(BuildSide s) {
  __shared__ int32_t part[1024];
  __shared__ int32_t carry, present;
  const int tile = blockIdx.x;
  const int32_t KD = s.KD;
  int32_t* prow = s.full + (int64_t)tile * (KD + 1);
  uint64_t* brow = s.bmp + (int64_t)tile * s.KW;
  if (threadIdx.x == 0) { carry = 0; present = 0; prow[0] = 0; }
  __syncthreads();
  for (int32_t base = 0; base < KD; base += 1024) {
    int32_t f = base + threadIdx.x;
    int32_t x = f < KD ? prow[f + 1] : 0;
    unsigned long long m = __ballot(x > 0);
    if (lane_id() == 0 && f < KD) {
      brow[f / 64] = m;
      atomicAdd(&present, __popcll(m));
    }
    part[threadIdx.x] = x;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      int32_t y = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0;
      __syncthreads();
      part[threadIdx.x] += y;
      __syncthreads();
    }
    if (f < KD) prow[f + 1] = carry + part[threadIdx.x] - x;
    __syncthreads();
    if (threadIdx.x == 1023) carry += part[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) s.ctot[tile] = present + 1;  // starts + end sentinel
}

// Scatter bitmap-tier entries into their tile blocks and write the cursor
// records of this side's cursor-tier entries.
__global__ __launch_bounds__(256) void k_tile_scatter(BuildSide s) {
  __shared__ int32_t rps[BAND_ROWS + 1];
  __shared__ unsigned int sfill;
  const int tile = blockIdx.x;
  int64_t r0 = s.tstart[tile], r1 = s.tstart[tile + 1];
  int nr = (int)(r1 - r0);
  int64_t e0 = s.rp[r0];
  for (int i = threadIdx.x; i <= nr; i += blockDim.x) rps[i] = (int32_t)(s.rp[r0 + i] - e0);
  if (threadIdx.x == 0) sfill = 0;
  __syncthreads();
  int32_t* prow = s.full + (int64_t)tile * (s.KD + 1);
  int64_t dbase = s.dtot[tile], sbase = s.stot[tile];
  int32_t ne = rps[nr];
  for (int32_t k = threadIdx.x; k < ne; k += blockDim.x) {
    int64_t e = e0 + k;
    int32_t c = s.col[e];
    if (c < s.p) continue;
    int32_t f = s.fmap[c];
    if (f < 0 && f != F_CURSOR) continue;
    int lo = 0, hi = nr;  // local row of entry k: last r with rps[r] <= k
    while (hi - lo > 1) {
      int m = (lo + hi) >> 1;
      if (rps[m] <= k) lo = m; else hi = m;
    }
    if (f >= 0) {
      int32_t q = atomicAdd(&prow[f + 1], 1);
      s.drow[dbase + q] = (uint8_t)lo;
      s.dval[dbase + q] = s.val[e];
      s.dfeat[dbase + q] = (uint16_t)f;
    } else if (s.srec) {  // cursor records only for the band (I) side
      // the record spans its whole list: the band kernel positions it at each
      // work item's first tile, and in a self-join the cells j <= i that the
      // diagonal tiles receive are never emitted
      unsigned int q = atomicAdd(&sfill, 1u);
      s.srec[sbase + q] = make_int4(lo, __float_as_int(s.val[e]), (int)s.sptr[c], (int)s.sptr[c + 1]);
    }
  }
}

// Order the cursor records of each band by the length class of their list
// (floor(log2(end - first)), longest first).  The band kernel hands records to
// lanes in this order, so the 64 walks of a wave are of similar length instead
// of every wave waiting for its longest walk.  Order inside a class is
// arbitrary (the scatter above is too).  Bands with more records than the LDS
// stage holds keep their order.
__global__ __launch_bounds__(256) void k_sort_records(const int64_t* __restrict__ sbase,
                                                      int4* __restrict__ srec) {
  __shared__ int4 buf[CAP_S];
  __shared__ unsigned int cnt[32], fill[32];
  const int band = blockIdx.x;
  const int64_t b0 = sbase[band];
  const int nS = (int)(sbase[band + 1] - b0);
  if (nS <= 1 || nS > CAP_S) return;
  if (threadIdx.x < 32) cnt[threadIdx.x] = 0;
  __syncthreads();
  for (int k = threadIdx.x; k < nS; k += blockDim.x) {
    const int4 r = srec[b0 + k];
    buf[k] = r;
    const int len = r.w - r.z;
    atomicAdd(&cnt[len > 0 ? 31 - __clz(len) : 0], 1u);
  }
  __syncthreads();
  if (threadIdx.x == 0) {  // classes in descending order
    unsigned int acc = 0;
    for (int c = 31; c >= 0; --c) { fill[c] = acc; acc += cnt[c]; }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nS; k += blockDim.x) {
    const int4 r = buf[k];
    const int len = r.w - r.z;
    const unsigned int slot = atomicAdd(&fill[len > 0 ? 31 - __clz(len) : 0], 1u);
    srec[b0 + slot] = r;
  }
}

// Compact the pointer row of each tile: segment starts of the present
// features (in id order) and the per-word prefix counts of the bitmap.
__global__ __launch_bounds__(256) void k_tile_compact(BuildSide s) {
  __shared__ int32_t wp[KW_MAX + 1];
  const int tile = blockIdx.x;
  const uint64_t* brow = s.bmp + (int64_t)tile * s.KW;
  const int32_t* prow = s.full + (int64_t)tile * (s.KD + 1);
  if (threadIdx.x == 0) {
    int32_t acc = 0;
    for (int w = 0; w < s.KW; ++w) { wp[w] = acc; acc += __popcll(brow[w]); }
    wp[s.KW] = acc;
  }
  __syncthreads();
  uint16_t* wrow = s.wpre + (int64_t)tile * s.KW;
  int32_t* crow = s.cst + s.ctot[tile];
  for (int w = threadIdx.x; w < s.KW; w += blockDim.x) wrow[w] = (uint16_t)wp[w];
  for (int32_t f = threadIdx.x; f < s.KD; f += blockDim.x) {
    uint64_t m = brow[f / 64];
    int b = f % 64;
    if ((m >> b) & 1ull) {
      int rank = wp[f / 64] + __popcll(m & ((1ull << b) - 1ull));
      crow[rank] = prow[f];
    }
  }
  if (threadIdx.x == 0) crow[wp[s.KW]] = prow[s.KD];
  // segment start of every feature (absent ones: empty segments), then the
  // total in the remaining slots of the row
  uint16_t* drow16 = s.dseg + (int64_t)tile * s.kds;
  for (int32_t f = threadIdx.x; f < s.kds; f += blockDim.x)
    drow16[f] = (uint16_t)prow[f < s.KD ? f : s.KD];
}

// ---------------------------------------------------------------------------
// dense MFMA head records: one wave per row assembles the row's 256-byte
// record in LDS (zero where the row lacks a head feature), then stores it.
//   DENSE_I8   X = rint(x * 2^B), |X| <= 2^28, written as balanced base-256
//              digits X = d0 2^24 + d1 2^16 + d2 2^8 + d3 (d1..d3 in
//              [-128, 127]); limb a of head feature k at byte 64 a + k.
//   DENSE_BF16 x rounded to BF16 (nearest even) at bytes 2k, 2k+1; dnorm[r]
//              = norm of the row's head values, rounded up.

__global__ __launch_bounds__(256) void k_dense_rows(int64_t n, const int64_t* __restrict__ rp,
                                                    const int32_t* __restrict__ col,
                                                    const float* __restrict__ val, int32_t p,
                                                    const int32_t* __restrict__ fmap, int32_t dm,
                                                    int32_t B, uint32_t* __restrict__ dense,
                                                    float* __restrict__ dnorm) {
  __shared__ uint32_t rec[4][DENSE_ROW_BYTES / 4];
  const int w = threadIdx.x >> 6, lane = lane_id();
  uint8_t* rb = reinterpret_cast<uint8_t*>(rec[w]);
  const int64_t waves = (int64_t)gridDim.x * 4;
  for (int64_t r = blockIdx.x * 4 + w; r < n; r += waves) {
    rec[w][lane] = 0u;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    double ss = 0.0;
    for (int64_t e = rp[r] + lane; e < rp[r + 1]; e += 64) {
      const int32_t c = col[e];
      if (c < p) continue;
      const int32_t f = fmap[c];
      if (f > F_DENSE0) continue;
      const int k = F_DENSE0 - f;
      const float x = val[e];
      if (dm == DENSE_I8) {
        long long X = llrint(ldexp((double)x, B));
#pragma unroll
        for (int a = 3; a >= 1; --a) {
          const int dgt = (int)(((X + 128) & 255) - 128);
          rb[64 * a + k] = (uint8_t)(int8_t)dgt;
          X = (X - dgt) >> 8;
        }
        rb[k] = (uint8_t)(int8_t)X;
      } else {
        uint32_t u = __float_as_uint(x);
        u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite values)
        rb[2 * k] = (uint8_t)((u >> 16) & 0xffu);
        rb[2 * k + 1] = (uint8_t)(u >> 24);
        ss += (double)x * (double)x;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    dense[r * (DENSE_ROW_BYTES / 4) + lane] = rec[w][lane];
    if (dm == DENSE_BF16) {
      for (int o = 32; o > 0; o >>= 1) ss += __shfl_xor(ss, o);
      if (lane == 0) dnorm[r] = round_up_f(sqrt(ss) * (1.0 + 1e-15));
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
}

// J tile and row within that tile of every cursor-list entry, packed as
// tile << 8 | local row (tiles have at most 64 rows, tile ids fit 24 bits),
// and the entry's value.
__global__ __launch_bounds__(256) void k_cursor_pos(int64_t n, const int2* __restrict__ sent,
                                                    const int32_t* __restrict__ row2tile,
                                                    const int64_t* __restrict__ tstart,
                                                    uint32_t* __restrict__ spos,
                                                    float* __restrict__ sval) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int2 x = sent[e];
    const int32_t t = row2tile[x.x];
    spos[e] = ((uint32_t)t << 8) | (uint32_t)(x.x - tstart[t]);
    sval[e] = __int_as_float(x.y);  // the values alone, for the band kernel's walks
  }
}

// Tile index of every row.
__global__ __launch_bounds__(256) void k_row2tile(const int64_t* __restrict__ tstart,
                                                  int32_t* __restrict__ row2tile) {
  const int t = blockIdx.x;
  for (int64_t r = tstart[t] + threadIdx.x; r < tstart[t + 1]; r += blockDim.x) row2tile[r] = t;
}

}  // namespace apss

namespace apss {

// Per tile (or band) of a tiling: the largest L1 norm and the largest |value|
// of its rows, rounded up.  For two blocks I, J every pair (i, j) has
// |s_ij| <= min(max_I |x|_1 * max_J max|y|, max_I max|x| * max_J |y|_1), the
// block form of the reference's L1 x max bound (filter_pair2,
// similarity_search.cu:533-558).
__global__ __launch_bounds__(256) void k_tile_l1max(const int64_t* __restrict__ tstart,
                                                    const int64_t* __restrict__ rp,
                                                    const float* __restrict__ val, float2* out) {
  // one wave per row (rows r0 + w, r0 + w + 4, ...), lanes over its entries:
  // the value loads of a wave are contiguous
  __shared__ float sl[4], sm[4];
  const int t = blockIdx.x;
  const int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
  const int64_t r0 = tstart[t], r1 = tstart[t + 1];
  float l1 = 0.f, mx = 0.f;
  for (int64_t r = r0 + w; r < r1; r += 4) {
    double sum = 0.0;
    float m = 0.f;
    for (int64_t k = rp[r] + ln; k < rp[r + 1]; k += 64) {
      const float v = fabsf(val[k]);
      sum += (double)v;
      m = fmaxf(m, v);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      sum += __shfl_xor(sum, o);
      m = fmaxf(m, __shfl_xor(m, o));
    }
    l1 = fmaxf(l1, (float)sum * (1.0f + 1e-6f));  // covers the rounding to FP32
    mx = fmaxf(mx, m);
  }
  if (ln == 0) {
    sl[w] = l1;
    sm[w] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    out[t] = make_float2(fmaxf(fmaxf(sl[0], sl[1]), fmaxf(sl[2], sl[3])),
                         fmaxf(fmaxf(sm[0], sm[1]), fmaxf(sm[2], sm[3])));
}

}  // namespace apss
