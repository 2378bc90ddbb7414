// Kernels (b) tail scan, (c) head dot and (d) filter + compaction, fused into
// one band kernel, plus the FP64 rescore of the error band.
//
// Reference structure (similarity_search.cu): the tail scan accumulates every
// partial product with a global float atomicAdd into a dense nA x nB (or
// n(n-1)/2) matrix (similarity_join1 :370-437, similarity_apss1 :642-735);
// filter_pair / filter_pair_apss (:510-531, :833-852) then sweep that matrix
// with the prefix-norm bound, and similarity_join_rest / similarity_apss_rest
// (:441-506, :854-915) dot the head of every survivor against a dense LDS copy
// of one row and test >= t, with the emission left commented out.
//
// Here a persistent workgroup (one per CU, 1024 threads) takes work items from
// a queue; an item is a band I (up to 128 rows) against a range of its partner
// tiles J (up to 64 rows each; J >= I's own tiles for the self-join).  For each
// tile pair it keeps the 128 x 64 score tile in LDS as int32 fixed point, so
// the n^2 matrix never exists in HBM:
//
//   1. cursor tier (rare features): every band record keeps a cursor into its
//      row-sorted inverted list; per chunk of up to 1024 tiles the records
//      walk their lists (count, then scatter) and bucket their products by J
//      tile in a per-workgroup HBM region; each tile pair adds its bucket;
//   2. bitmap tier (frequent features): every band entry reads its feature's
//      segment in the J tile block (segment starts by bitmap index) and adds
//      the products with integer LDS atomics;
//   3. dense head (most frequent features): the band's and the tile's 256-byte
//      head records are contracted on the matrix cores, as exact int8 limbs or
//      in BF16;
//   4. head (p > 0): the prefix-norm bound partial + |x_head||y_head| >= t
//      (rigorous, with the FP32 error bound) selects the cells whose
//      forward-list head dot is computed; others keep their partial score;
//   5. threshold: each lane reads and clears its eight cells and settles the
//      ones that can reach t against their rigorous error bound: certain hits
//      go to the output, cells within the bound to the band list that the FP64
//      rescore kernel settles exactly.  Both go through LDS staging buffers
//      and reach HBM with one global atomic per buffer per flush.
#include "apss_internal.h"

namespace apss {

__device__ __forceinline__ float head_dot(const int32_t* __restrict__ ac,
                                          const float* __restrict__ av, int64_t an,
                                          const int32_t* __restrict__ bc,
                                          const float* __restrict__ bv, int64_t bn, int32_t p) {
  // forward-list dot of the head (ids < p) of two rows, ascending ids
  float s = 0.f;
  int64_t x = 0, y = 0;
  while (x < an && y < bn) {
    int32_t cx = ac[x], cy = bc[y];
    if (cx >= p || cy >= p) break;
    if (cx < cy) ++x;
    else if (cx > cy) ++y;
    else { s += av[x] * bv[y]; ++x; ++y; }
  }
  return s;
}

// Tile-block staging.  Block sizes are padded to multiples of 4 entries so
// the entries move as 32-bit / 128-bit words.  The loads for the next tile
// are issued into registers before the current tile's product phase and
// written to LDS after it, so their latency hides behind the products.
// LDS staging buffers of emitted pairs and of error-band pairs.  The dense-head
// kernels keep the first one small and give the LDS to a larger bitmap tier
// (their tile blocks take it).
template <int DM>
constexpr int obuf_cap() { return DM == DENSE_OFF ? 512 : 256; }
template <int DM>
constexpr int bbuf_cap() { return 256; }
static_assert(KD_MAX < 65536 && CAP_J < 65536, "16-bit segment starts");

// Per-mode sizes: bitmap-tier entries per thread of a band and of a tile.
template <int DM>
struct Sizes {
  static constexpr int CI = cap_i_for(DM), CJ = cap_j_for(DM), KDM = kd_max_for(DM);
  static constexpr int IPT = CI / SCAN_THREADS;
  static constexpr int QPT = (CJ / 4 + SCAN_THREADS - 1) / SCAN_THREADS;
  static_assert(CI % SCAN_THREADS == 0 && IPT % 2 == 0, "band entries split in two halves");
  static_assert(KDM + 1 <= 4 * SCAN_THREADS, "one 4-entry chunk of segment starts per thread");
};

// Block metadata of one J tile: first bitmap-tier entry (a multiple of 4) and
// first row, entries / 4 and rows.  Loaded one tile ahead of its prefetch.
struct TileMeta {
  int64_t d0, r0;
  int32_t nd4, nrows;
};

// Read-only arrays read with a uniform index go through the constant address
// space, which the compiler loads with scalar loads into scalar registers
// (counted by the LDS/scalar counter, so waiting for them never waits for the
// vector loads of the prefetch).
typedef __attribute__((address_space(4))) const int64_t* const_i64_ptr;
__device__ __forceinline__ int64_t sload_i64(const int64_t* p, int64_t i) {
  return ((const_i64_ptr)(p))[i];
}

__device__ __forceinline__ TileMeta tile_meta(const TileSide& S, int32_t t) {
  TileMeta m;
  m.d0 = sload_i64(S.dbase, t);
  m.r0 = sload_i64(S.tstart, t);
  m.nd4 = (int32_t)((sload_i64(S.dbase, t + 1) - m.d0) >> 2);
  m.nrows = (int32_t)(sload_i64(S.tstart, t + 1) - m.r0);
  return m;
}

template <int DM>
struct Prefetch {
  uint2 seg;                      // four 16-bit segment starts
  uint32_t rows[Sizes<DM>::QPT];
  float4 vals[Sizes<DM>::QPT];
  uint4 dense;                    // one 16-byte chunk of the tile's head records
  int2 rq;                        // settle bounds {E, Q} of one row
  uint32_t aux;                   // one row value: head norm (split p > 0) or head-part norm (BF16)
  TileMeta m;
};

// Prefetch of tile t into registers.  Every load is unconditional, from an
// index clamped to a valid element when the lane has nothing to fetch, so no
// loaded value is merged with a default before its use (such merges made the
// compiler wait for the loads on the spot); validity is applied when the
// registers are written to LDS.
template <int DM, bool HM>
__device__ __forceinline__ void prefetch_tile(const TileSide& S, const TileMeta& m, int32_t t,
                                              int tid, Prefetch<DM>& pf) {
  pf.m = m;
  // uniform tile bases (scalar registers) and 32-bit lane offsets
  const uint2* segb = reinterpret_cast<const uint2*>(S.dseg + (int64_t)t * S.kds);
  pf.seg = segb[4 * tid < S.kds ? (unsigned)tid : 0u];
  const uint32_t* rowb = reinterpret_cast<const uint32_t*>(S.drow) + (m.d0 >> 2);
  const float4* valb = reinterpret_cast<const float4*>(S.dval) + (m.d0 >> 2);
#pragma unroll
  for (int i = 0; i < Sizes<DM>::QPT; ++i) {
    const int k = tid + i * SCAN_THREADS;
    const unsigned q = k < m.nd4 ? (unsigned)k : 0u;
    pf.rows[i] = rowb[q];
    pf.vals[i] = valb[q];
  }
  if constexpr (DM != DENSE_OFF) {  // thread -> (row tid / 16, chunk tid % 16)
    const int row = tid >> 4;
    const uint4* db = S.dense + m.r0 * 16;
    pf.dense = db[(row < m.nrows ? (unsigned)row : 0u) * 16u + (unsigned)(tid & 15)];
  }
  // per-row values, one row per lane: the settle bounds, and the head norm
  // (second wave) or the head-part norm (fourth wave) where the mode uses them
  const int row = tid & 63;
  const unsigned r = row < m.nrows ? (unsigned)row : 0u;
  pf.rq = (S.rq + m.r0)[r];
  if constexpr (HM || DM == DENSE_BF16) {
    const int which = __builtin_amdgcn_readfirstlane(tid >> 6);
    const float* src = (HM && which == 1) ? S.hnorm : DM == DENSE_BF16 ? S.dnorm : S.hnorm;
    pf.aux = __float_as_uint((src + m.r0)[r]);
  }
}

// Saturating conversion of an integral double to int32 (thresholds in
// fixed point: a clamped value only makes a settle decision stricter).
__device__ __forceinline__ int32_t sat_i32(double x) {
  return x >= 2147483647.0 ? INT32_MAX : x <= -2147483647.0 ? -2147483647 : (int32_t)x;
}

// BF16 head: the per-cell part c_bf |x_H| |y_H| of the error bound, in
// fixed-point units, rounded up (cbf_q = c_bf 2^S, rounded up, with a margin
// for the two FP32 products).
__device__ __forceinline__ int32_t bf_cell(float cbf_q, float dn_i, float dn_j) {
  return (int32_t)ceilf(cbf_q * dn_i * dn_j * (1.0f + 0x1p-20f)) + 1;
}

// BF16 head with no negative value on either side: every product of a cell
// is non-negative, so the head part A of its score is at most the whole score
// s.  With q the cell's fixed-point score and e its other error terms,
// s <= (q + e) 2^-S + c_bf A gives A <= (q + e) 2^-S / (1 - c_bf), so the head
// error c_bf A is at most c_bf / (1 - c_bf) (q + e) fixed-point units
// (cabs32 = that factor x 2^32 rounded up; +2 covers the floor of the high
// product and a margin).  This replaces the norm bound c_bf |x_H||y_H|, which
// is wider for the cells near t.
__device__ __forceinline__ int32_t bf_abs(uint32_t cabs32, int32_t q, int64_t e) {
  const int64_t qe = (int64_t)q + e;
  const uint32_t u = qe <= 0 ? 0u : qe >= (int64_t)0xffffffffLL ? 0xffffffffu : (uint32_t)qe;
  return (int32_t)__umulhi(u, cabs32) + 2;
}

// The J side of the kernel's arguments, copied from the kernarg segment where
// it is used: the opaque pointer keeps the compiler from hoisting the loads
// out of the tile loop, so these pointers hold no scalar registers between
// their uses (fewer scalar registers spill to vector-register lanes).
typedef __attribute__((address_space(4))) const ScanArgs* KArgs;
__device__ __forceinline__ TileSide kargs_J() {
  KArgs p = (KArgs)__builtin_amdgcn_kernarg_segment_ptr();
  asm volatile("" : "+s"(p));
  return ((const ScanArgs*)(const void*)p)->J;
}

// Workgroup barrier for hand-offs through LDS only: it waits for this wave's
// LDS operations but not for its global loads and stores in flight (the
// prefetch of the next tile, output stores), which __syncthreads() would.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

typedef int v4i __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

// Head records sit in LDS as 16 chunks of 16 bytes per row, chunk c of row r
// at slot c ^ (r % 16): the 16 rows one MFMA operand load touches then spread
// over the banks.
__device__ __forceinline__ int dslot(int row, int chunk) { return row * 16 + (chunk ^ (row & 15)); }

// Exclusive scan over the workgroup (one value per thread); *total gets the
// sum.  wsum holds 17 words of LDS.
__device__ __forceinline__ uint32_t block_exscan(uint32_t v, int tl, uint32_t* wsum,
                                                 uint32_t* total) {
  const int lane = lane_id(), w = tl >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  if (tl == 0) {
    uint32_t acc = 0;
    for (int k = 0; k < SCAN_THREADS / 64; ++k) {
      const uint32_t c = wsum[k];
      wsum[k] = acc;
      acc += c;
    }
    wsum[SCAN_THREADS / 64] = acc;
  }
  __syncthreads();
  *total = wsum[SCAN_THREADS / 64];
  return wsum[w] + x - v;
}

// Cursor walks over one record's list from position cur, visiting the entries
// whose J tile lies below ce (lists are row-sorted, so those come first).
// Entries are read as aligned 16-byte words, WALK_BATCH per step from the
// multiple of four at or below cur: a lane's step touches a few cache lines
// instead of WALK_BATCH separate words.  Entries before cur are skipped (first
// step only), entries at or after end stop the walk; the list arrays are padded
// by WALK_PAD entries, so a step never reads past them.  Callers guarantee
// cur < end.
constexpr int WALK_BATCH = 16;
static_assert(WALK_BATCH % 4 == 0 && WALK_PAD >= WALK_BATCH,
              "a walk step reads up to WALK_BATCH - 1 entries past end - 1");

__device__ __forceinline__ void walk_count(const uint32_t* __restrict__ pos, int cur, int end,
                                           int32_t cs, int32_t ce, uint32_t* bkt,
                                           unsigned int& n_read) {
  const uint4* pos4 = reinterpret_cast<const uint4*>(pos);
  for (int b = cur & ~3;; b += WALK_BATCH) {
    uint32_t pt[WALK_BATCH];
#pragma unroll
    for (int w = 0; w < WALK_BATCH / 4; ++w) {
      const uint4 x = pos4[(b >> 2) + w];
      pt[4 * w] = x.x; pt[4 * w + 1] = x.y; pt[4 * w + 2] = x.z; pt[4 * w + 3] = x.w;
    }
    bool go = true;
#pragma unroll
    for (int q = 0; q < WALK_BATCH; ++q) {
      if (b + q < cur) continue;
      const int32_t t = b + q < end ? (int32_t)(pt[q] >> 8) : INT32_MAX;
      n_read += (go && b + q < end) ? 1u : 0u;  // list entries this walk reads
      go = go && t < ce;
      if (go) atomicAdd(&bkt[t - cs], 1u);
    }
    if (!go) return;
  }
}

// Writes the products and returns the first position not consumed; *t_next
// receives the J tile of that position (INT32_MAX at the list end).
__device__ __forceinline__ int walk_scatter(const uint32_t* __restrict__ pos,
                                            const float* __restrict__ val, int cur, int end,
                                            int32_t cs, int32_t ce, int il, float x,
                                            uint32_t* bfill, int2* __restrict__ cprod,
                                            unsigned int& n_prod, int32_t* t_next,
                                            unsigned int& n_read) {
  const uint4* pos4 = reinterpret_cast<const uint4*>(pos);
  const float4* val4 = reinterpret_cast<const float4*>(val);
  int consumed = 0;
  for (int b = cur & ~3;; b += WALK_BATCH) {
    uint32_t pt[WALK_BATCH];
    float ev[WALK_BATCH];
#pragma unroll
    for (int w = 0; w < WALK_BATCH / 4; ++w) {
      const uint4 x = pos4[(b >> 2) + w];
      pt[4 * w] = x.x; pt[4 * w + 1] = x.y; pt[4 * w + 2] = x.z; pt[4 * w + 3] = x.w;
      const float4 v = val4[(b >> 2) + w];
      ev[4 * w] = v.x; ev[4 * w + 1] = v.y; ev[4 * w + 2] = v.z; ev[4 * w + 3] = v.w;
    }
    bool go = true;
    int used = 0;
    int32_t stop_t = INT32_MAX;
#pragma unroll
    for (int q = 0; q < WALK_BATCH; ++q) {
      if (b + q < cur) continue;
      const int32_t t = b + q < end ? (int32_t)(pt[q] >> 8) : INT32_MAX;
      const bool was = go;
      go = go && t < ce;
      if (go) {
        const uint32_t slot = atomicAdd(&bfill[t - cs], 1u);
        cprod[slot] = make_int2(il + (int)(pt[q] & 0xffu) * TSTRIDE, (int)(x * ev[q]));
        ++used;
      } else if (was) {
        stop_t = t;
      }
    }
    consumed += used;
    if (!go) {
      n_prod += (unsigned)consumed;
      *t_next = stop_t;
      n_read += (unsigned)consumed + (cur + consumed < end ? 1u : 0u);
      return cur + consumed;
    }
  }
}

// LDS staging of the workgroup's results: emitted pairs and error-band pairs
// collect here (one LDS atomic per wave and round) and move to HBM in one
// reserved range per buffer, so the global counters see one atomic per flush.
template <int OC, int BC>
struct StageBufs {
  static constexpr int OUT_CAP = OC, BAND_CAP = BC;
  apss_pair out[OC];
  unsigned long long band[BC];
  unsigned int n_out, n_band;          // may exceed the capacity: the excess went to HBM directly
  unsigned long long base_out, base_band;
};

// Copies both buffers to HBM and empties them.  Called by the whole workgroup
// with the counts stable (after a barrier).
template <class SB>
__device__ __forceinline__ void flush_stage(const ScanArgs& a, int tl, SB& sb) {
  const unsigned int no = sb.n_out < (unsigned)SB::OUT_CAP ? sb.n_out : (unsigned)SB::OUT_CAP;
  const unsigned int nb = sb.n_band < (unsigned)SB::BAND_CAP ? sb.n_band : (unsigned)SB::BAND_CAP;
  if (tl == 0) {
    sb.base_out = no ? atomicAdd(a.out_cnt, (unsigned long long)no) : 0ull;
    sb.base_band = nb ? atomicAdd(a.band_cnt, (unsigned long long)nb) : 0ull;
  }
  lds_barrier();
  const unsigned long long ob = sb.base_out, bb = sb.base_band;
  for (unsigned int k = tl; k < no; k += SCAN_THREADS)
    if (ob + k < (unsigned long long)a.out_cap) a.out[ob + k] = sb.out[k];
  for (unsigned int k = tl; k < nb; k += SCAN_THREADS)
    if (bb + k < (unsigned long long)a.band_cap) a.band[bb + k] = sb.band[k];
  lds_barrier();
  if (tl == 0) {
    sb.n_out = 0;
    sb.n_band = 0;
  }
}

// Appends this lane's record (when want) to an LDS staging buffer with one
// LDS atomic per wave; the part of a reservation beyond the capacity goes
// straight to HBM through the global counter.
template <class T>
__device__ __forceinline__ void stage_append(bool want, const T& rec, T* buf, unsigned int cap,
                                             unsigned int* n, unsigned long long* counter,
                                             T* dst, long long dst_cap) {
  const unsigned long long m = __ballot(want);
  if (!m) return;
  const int tot = __popcll(m);
  const int leader = __ffsll((long long)m) - 1;
  unsigned int lb = 0;
  if (lane_id() == leader) lb = atomicAdd(n, (unsigned)tot);
  lb = (unsigned)__builtin_amdgcn_readlane((int)lb, leader);  // leader is wave-uniform
  const unsigned int pos = lb + (unsigned)lane_rank(m);
  if (want && pos < cap) buf[pos] = rec;
  if (lb + tot > cap) {
    int t2;
    const bool spill = want && pos >= cap;
    const long long slot = wave_reserve(counter, spill, &t2);
    if (spill && slot < dst_cap) dst[slot] = rec;
  }
}

// Persistent band kernel: one workgroup per CU takes bands from a queue.
// DM selects the dense head (DENSE_OFF, DENSE_I8, DENSE_BF16).
template <int DM, bool HM, bool SP, bool ST>
__global__ __launch_bounds__(SCAN_THREADS) void k_band_scan(ScanArgs a) {
  using Z = Sizes<DM>;
  __shared__ __align__(16) int32_t tile[TILE_ROWS * TSTRIDE];  // fixed-point scores (x 2^S), transposed
  __shared__ __align__(8) uint16_t cJ[4 * SCAN_THREADS];  // segment starts of tile J by bitmap index
  __shared__ __align__(16) int2 eJ[Z::CJ];           // {local row, value bits}
  __shared__ uint32_t bkt[CH_TILES + 1];             // cursor-product bucket starts of the chunk
  __shared__ uint32_t bfill[CH_TILES];               // bucket fill pointers while scattering
  __shared__ uint32_t wsum[SCAN_THREADS / 64 + 1];
  __shared__ StageBufs<obuf_cap<DM>(), bbuf_cap<DM>()> stage;  // results waiting for a flush
  // per-row values of tile J in two buffers (by J % 2): a wave that settles the
  // cells of tile J may still read them while the next tile is written
  // settle bounds: band rows {E, Q, ceil(t 2^S), floor(t 2^S)}, tile rows {E, Q}
  // (see k_row_bounds); head norms for the split p > 0
  __shared__ int4 rqI[BAND_ROWS];
  __shared__ int2 rqJ[2][TILE_ROWS];
  __shared__ float hnI[HM ? BAND_ROWS : 1], hnJ[2][HM ? TILE_ROWS : 1];
  __shared__ float vI[Z::CI];                        // band entry values * 2^S, [q][tid]
  __shared__ uint4 dA[DM ? BAND_ROWS * 16 : 1];      // head records of the band
  __shared__ uint4 dB[DM ? TILE_ROWS * 16 : 1];      // head records of tile J
  __shared__ float dnI[DM == DENSE_BF16 ? BAND_ROWS : 1], dnJ[2][DM == DENSE_BF16 ? TILE_ROWS : 1];
  __shared__ int4 item_s;
  __shared__ int32_t ch_end_s;
  __shared__ unsigned int rec_next;                  // cursor records handed out so far
  __shared__ unsigned long long acc_products, acc_head, acc_visits, acc_cursor;
  // entries staged once per work item (band bitmap-tier entries, cursor
  // records) and cursor-list entries read by the count and scatter walks
  __shared__ unsigned long long acc_band_e, acc_band_s, acc_walk_c, acc_walk_s;
  __shared__ unsigned long long bsum;  // sparse bands: sum of |products| of one tile
  static_assert(CH_TILES == SCAN_THREADS, "one bucket per thread in the chunk scan");
  static_assert(BAND_ROWS == 128 && TILE_ROWS == 64 && SCAN_THREADS == 1024,
                "MFMA layout: 16 waves x two 16x16 blocks cover the 128 x 64 tile");

  const int tid0 = threadIdx.x;
  int2* const cprod = a.cprod + (int64_t)blockIdx.x * a.cprod_cap;
  int4* const scs = a.scs + (int64_t)blockIdx.x * a.scs_cap;

  // the tile is cleared cell by cell by the threshold pass, so once is enough
  for (int c = tid0; c < TILE_ROWS * TSTRIDE; c += SCAN_THREADS) tile[c] = 0;
  if (tid0 == 0) {
    acc_products = 0; acc_head = 0; acc_visits = 0; acc_cursor = 0;
    acc_band_e = 0; acc_band_s = 0; acc_walk_c = 0; acc_walk_s = 0;
    stage.n_out = 0;
    stage.n_band = 0;
  }
  unsigned int my_products = 0, my_visits = 0;
  unsigned long long n_skip = 0, n_bound = 0;  // sparse-band tile pairs skipped (thread 0)

  for (;;) {
  int tid = tid0;
  asm volatile("" : "+v"(tid));
  // work item: band I against the partner tiles [J_first, J_end)
  if (tid == 0) {
    const unsigned long long w = atomicAdd(a.item_next, 1ull);
    item_s = w < (unsigned long long)a.n_items ? a.items[w] : make_int4(-1, 0, 0, 0);
  }
  __syncthreads();
  // uniform values: the band, its rows and the partner range live in SGPRs
  const int32_t I = __builtin_amdgcn_readfirstlane(item_s.x);
  if (I < 0) break;
  const int64_t i0 = a.I.tstart[I];
  const int nI = (int)(a.I.tstart[I + 1] - i0);
  const int32_t J_first = __builtin_amdgcn_readfirstlane(item_s.y);
  const int32_t J_end = __builtin_amdgcn_readfirstlane(item_s.z);
  // batched runs and packed queries give each band its own threshold
  const double t_band = a.band_t ? a.band_t[I] : a.t;
  int32_t q_lo = a.q_lo;
  if (a.band_t) {
    const double ql = floor((t_band - 2.0 * a.eps_max) * (double)a.qscale);
    q_lo = (int32_t)fmax(-2147483647.0, fmin(2147483647.0, ql));
  }

  // ---- band set-up: this thread's band entries live in registers ---------
  // Entry k = tid + q * SCAN_THREADS of the band block: (band row, bitmap
  // index, value * 2^S).  The block is ordered by frequency rank, so
  // the lanes of a wave hold entries of the same or similar features.
  int32_t e_rf[Z::IPT];  // band row << 12 | bitmap index (KD past the block)
  const int nE = a.I.dseg[(int64_t)I * a.I.kds + a.KD];  // entries of the band block
  if constexpr (!SP) {
    const int64_t d0 = a.I.dbase[I];
#pragma unroll
    for (int q = 0; q < Z::IPT; ++q) {
      const int k = tid + q * SCAN_THREADS;
      const bool ok = k < nE;
      // bit 26 marks the first entry of a feature's band segment (it counts
      // the tile segment once in the list-visit statistics)
      const int32_t f = ok ? (int32_t)a.I.dfeat[d0 + k] : 0;
      const bool seg_head = ok && (k == 0 || (int32_t)a.I.dfeat[d0 + k - 1] != f);
      // entries past the block point at bitmap index KD, whose segment is
      // empty in every tile (the segment-start rows end with the block total)
      e_rf[q] = ok ? ((((int32_t)a.I.drow[d0 + k]) << 12) | f | (seg_head ? (1 << 26) : 0)) : a.KD;
      vI[k] = ok ? a.I.dval[d0 + k] * a.qscale : 0.f;  // exact: power of 2
    }
  }
  if (tid < BAND_ROWS) {
    const bool ok = tid < nI;
    const int64_t r = i0 + tid;
    const int2 rq = ok ? a.I.rq[r] : make_int2(0, 0);
    // the row's threshold in fixed point (exact: 2^S is a power of two)
    const double tq = ((ok && a.row_t) ? a.row_t[r] : t_band) * (double)a.qscale;
    rqI[tid] = make_int4(rq.x, rq.y, sat_i32(ceil(tq)), sat_i32(floor(tq)));
    if constexpr (HM) hnI[tid] = ok ? a.I.hnorm[r] : 0.f;
    if constexpr (DM == DENSE_BF16) dnI[tid] = ok ? a.I.dnorm[r] : 0.f;
  }
  if constexpr (DM != DENSE_OFF) {  // the band's head records (zero beyond its rows)
    for (int c = tid; c < BAND_ROWS * 16; c += SCAN_THREADS) {
      const int row = c >> 4;
      dA[dslot(row, c & 15)] = row < nI ? a.I.dense[(i0 + row) * 16 + (c & 15)] : make_uint4(0, 0, 0, 0);
    }
  }
  // cursor records: the first position whose J tile is >= J_first (lists are
  // row-sorted, so positions are ascending in tile << 8 | row) and its tile
  const int64_t sb = a.I.sbase[I];
  const int nS = (int)(a.I.sbase[I + 1] - sb);
  if (tid == 0) {
    acc_band_e += (unsigned long long)a.I.dseg[(int64_t)I * a.I.kds + a.KD];
    acc_band_s += (unsigned long long)nS;
  }
  for (int k = tid; k < nS; k += SCAN_THREADS) {
    const int4 r = a.I.srec[sb + k];
    int lo = r.z, hi = r.w;
    const uint32_t key = (uint32_t)J_first << 8;
    while (lo < hi) {
      const int m = (lo + hi) >> 1;
      if (a.S.pos[m] < key) lo = m + 1; else hi = m;
    }
    scs[k] = make_int4(lo, r.w, lo < r.w ? (int32_t)(a.S.pos[lo] >> 8) : INT32_MAX, 0);
  }
  __syncthreads();

  int32_t ch_lo = J_first, ch_hi = nS > 0 ? J_first : J_end;  // bucket chunk [ch_lo, ch_hi)

  // ---- 1. cursor tier: bucket the products of the next chunk of tiles ----
  // Every record walks its row-sorted list up to the chunk end twice: once
  // to count the products per J tile, once to write them (fixed point, with
  // their cell) into this workgroup's region, grouped by tile.  Threads take
  // records one at a time from a shared counter, so long lists spread over
  // the workgroup.  A chunk whose products exceed the region ends early at a
  // tile boundary.
  // The chunk starts at tile Jc; ch_lo / ch_hi receive its range.
  auto walk_chunk = [&](int32_t Jc, int tl) {
    ch_lo = Jc;
    const int32_t ce0 = Jc + CH_TILES < J_end ? Jc + CH_TILES : J_end;
    bkt[tl] = 0;
    if (tl == 0) { bkt[CH_TILES] = 0; ch_end_s = ce0; rec_next = 0; }
    __syncthreads();
    unsigned int n_rd = 0;
    for (;;) {
      const int k = (int)atomicAdd(&rec_next, 1u);
      if (k >= nS) break;
      const int4 st = scs[k];
      if (st.z < ce0) walk_count(a.S.pos, st.x, st.y, ch_lo, ce0, bkt, n_rd);
    }
    if (n_rd) atomicAdd(&acc_walk_c, (unsigned long long)n_rd);
    __syncthreads();
    if (tl == 0) rec_next = 0;
    const uint32_t cnt = bkt[tl];
    uint32_t total;
    const uint32_t ex = block_exscan(cnt, tl, wsum, &total);
    bkt[tl] = ex;
    bfill[tl] = ex;
    if (tl == 0) bkt[CH_TILES] = total;
    if ((int64_t)ex + cnt > a.cprod_cap) atomicMin(&ch_end_s, ch_lo + tl);
    __syncthreads();
    const int32_t ce = ch_end_s;
    for (;;) {
      const int k = (int)atomicAdd(&rec_next, 1u);
      if (k >= nS) break;
      const int4 st = scs[k];
      if (st.z >= ce) continue;
      const int4 r = a.I.srec[sb + k];
      int32_t tn;
      unsigned int np = 0, nr = 0;
      const int cur = walk_scatter(a.S.pos, a.S.val, st.x, st.y, ch_lo, ce, r.x,
                                   __int_as_float(r.y) * a.qscale, bfill, cprod, np, &tn, nr);
      scs[k] = make_int4(cur, st.y, tn, 0);
      if (np) atomicAdd(&acc_cursor, (unsigned long long)np);
      if (nr) atomicAdd(&acc_walk_s, (unsigned long long)nr);
    }
    __syncthreads();
    ch_hi = ce;
  };

  // ---- 4. threshold and emission ---------------------------------------------
  // Each thread reads and clears its eight cells (+ their dense parts) and
  // settles the ones that can reach t itself, one candidate per lane per
  // round.  Valid cells: row below the band height, column below the tile
  // height, and for self-joins i0 + il < j0 + jl.  Emitted pairs go to the
  // LDS output buffer, cells within the error bound to the band list.  No
  // barrier follows: the next tile pair's first barrier orders the cleared
  // cells and the buffer appends before their next use.
  // (tile J's first row j0, its height nJ, its per-row buffer jb; dv: the
  // cells' dense-head parts, added to the score tile's)
  auto settle = [&](int64_t j0, int nJ, int jb, int32_t (&dv)[8], int tl) {
    const int wv = tl >> 6, ln = tl & 63;
    const int il_b = 16 * (wv & 7) + 4 * (ln >> 4);
    const int jl_b = 32 * (wv >> 3) + (ln & 15);
    const int64_t dj = j0 - i0;
    // Cells past the band or tile height receive no products and a zero head,
    // so with q_lo > 0 only the self-join diagonal needs an explicit validity
    // test: off the diagonal one max over the eight cells rejects them all.
    const bool plain = q_lo > 0 && (!a.self || dj >= BAND_ROWS);
    int32_t qmax = INT32_MIN;
#pragma unroll
    for (int blk = 0; blk < 2; ++blk) {  // four band rows of one tile row: one 16-byte word
      int4* w4 = reinterpret_cast<int4*>(&tile[(jl_b + 16 * blk) * TSTRIDE + il_b]);
      const int4 c4 = *w4;
      *w4 = make_int4(0, 0, 0, 0);
      dv[4 * blk + 0] += c4.x;  // full fixed-point score of the cell
      dv[4 * blk + 1] += c4.y;
      dv[4 * blk + 2] += c4.z;
      dv[4 * blk + 3] += c4.w;
    }
#pragma unroll
    for (int m = 0; m < 8; ++m) qmax = dv[m] > qmax ? dv[m] : qmax;
    unsigned cmask = 0;
    if (plain) {
      // off the self-join diagonal, with q_lo > 0: cells past the band or
      // tile height hold zero and fail the test by themselves
      if (__ballot(qmax >= q_lo)) {
#pragma unroll
        for (int m = 0; m < 8; ++m) cmask |= dv[m] >= q_lo ? (1u << m) : 0u;
      }
    } else {
      // valid rows of this thread's two blocks: v < lim[blk] (v = m % 4)
      int lim[2];
#pragma unroll
      for (int blk = 0; blk < 2; ++blk) {
        const int jl = jl_b + 16 * blk;
        int l = jl < nJ ? nI - il_b : 0;
        if (a.self) {  // i0 + il < j0 + jl  <=>  il - il_b < dj + jl - il_b
          const int64_t dl = dj + jl - il_b;
          l = dl < l ? (int)(dl < 0 ? 0 : dl) : l;
        }
        lim[blk] = l < 0 ? 0 : (l > 4 ? 4 : l);
      }
#pragma unroll
      for (int m = 0; m < 8; ++m)
        cmask |= ((m & 3) < lim[m >> 2] && dv[m] >= q_lo) ? (1u << m) : 0u;
    }
    while (__ballot(cmask != 0)) {  // rare: rounds over the candidate cells
      const bool is_cand = cmask != 0;
      const int m = is_cand ? __builtin_ctz(cmask) : 0;
      cmask &= cmask - 1u;
      int32_t q = dv[0];
#pragma unroll
      for (int k = 1; k < 8; ++k) q = m == k ? dv[k] : q;
      const int il = il_b + (m & 3), jl = jl_b + 16 * (m >> 2);
      bool emit = false, to_band = false;
      if (is_cand) {
        // the cell's error bound e (fixed point) is the smaller of its two
        // rows' bounds; with rqI.z = ceil(t 2^S) and rqI.w = floor(t 2^S):
        // q - e >= ceil(t 2^S) makes s >= t certain, q + e < floor(t 2^S)
        // makes s < t certain, and between them the FP64 rescore decides
        const int4 bi = rqI[il];
        const int2 bj = rqJ[jb][jl];
        int64_t e = bi.x < bj.x ? bi.x : bj.x;
        if constexpr (DM == DENSE_BF16)
          e += a.bf_nonneg ? bf_abs(a.cabs32, q, e) : bf_cell(a.cbf_q, dnI[il], dnJ[jb][jl]);
        if ((int64_t)q - e >= (int64_t)bi.z) {
          // certain hit; its score is final only if the bound meets the
          // relative tolerance (q >= the rows' Q, which is e scaled by it)
          if (a.rescore_all || q < (bi.y < bj.y ? bi.y : bj.y)) to_band = true; else emit = true;
        } else if ((int64_t)q + e >= (int64_t)bi.w) {
          to_band = true;
        }
      }
      // emitted pairs and error-band pairs go to the LDS staging buffers
      apss_pair pr;
      pr.i = (uint32_t)(i0 + il + a.i_off);
      pr.j = (uint32_t)(j0 + jl + a.j_off);
      pr.score = (float)q * a.qinv_f;  // = (float)(q 2^-S): 2^-S is a power of two
      stage_append(emit, pr, stage.out, obuf_cap<DM>(), &stage.n_out, a.out_cnt, a.out, a.out_cap);
      const unsigned long long key = ((unsigned long long)(i0 + il) << 32) | (unsigned long long)(j0 + jl);
      stage_append(to_band, key, stage.band, bbuf_cap<DM>(), &stage.n_band, a.band_cnt, a.band, a.band_cap);
    }
  };

  if constexpr (SP) {
    // ---- sparse band (SP): tiles with cursor products only -------------------
    // The host sends here the bands whose rows hold cursor-tier features only
    // (no bitmap-tier entry, no dense-head or head feature) and whose
    // thresholds are above zero (q_lo > 0).  Their cells receive cursor
    // products only, so a tile with no product in the band's bucket holds no
    // pair that can reach t (the two blocks share no feature) and is skipped
    // whole.
    // Per chunk of tiles: bucket the products (walk_chunk), then for each tile
    // with a non-empty bucket add its products into the score tile and settle
    // it -- unless a block bound shows no cell can be a candidate: the L1 x
    // max bound of the two blocks (k_tile_l1max) or, once the products are
    // in, their sum of magnitudes (every cell's score is at most that sum).
    // The tile block, the bitmap tier and the dense head are never touched.
    int tl = tid;
    asm volatile("" : "+v"(tl));
    const double qs = (double)a.qscale;
    const float2 bI = a.I.l1m[I];
    unsigned long long my_skip = 0, my_bound = 0;
    for (int32_t J = J_first; J < J_end;) {
      if (nS > 0) walk_chunk(J, tl);
      const int32_t ce = nS > 0 ? ch_hi : J_end;
      for (; J < ce; ++J) {
        const int32_t pb = nS > 0 ? (int32_t)bkt[J - ch_lo] : 0;
        const int32_t pe = nS > 0 ? (int32_t)bkt[J - ch_lo + 1] : 0;
        if (pb == pe) {  // the blocks share no feature
          ++my_skip;
          continue;
        }
        const float2 bJ = a.J.l1m[J];
        const double l1m = fmin((double)bI.x * (double)bJ.y, (double)bI.y * (double)bJ.x);
        if ((l1m * (1.0 + 1e-6) + a.eps_max) * qs < (double)q_lo) {
          ++my_bound;
          continue;
        }
        const int64_t j0 = a.J.tstart[J];
        const int nJ = (int)(a.J.tstart[J + 1] - j0);
        if (tl < TILE_ROWS) rqJ[0][tl] = tl < nJ ? a.J.rq[j0 + tl] : make_int2(0, 0);
        if (tl == 0) bsum = 0;
        __syncthreads();
        // staging buffers three quarters full: empty them (every append of
        // the earlier tiles is complete behind the barrier above)
        if (stage.n_out >= obuf_cap<DM>() * 3 / 4 || stage.n_band >= bbuf_cap<DM>() * 3 / 4)
          flush_stage(a, tl, stage);
        unsigned long long my_abs = 0;
        for (int32_t k = pb + tl; k < pe; k += SCAN_THREADS) {
          const int2 q = cprod[k];
          atomicAdd(&tile[q.x], q.y);
          my_abs += (unsigned long long)(q.y < 0 ? -(int64_t)q.y : (int64_t)q.y);
        }
        if (my_abs) atomicAdd(&bsum, my_abs);
        __syncthreads();
        if (bsum >= (unsigned long long)q_lo) {
          int32_t dv[8] = {0, 0, 0, 0, 0, 0, 0, 0};
          settle(j0, nJ, 0, dv, tl);
        } else {  // no cell reaches q_lo: clear the cells the products touched
          for (int32_t k = pb + tl; k < pe; k += SCAN_THREADS) tile[cprod[k].x] = 0;
          ++my_bound;
        }
        __syncthreads();  // cells cleared and appends done before the next tile
      }
    }
    n_skip += my_skip;
    n_bound += my_bound;
    __syncthreads();
    continue;
  }

  Prefetch<DM> pf;
  TileMeta meta_next{};  // metadata of tile J + 1 while tile J is processed
  if (J_first < J_end) {
    prefetch_tile<DM, HM>(a.J, tile_meta(a.J, J_first), J_first, tid, pf);
    meta_next = tile_meta(a.J, J_first + 1 < J_end ? J_first + 1 : J_first);
  }
  // the band's MFMA operands stay in registers for the whole item
  v4i avp[4];
  if constexpr (DM != DENSE_OFF) {
    const int wv0 = tid >> 6, ln0 = tid & 63;
    const int rowA = 16 * (wv0 & 7) + (ln0 & 15);
#pragma unroll
    for (int q = 0; q < 4; ++q) avp[q] = __builtin_bit_cast(v4i, dA[dslot(rowA, 4 * q + (ln0 >> 4))]);
  }

  for (int32_t J = J_first; J < J_end; ++J) {
    // per-thread addresses are recomputed from an opaque thread index in every
    // pass instead of being kept live across the loop (register budget)
    int tl = tid;
    asm volatile("" : "+v"(tl));

    if (J == ch_hi) walk_chunk(J, tl);

    const int64_t j0 = pf.m.r0;
    const int nJ = pf.m.nrows;

    // ---- write the prefetched tile J into LDS --------------------------------
    const int32_t kds = DM != DENSE_BF16 ? kargs_J().kds : a.J.kds;
    if (4 * tl < kds) reinterpret_cast<uint2*>(cJ)[tl] = pf.seg;
#pragma unroll
    for (int i = 0; i < Z::QPT; ++i) {
      const int k = tl + i * SCAN_THREADS;
      if (k < pf.m.nd4) {
        const float vv[4] = {pf.vals[i].x, pf.vals[i].y, pf.vals[i].z, pf.vals[i].w};
#pragma unroll
        for (int q = 0; q < 4; ++q)
          eJ[4 * k + q] = make_int2((int)((pf.rows[i] >> (8 * q)) & 0xffu) * TSTRIDE, __float_as_int(vv[q]));
      }
    }
    const int jb = J & 1;  // buffer of this tile's per-row values
    if (tl < 4 * TILE_ROWS) {  // rows past the tile read as zero
      const int row = tl & 63, which = tl >> 6;
      const bool ok = row < nJ;
      if (which == 0) rqJ[jb][row] = ok ? pf.rq : make_int2(0, 0);
      if constexpr (HM)
        if (which == 1) hnJ[jb][row] = ok ? __uint_as_float(pf.aux) : 0.f;
      if constexpr (DM == DENSE_BF16)
        if (which == 3) dnJ[jb][row] = ok ? __uint_as_float(pf.aux) : 0.f;
    }
    if constexpr (DM != DENSE_OFF)
      dB[dslot(tl >> 4, tl & 15)] = (tl >> 4) < nJ ? pf.dense : make_uint4(0, 0, 0, 0);

    lds_barrier();

    // flush the staging buffers once one is three quarters full: every append of the
    // earlier tile pairs is complete here (the next appends come after the
    // MFMA barrier, which also orders the reset of the counts before them)
    if (stage.n_out >= obuf_cap<DM>() * 3 / 4 || stage.n_band >= bbuf_cap<DM>() * 3 / 4)
      flush_stage(a, tl, stage);

    // this thread's first cursor product of tile J: its load completes behind
    // the bitmap-tier products.  It is issued before the prefetch of the next
    // tile: loads complete in order, so waiting for it does not wait for the
    // prefetch
    int32_t pb = 0, pe = 0;
    if (nS > 0) { pb = (int32_t)bkt[J - ch_lo]; pe = (int32_t)bkt[J - ch_lo + 1]; }
    const int32_t k0 = pb + tl;
    const int2 hp = cprod[k0 < pe ? k0 : 0];  // unconditional (see prefetch_tile)

    // prefetch tile J+1 (registers; written to LDS at the top of the next
    // pass) and the metadata of tile J+2.  Both are unconditional (the last
    // tile is fetched again at the end of the range), so the code between
    // here and the uses is straight-line and the compiler waits for exactly
    // the loads it needs.
    {
      const int32_t t1 = J + 1 < J_end ? J + 1 : J;
      if constexpr (DM != DENSE_BF16) {
        // (1.3% faster for the int8 kernel; the BF16 kernel's copy would
        // live in scratch memory, so it reads the hoisted argument)
        const TileSide Js = kargs_J();
        prefetch_tile<DM, HM>(Js, meta_next, t1, tl, pf);
        meta_next = tile_meta(Js, J + 2 < J_end ? J + 2 : J_end - 1);
      } else {
        prefetch_tile<DM, HM>(a.J, meta_next, t1, tl, pf);
        meta_next = tile_meta(a.J, J + 2 < J_end ? J + 2 : J_end - 1);
      }
    }

    // ---- 2. bitmap tier: every band entry meets its feature's J segment -----
    // Two passes of half the entries: the LDS lookups of a pass are
    // issued together, then its segments are walked.
#pragma unroll
    for (int h = 0; h < Z::IPT; h += Z::IPT / 2) {
      int y0[Z::IPT / 2], y1[Z::IPT / 2];
#pragma unroll
      for (int q = 0; q < Z::IPT / 2; ++q) {
        int rf = e_rf[h + q];
        asm volatile("" : "+v"(rf));  // no hoisting of per-entry values out of the J loop
        const int f = rf & 0xfff;  // bitmap index (bits 0..11)
        y0[q] = cJ[f];
        y1[q] = cJ[f + 1];
      }
#pragma unroll
      for (int q = 0; q < Z::IPT / 2; ++q) {
        if (y1[q] <= y0[q]) continue;
        // tile-side list entries: counted once per feature and tile pair
        // (ST: the work counters cost about 2% of the scan, so they are a form
        // of their own)
        if constexpr (ST) my_visits += (unsigned)((e_rf[h + q] >> 26) & 1) * (unsigned)(y1[q] - y0[q]);
        const int ib = (e_rf[h + q] >> 12) & 0x3fff;  // band row
        const float u = vI[tl + (h + q) * SCAN_THREADS];
        // four entries per step: their LDS reads issue together, then the atomics
        int y = y0[q];
        for (; y + 4 <= y1[q]; y += 4) {
          const int2 e0 = eJ[y], e1 = eJ[y + 1], e2 = eJ[y + 2], e3 = eJ[y + 3];
          atomicAdd(&tile[ib + e0.x], (int)(u * __int_as_float(e0.y)));
          atomicAdd(&tile[ib + e1.x], (int)(u * __int_as_float(e1.y)));
          atomicAdd(&tile[ib + e2.x], (int)(u * __int_as_float(e2.y)));
          atomicAdd(&tile[ib + e3.x], (int)(u * __int_as_float(e3.y)));
        }
        for (; y < y1[q]; ++y) {
          const int2 e = eJ[y];
          atomicAdd(&tile[ib + e.x], (int)(u * __int_as_float(e.y)));
        }
        if constexpr (ST) my_products += (unsigned)(y1[q] - y0[q]);
      }
    }

    // ---- 1b. cursor tier: add this tile's bucket -----------------------------
    if (k0 < pe) {
      atomicAdd(&tile[hp.x], hp.y);
      for (int32_t k = k0 + SCAN_THREADS; k < pe; k += SCAN_THREADS) {
        const int2 q = cprod[k];
        atomicAdd(&tile[q.x], q.y);
      }
    }

    // ---- 2b. dense head on the matrix cores ------------------------------------
    // Wave w owns the 16x16 blocks (rows 16*(w%8).., columns 32*(w/8) + 16*blk..)
    // of the tile, blk = 0, 1; in the MFMA output layout lane l holds column
    // l%16 and rows 4*(l/16) + v, v = 0..3, i.e. cell m = 4*blk + v.  The
    // operands: lane l reads row l%16 of the block, chunk 4*s + l/16 of the
    // head record (s = limb for int8, 32-feature step for BF16); the band's
    // operands were read once per work item (avp).
    const int wv = tl >> 6, ln = tl & 63;
    int32_t dv[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) dv[m] = 0;
    if constexpr (DM != DENSE_OFF) {
      const int r15 = ln & 15, g = ln >> 4;
      v4i av[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) av[q] = avp[q];
#pragma unroll
      for (int blk = 0; blk < 2; ++blk) {
        const int rowB = 16 * (2 * (wv >> 3) + blk) + r15;
        v4i bv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) bv[q] = __builtin_bit_cast(v4i, dB[dslot(rowB, 4 * q + g)]);
        if constexpr (DM == DENSE_I8) {
          // limb products grouped by weight 2^(8*(6-w)); groups w <= 4 are
          // kept (the dropped w = 5, 6 are inside eps_dense)
          // hi = (c0 << 16) + (c1 << 8) + c2 (weight 2^32) is folded group by
          // group (shifts distribute over wrapping addition); lo = (c3 << 8) +
          // c4 (weight 2^16) is truncated by a right shift, so it is formed whole
          const v4i z = {0, 0, 0, 0};
          const v4i c0 = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[0], bv[0], z, 0, 0, 0);
          v4i c1 = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[0], bv[1], z, 0, 0, 0);
          c1 = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[1], bv[0], c1, 0, 0, 0);
          int32_t hv[4];
#pragma unroll
          for (int v = 0; v < 4; ++v) hv[v] = (c0[v] << 16) + (c1[v] << 8);
          v4i c2 = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[0], bv[2], z, 0, 0, 0);
          c2 = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[1], bv[1], c2, 0, 0, 0);
          c2 = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[2], bv[0], c2, 0, 0, 0);
#pragma unroll
          for (int v = 0; v < 4; ++v) hv[v] = (hv[v] + c2[v]) << a.dsh_hi;
          v4i c3 = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[0], bv[3], z, 0, 0, 0);
          c3 = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[1], bv[2], c3, 0, 0, 0);
          c3 = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[2], bv[1], c3, 0, 0, 0);
          c3 = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[3], bv[0], c3, 0, 0, 0);
          v4i c4 = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[1], bv[3], z, 0, 0, 0);
          c4 = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[2], bv[2], c4, 0, 0, 0);
          c4 = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[3], bv[1], c4, 0, 0, 0);
#pragma unroll
          for (int v = 0; v < 4; ++v)
            dv[4 * blk + v] = hv[v] + (((c3[v] << 8) + c4[v]) >> a.dsh_lo);
        } else {
          v4f c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int q = 0; q < 4; ++q)
            c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, av[q]),
                                                        __builtin_bit_cast(v8bf, bv[q]), c, 0, 0, 0);
#pragma unroll
          for (int v = 0; v < 4; ++v) dv[4 * blk + v] = (int)(c[v] * a.qscale);
        }
      }
    }
    lds_barrier();

    // this thread's eight cells in the MFMA output layout
    const int il_b = 16 * (wv & 7) + 4 * (ln >> 4);
    const int jl_b = 32 * (wv >> 3) + (ln & 15);

    // ---- 3. head: prefix-norm filter + forward-list dot -------------------
    if constexpr (HM) {
      unsigned int my_head = 0;
      if constexpr (DM != DENSE_OFF) {  // the filter needs the dense part in the tile
#pragma unroll
        for (int m = 0; m < 8; ++m) {
          tile[(jl_b + 16 * (m >> 2)) * TSTRIDE + il_b + (m & 3)] += dv[m];
          dv[m] = 0;
        }
        lds_barrier();
      }
#pragma unroll 1
      for (int c = tl; c < BAND_ROWS * TILE_ROWS; c += SCAN_THREADS) {
        const int jl = c / BAND_ROWS, il = c % BAND_ROWS;
        if (il >= nI || jl >= nJ || (a.self && i0 + il >= j0 + jl)) continue;
        int32_t* cell = &tile[jl * TSTRIDE + il];
        const int32_t part = *cell;
        // pruned when even the Cauchy-Schwarz head bound, on top of the
        // partial score and its error bound, cannot lift it to t (floor(t 2^S)
        // 2^-S <= t, so the test only prunes less)
        const int4 bi = rqI[il];
        int64_t e = bi.x < rqJ[jb][jl].x ? bi.x : rqJ[jb][jl].x;
        if constexpr (DM == DENSE_BF16)
          e += a.bf_nonneg ? bf_abs(a.cabs32, part, e) : bf_cell(a.cbf_q, dnI[il], dnJ[jb][jl]);
        const double bound = (double)((int64_t)part + e) * a.qinv + (double)hnI[il] * (double)hnJ[jb][jl];
        if (bound < (double)bi.w * a.qinv) continue;
        const int64_t ri = i0 + il, rj = j0 + jl;
        const int64_t ai = a.I.rp[ri], aj = a.J.rp[rj];
        const float h = head_dot(a.I.col + ai, a.I.val + ai, a.I.rp[ri + 1] - ai, a.J.col + aj,
                                 a.J.val + aj, a.J.rp[rj + 1] - aj, a.p);
        *cell = part + __float2int_rn(h * a.qscale);
        ++my_head;
      }
      if (my_head) atomicAdd(&acc_head, (unsigned long long)my_head);
      lds_barrier();
    }

    // ---- 4. threshold and emission (settle, above) --------------------------
    settle(j0, nJ, jb, dv, tl);
  }
  __syncthreads();
  }  // band queue

  // pairs still in the staging buffers (the barrier above ends every append)
  if (stage.n_out > 0 || stage.n_band > 0) flush_stage(a, tid0, stage);
  atomicAdd(&acc_products, (unsigned long long)my_products);
  atomicAdd(&acc_visits, (unsigned long long)my_visits);
  __syncthreads();
  if (tid0 == 0) {
    atomicAdd(&a.stats[0], acc_products + acc_cursor);
    atomicAdd(&a.stats[1], acc_head);
    atomicAdd(&a.stats[3], acc_visits);
    atomicAdd(&a.stats[4], acc_cursor);
    atomicAdd(&a.stats[2], acc_band_e);
    atomicAdd(&a.stats[5], acc_band_s);
    atomicAdd(&a.stats[6], acc_walk_c);
    atomicAdd(&a.stats[8], acc_walk_s);
    atomicAdd(&a.stats[STAT_SKIP_NO_COMMON], n_skip);
    atomicAdd(&a.stats[STAT_SKIP_BOUND], n_bound);
  }
}

template __global__ void k_band_scan<DENSE_OFF, false, false, false>(ScanArgs);
template __global__ void k_band_scan<DENSE_I8, false, false, false>(ScanArgs);
template __global__ void k_band_scan<DENSE_BF16, false, false, false>(ScanArgs);
template __global__ void k_band_scan<DENSE_OFF, true, false, false>(ScanArgs);
template __global__ void k_band_scan<DENSE_I8, true, false, false>(ScanArgs);
template __global__ void k_band_scan<DENSE_BF16, true, false, false>(ScanArgs);
template __global__ void k_band_scan<DENSE_OFF, false, false, true>(ScanArgs);
template __global__ void k_band_scan<DENSE_I8, false, false, true>(ScanArgs);
template __global__ void k_band_scan<DENSE_BF16, false, false, true>(ScanArgs);
template __global__ void k_band_scan<DENSE_OFF, true, false, true>(ScanArgs);
template __global__ void k_band_scan<DENSE_I8, true, false, true>(ScanArgs);
template __global__ void k_band_scan<DENSE_BF16, true, false, true>(ScanArgs);
template __global__ void k_band_scan<DENSE_OFF, false, true, false>(ScanArgs);

// FP64 rescore of the band: exact products summed in ascending feature id,
// the same order as the FP64 oracle, so the decision s >= t is settled exactly.
// Grid-stride over the band count the scan left in device memory (capped at
// the list's capacity), so the scan and the rescore follow each other on the
// stream with no host round trip and can be captured in one graph.
//
// Each lane merges its pair's two sorted rows through windows of four column
// ids and four values, read as aligned 16-byte words (the window holding the
// merge position, and the next one, loaded a window ahead) and advanced every
// fourth step of a row: the steps between, matches included, touch registers
// only.  Windows at an array's end, or of arrays that are not 16-byte aligned,
// are read entry by entry.
constexpr unsigned int RS_STAGE = 1024;  // emitted pairs staged per workgroup of the rescore
struct RowWindow {
  int4 c;     // column ids of the aligned word holding the merge position
  float4 v;   // their values
};
// the aligned 4-entry word at col + b (b a multiple of 4); `room` = entries
// of the array from b on
__device__ __forceinline__ RowWindow row_window(const int32_t* __restrict__ col,
                                                const float* __restrict__ val, int64_t b,
                                                int64_t room, bool vec) {
  RowWindow w;
  if (vec && room >= 4) {
    w.c = *reinterpret_cast<const int4*>(col + b);
    w.v = *reinterpret_cast<const float4*>(val + b);
  } else {
    w.c.x = room > 0 ? col[b] : INT32_MAX;
    w.c.y = room > 1 ? col[b + 1] : INT32_MAX;
    w.c.z = room > 2 ? col[b + 2] : INT32_MAX;
    w.c.w = room > 3 ? col[b + 3] : INT32_MAX;
    w.v.x = room > 0 ? val[b] : 0.f;
    w.v.y = room > 1 ? val[b + 1] : 0.f;
    w.v.z = room > 2 ? val[b + 2] : 0.f;
    w.v.w = room > 3 ? val[b + 3] : 0.f;
  }
  return w;
}
// entry k of a window: two levels of selects (no branches)
__device__ __forceinline__ int32_t pick_i(const int4& w, int k) {
  const int32_t lo = (k & 1) ? w.y : w.x, hi = (k & 1) ? w.w : w.z;
  return (k & 2) ? hi : lo;
}
__device__ __forceinline__ float pick_f(const float4& w, int k) {
  const float lo = (k & 1) ? w.y : w.x, hi = (k & 1) ? w.w : w.z;
  return (k & 2) ? hi : lo;
}

__global__ __launch_bounds__(256) void k_rescore(const unsigned long long* __restrict__ band,
                                                 const unsigned long long* __restrict__ band_cnt,
                                                 int64_t band_cap, const int64_t* __restrict__ rpI,
                                                 const int32_t* __restrict__ colI,
                                                 const float* __restrict__ valI, int64_t nnzI,
                                                 const int64_t* __restrict__ rpJ,
                                                 const int32_t* __restrict__ colJ,
                                                 const float* __restrict__ valJ, int64_t nnzJ,
                                                 double t, const double* __restrict__ row_t,
                                                 int64_t i_off, int64_t j_off, apss_pair* out,
                                                 unsigned long long* out_cnt, int64_t out_cap) {
  const int64_t c = (int64_t)*band_cnt;
  const int64_t nband = c < band_cap ? c : band_cap;
  const bool vecI = ((reinterpret_cast<uintptr_t>(colI) | reinterpret_cast<uintptr_t>(valI)) & 15) == 0;
  const bool vecJ = ((reinterpret_cast<uintptr_t>(colJ) | reinterpret_cast<uintptr_t>(valJ)) & 15) == 0;
  // emitted pairs are staged in LDS and leave with one global reservation per
  // flush (a per-wave reservation on the one output counter serialised the
  // kernel); the loop bound is uniform per workgroup, so every thread reaches
  // the barriers below
  __shared__ apss_pair buf[RS_STAGE];
  __shared__ unsigned int n_buf, wcnt[4];
  __shared__ unsigned long long gbase;
  const int w = threadIdx.x >> 6;
  if (threadIdx.x == 0) n_buf = 0;
  __syncthreads();
  auto flush = [&]() {
    if (threadIdx.x == 0) gbase = n_buf ? atomicAdd(out_cnt, (unsigned long long)n_buf) : 0ull;
    __syncthreads();
    const unsigned int nb_ = n_buf;
    for (unsigned int q = threadIdx.x; q < nb_; q += blockDim.x) {
      const unsigned long long slot = gbase + q;
      if (slot < (unsigned long long)out_cap) out[slot] = buf[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) n_buf = 0;
    __syncthreads();
  };
  for (int64_t base = blockIdx.x * (int64_t)blockDim.x; base < nband;
       base += (int64_t)gridDim.x * blockDim.x) {
  const int64_t k = base + threadIdx.x;
  bool emit = false;
  uint32_t i = 0, j = 0;
  double s = 0.0;
  if (k < nband) {
    unsigned long long pr = band[k];
    i = (uint32_t)(pr >> 32);
    j = (uint32_t)(pr & 0xffffffffull);
    const int64_t x = rpI[i], y = rpJ[j];
    int na = (int)(rpI[i + 1] - x), nb = (int)(rpJ[j + 1] - y);  // entries left in each row
    if (na > 0 && nb > 0) {
      int64_t ba = x & ~(int64_t)3, bb = y & ~(int64_t)3;  // aligned window starts
      int ka = (int)(x & 3), kb = (int)(y & 3);              // merge positions in the windows
      // the current window of each row and the next one, loaded a window ahead
      RowWindow wa = row_window(colI, valI, ba, nnzI - ba, vecI);
      RowWindow wb = row_window(colJ, valJ, bb, nnzJ - bb, vecJ);
      RowWindow na_w = row_window(colI, valI, ba + 4, nnzI - ba - 4, vecI);
      RowWindow nb_w = row_window(colJ, valJ, bb + 4, nnzJ - bb - 4, vecJ);
      while (true) {
        const int32_t cx = pick_i(wa.c, ka), cy = pick_i(wb.c, kb);
        if (cx == cy) s += (double)pick_f(wa.v, ka) * (double)pick_f(wb.v, kb);
        if (cx <= cy) {
          if (--na == 0) break;
          if (++ka == 4) {
            ka = 0;
            ba += 4;
            wa = na_w;
            na_w = row_window(colI, valI, ba + 4, nnzI - ba - 4, vecI);
          }
        }
        if (cy <= cx) {
          if (--nb == 0) break;
          if (++kb == 4) {
            kb = 0;
            bb += 4;
            wb = nb_w;
            nb_w = row_window(colJ, valJ, bb + 4, nnzJ - bb - 4, vecJ);
          }
        }
      }
    }
    emit = s >= (row_t ? row_t[i] : t);
  }
  // this block's emitted pairs of the round, appended behind the buffer's
  // contents (a round adds at most blockDim.x <= 256 pairs; the buffer is
  // flushed before a round could overflow it)
  const unsigned long long m = __ballot(emit);
  if (lane_id() == 0) wcnt[w] = (unsigned int)__popcll(m);
  __syncthreads();
  unsigned int pos = n_buf + (unsigned int)lane_rank(m);
  for (int v = 0; v < w; ++v) pos += wcnt[v];
  if (emit) {
    apss_pair p;
    p.i = (uint32_t)(i + i_off);
    p.j = (uint32_t)(j + j_off);
    p.score = (float)s;
    buf[pos] = p;
  }
  __syncthreads();
  if (threadIdx.x == 0) n_buf += wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
  __syncthreads();
  if (n_buf > RS_STAGE - 256) flush();
  }
  if (n_buf > 0) flush();
}

// Settle bounds of every row of one matrix side for one launch.  E bounds the
// error of any score-tile cell the row takes part in, in fixed-point units
// (2^-S), whatever the other row: the cell bound u |x||y| + (k + 2) 2^-S
// (+ the dense-head and head-dot terms) with the other row's norms replaced by
// the other side's largest ones and k = min(len_i, len_j) by this row's own
// length; so the smaller of a cell's two row bounds bounds the cell.  Q is the
// least fixed-point score at which E meets the relative score tolerance
// (Q = E (rel_k + 1): q - E >= rel_k E means E <= (q - E) / rel_k), or
// INT32_MIN when emitted scores carry no relative tolerance (BF16 head).
__global__ __launch_bounds__(256) void k_row_bounds(int64_t n, const int64_t* __restrict__ rp,
                                                    const float* __restrict__ norm,
                                                    const float* __restrict__ hnorm,
                                                    RowBoundArgs b, int2* __restrict__ out) {
  const double u = 5.9604644775390625e-08;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * blockDim.x) {
    const int64_t len = rp[r + 1] - rp[r];
    double eps = u * (double)norm[r] * b.other_norm * (1.0 + 1e-6) + (double)(len + 2) * b.qinv + b.e_fixed;
    if (b.head) eps += gamma32(len + 1) * (double)hnorm[r] * b.other_norm * (1.0 + 1e-6);
    const double eq = ceil(eps * b.qscale * (1.0 + 1e-12)) + 1.0;
    const int32_t E = eq >= 1073741824.0 ? 1073741824 : (int32_t)eq;
    int32_t Q = INT32_MIN;
    if (b.rel_k > 0.0) {
      const double qq = (double)E * (b.rel_k + 1.0);
      Q = qq >= 2147483647.0 ? INT32_MAX : (int32_t)qq;
    }
    out[r] = make_int2(E, Q);
  }
}

}  // namespace apss
