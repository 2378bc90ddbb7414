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
// Here one workgroup owns one band: a tile I of 128 rows.  It sweeps its
// partner tiles J (J >= I for the self-join) and for each keeps the 128x128
// score tile in LDS, so the n^2 matrix never exists in HBM:
//
//   1. cursor tier: every short-list entry of band I keeps a cursor into its
//      (row-sorted) inverted list; the cursors that point into J advance and
//      add their products to the tile;
//   2. bitmap tier: AND of the two tiles' presence bitmaps gives the common
//      long-list features; each contributes the outer product of its two
//      tile segments (wave per feature when large, lane per feature when
//      small);
//   3. head (p > 0): the prefix-norm bound partial + |x_head||y_head| >= t
//      (rigorous, with the FP32 error bound) selects the cells whose
//      forward-list head dot is computed; others keep their partial score;
//   4. threshold: every valid cell is compared with t using its rigorous FP32
//      error bound; certain hits are appended to the output with one atomic
//      per wave (ballot + mbcnt ranks), cells within the bound go to the band
//      list that the FP64 rescore kernel settles exactly.
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
constexpr int OBUF_CAP = 512;               // LDS output buffer (pairs)
constexpr int I_PER_THREAD = CAP_I / SCAN_THREADS;
constexpr int CST_PAIRS = (KD_MAX + 2 + 2 * SCAN_THREADS - 1) / (2 * SCAN_THREADS);
constexpr int QUADS_PER_THREAD = (CAP_J / 4 + SCAN_THREADS - 1) / SCAN_THREADS;
static_assert(KD_MAX < 65536 && CAP_J < 65536, "16-bit segment starts");

struct Prefetch {
  uint64_t bmp;
  uint32_t wpre;
  uint32_t cst[CST_PAIRS];  // two 16-bit segment starts per register
  uint32_t rows[QUADS_PER_THREAD];
  float4 vals[QUADS_PER_THREAD];
  int32_t ncs, nd4, nrows;
};

__device__ __forceinline__ void prefetch_tile(const TileSide& S, int32_t t, int KW, int tid,
                                              Prefetch& pf) {
  const int64_t c0 = S.cbase[t];
  pf.ncs = (int)(S.cbase[t + 1] - c0);  // present features + 1
  const int64_t d0 = S.dbase[t];
  pf.nd4 = (int)((S.dbase[t + 1] - d0) >> 2);
  pf.nrows = (int)(S.tstart[t + 1] - S.tstart[t]);
  if (tid < KW) {
    pf.bmp = S.bmp[(int64_t)t * KW + tid];
    pf.wpre = S.wpre[(int64_t)t * KW + tid];
  }
#pragma unroll
  for (int i = 0; i < CST_PAIRS; ++i) {
    const int k = 2 * (tid + i * SCAN_THREADS);
    uint32_t lo = k < pf.ncs ? (uint32_t)S.cst[c0 + k] : 0u;
    uint32_t hi = k + 1 < pf.ncs ? (uint32_t)S.cst[c0 + k + 1] : 0u;
    pf.cst[i] = lo | (hi << 16);
  }
#pragma unroll
  for (int i = 0; i < QUADS_PER_THREAD; ++i) {
    const int k = tid + i * SCAN_THREADS;
    if (k < pf.nd4) {
      pf.rows[i] = reinterpret_cast<const uint32_t*>(S.drow + d0)[k];
      pf.vals[i] = reinterpret_cast<const float4*>(S.dval + d0)[k];
    }
  }
}

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
// Positions are read four at a time so one memory round trip covers up to
// four entries.
__device__ __forceinline__ void walk_count(const uint32_t* __restrict__ pos, int cur, int end,
                                           int32_t cs, int32_t ce, uint32_t* bkt) {
  for (;;) {
    uint32_t pt[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) pt[q] = cur + q < end ? pos[cur + q] : 0xffffffffu;
    bool go = true;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int32_t t = (int32_t)(pt[q] >> 8);
      go = go && t < ce;
      if (go) atomicAdd(&bkt[t - cs], 1u);
    }
    if (!go) return;
    cur += 4;
  }
}

__device__ __forceinline__ int walk_scatter(const uint32_t* __restrict__ pos,
                                            const int2* __restrict__ ent, int cur, int end,
                                            int32_t cs, int32_t ce, int rowc, float x,
                                            uint32_t* bfill, int2* __restrict__ cprod,
                                            unsigned int& n_prod) {
  for (;;) {
    uint32_t pt[4];
    int2 e[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const bool ok = cur + q < end;
      pt[q] = ok ? pos[cur + q] : 0xffffffffu;
      e[q] = ok ? ent[cur + q] : make_int2(0, 0);
    }
    bool go = true;
    int used = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int32_t t = (int32_t)(pt[q] >> 8);
      go = go && t < ce;
      if (go) {
        const uint32_t slot = atomicAdd(&bfill[t - cs], 1u);
        cprod[slot] = make_int2(rowc + (int)(pt[q] & 0xffu), (int)(x * __int_as_float(e[q].y)));
        ++used;
      }
    }
    cur += used;
    n_prod += (unsigned)used;
    if (!go) return cur;
  }
}

// Persistent band kernel: one workgroup per CU takes bands from a queue.
__global__ __launch_bounds__(SCAN_THREADS) void k_band_scan(ScanArgs a) {
  __shared__ int32_t tile[BAND_ROWS * TSTRIDE];     // fixed-point scores (x 2^S)
  __shared__ uint64_t bJ[KW_MAX];
  __shared__ uint16_t wpJ[KW_MAX];
  __shared__ uint16_t cJ[KD_MAX + 1];
  __shared__ __align__(16) int2 eJ[CAP_J];          // {local row, value bits}
  __shared__ uint32_t bkt[CH_TILES + 1];             // cursor-product bucket starts of the chunk
  __shared__ uint32_t bfill[CH_TILES];               // bucket fill pointers while scattering
  __shared__ uint32_t wsum[SCAN_THREADS / 64 + 1];
  __shared__ apss_pair obuf[OBUF_CAP];               // emitted pairs waiting for a flush
  __shared__ unsigned int obuf_n;
  __shared__ unsigned long long obuf_base;
  __shared__ float normI[BAND_ROWS], hnI[BAND_ROWS], normJ[TILE_ROWS], hnJ[TILE_ROWS];
  __shared__ int32_t lenI[BAND_ROWS], lenJ[TILE_ROWS];
  __shared__ float vI[CAP_I];                        // band entry values * 2^S, [q][tid]
  __shared__ int32_t band_s, J_first_s, ch_end_s;
  __shared__ unsigned long long acc_products, acc_head, acc_cells, acc_visits, acc_cursor;
  static_assert(CH_TILES == SCAN_THREADS, "one bucket per thread in the chunk scan");

  const int tid0 = threadIdx.x;
  const int KW = a.KW;
  int2* const cprod = a.cprod + (int64_t)blockIdx.x * a.cprod_cap;

  // the tile is cleared cell by cell by the threshold pass, so once is enough
  for (int c = tid0; c < BAND_ROWS * TSTRIDE; c += SCAN_THREADS) tile[c] = 0;
  if (tid0 == 0) {
    acc_products = 0; acc_head = 0; acc_cells = 0; acc_visits = 0; acc_cursor = 0; obuf_n = 0;
  }
  unsigned int my_products = 0, my_head = 0, my_cells = 0, my_visits = 0, my_cursor = 0;

  for (;;) {
  int tid = tid0;
  asm volatile("" : "+v"(tid));
  if (tid == 0) band_s = a.band_first + (int32_t)atomicAdd(a.band_next, 1ull);
  __syncthreads();
  const int32_t I = band_s;
  if (I >= a.band_first + a.band_count || I >= a.I.T) break;
  const int64_t i0 = a.I.tstart[I];
  const int nI = (int)(a.I.tstart[I + 1] - i0);
  // batched runs give each band its own partner range and threshold
  const int32_t J_end = a.band_jend ? a.band_jend[I] : a.J.T;
  const double t_band = a.band_t ? a.band_t[I] : a.t;
  int32_t q_lo = a.q_lo;
  if (a.band_t) {
    const double ql = floor((t_band - 2.0 * a.eps_max) * (double)a.qscale);
    q_lo = (int32_t)fmax(-2147483647.0, fmin(2147483647.0, ql));
  }

  // ---- band set-up: this thread's band entries live in registers ---------
  // Entry k = tid + q * SCAN_THREADS of the band block: (row * TSTRIDE,
  // bitmap index, value * 2^S).  The block is ordered by frequency rank, so
  // the lanes of a wave hold entries of the same or similar features.
  int32_t e_rf[I_PER_THREAD];  // row * TSTRIDE << 12 | bitmap index (or -1)
  {
    const int64_t d0 = a.I.dbase[I];
    const int nE = a.I.cst[a.I.cbase[I + 1] - 1];  // entries (end sentinel)
#pragma unroll
    for (int q = 0; q < I_PER_THREAD; ++q) {
      const int k = tid + q * SCAN_THREADS;
      const bool ok = k < nE;
      // bit 26 marks the first entry of a feature's band segment (it counts
      // the tile segment once in the list-visit statistics)
      const int32_t f = ok ? (int32_t)a.I.dfeat[d0 + k] : 0;
      const bool seg_head = ok && (k == 0 || (int32_t)a.I.dfeat[d0 + k - 1] != f);
      e_rf[q] = ok ? ((((int32_t)a.I.drow[d0 + k] * TSTRIDE) << 12) | f | (seg_head ? (1 << 26) : 0)) : -1;
      vI[k] = ok ? a.I.dval[d0 + k] * a.qscale : 0.f;  // exact: power of 2
    }
  }
  if (tid < BAND_ROWS) {
    bool ok = tid < nI;
    int64_t r = i0 + tid;
    normI[tid] = ok ? a.I.norm[r] : 0.f;
    hnI[tid] = ok ? a.I.hnorm[r] : 0.f;
    lenI[tid] = ok ? (int32_t)(a.I.rp[r + 1] - a.I.rp[r]) : 0;
  }
  if (tid == 0) {
    int32_t jf = 0;
    if (a.self) {  // first tile whose rows end after the band's first row
      int32_t lo = 0, hi = a.J.T;
      while (lo < hi) {
        int32_t m = (lo + hi) >> 1;
        if (a.J.tstart[m + 1] <= i0) lo = m + 1; else hi = m;
      }
      jf = lo;
    }
    J_first_s = jf;
  }
  // cursor records: position and the J tile it points into
  const int64_t sb = a.I.sbase[I];
  const int nS = (int)(a.I.sbase[I + 1] - sb);
  for (int k = tid; k < nS; k += SCAN_THREADS) {
    const int4 r = a.I.srec[sb + k];
    a.scur[sb + k] = r.z;
    a.snext[sb + k] = r.z < r.w ? (int32_t)(a.S.pos[r.z] >> 8) : INT32_MAX;
  }
  __syncthreads();
  const int32_t J_first = J_first_s;

  Prefetch pf;
  if (J_first < J_end) prefetch_tile(a.J, J_first, KW, tid, pf);
  int32_t ch_lo = J_first, ch_hi = nS > 0 ? J_first : J_end;  // bucket chunk [ch_lo, ch_hi)

  for (int32_t J = J_first; J < J_end; ++J) {
    // per-thread addresses are recomputed from an opaque thread index in every
    // pass instead of being kept live across the loop (register budget)
    int tl = tid;
    asm volatile("" : "+v"(tl));

    // ---- 1. cursor tier: bucket the products of the next chunk of tiles ----
    // Every record walks its row-sorted list up to the chunk end twice: once
    // to count the products per J tile, once to write them (fixed point, with
    // their cell) into this workgroup's region, grouped by tile.  A chunk
    // whose products exceed the region ends early at a tile boundary.
    if (J == ch_hi) {
      ch_lo = J;
      const int32_t ce0 = J + CH_TILES < J_end ? J + CH_TILES : J_end;
      bkt[tl] = 0;
      if (tl == 0) { bkt[CH_TILES] = 0; ch_end_s = ce0; }
      __syncthreads();
      for (int k = tl; k < nS; k += SCAN_THREADS) {
        if (a.snext[sb + k] >= ce0) continue;
        walk_count(a.S.pos, a.scur[sb + k], a.I.srec[sb + k].w, ch_lo, ce0, bkt);
      }
      __syncthreads();
      const uint32_t cnt = bkt[tl];
      uint32_t total;
      const uint32_t ex = block_exscan(cnt, tl, wsum, &total);
      bkt[tl] = ex;
      bfill[tl] = ex;
      if (tl == 0) bkt[CH_TILES] = total;
      if ((int64_t)ex + cnt > a.cprod_cap) atomicMin(&ch_end_s, ch_lo + tl);
      __syncthreads();
      const int32_t ce = ch_end_s;
      for (int k = tl; k < nS; k += SCAN_THREADS) {
        if (a.snext[sb + k] >= ce) continue;
        const int4 r = a.I.srec[sb + k];
        const int cur = walk_scatter(a.S.pos, a.S.ent, a.scur[sb + k], r.w, ch_lo, ce,
                                     r.x * TSTRIDE, __int_as_float(r.y) * a.qscale, bfill, cprod,
                                     my_cursor);
        a.scur[sb + k] = cur;
        a.snext[sb + k] = cur < r.w ? (int32_t)(a.S.pos[cur] >> 8) : INT32_MAX;
      }
      __syncthreads();
      ch_hi = ce;
    }

    const int64_t j0 = a.J.tstart[J];
    const int nJ = pf.nrows;

    // ---- write the prefetched tile J into LDS --------------------------------
    if (tl < KW) { bJ[tl] = pf.bmp; wpJ[tl] = pf.wpre; }
#pragma unroll
    for (int i = 0; i < CST_PAIRS; ++i) {
      const int k = 2 * (tl + i * SCAN_THREADS);
      if (k < pf.ncs) cJ[k] = (uint16_t)(pf.cst[i] & 0xffffu);
      if (k + 1 < pf.ncs) cJ[k + 1] = (uint16_t)(pf.cst[i] >> 16);
    }
#pragma unroll
    for (int i = 0; i < QUADS_PER_THREAD; ++i) {
      const int k = tl + i * SCAN_THREADS;
      if (k < pf.nd4) {
        const float vv[4] = {pf.vals[i].x, pf.vals[i].y, pf.vals[i].z, pf.vals[i].w};
#pragma unroll
        for (int q = 0; q < 4; ++q)
          eJ[4 * k + q] = make_int2((int)((pf.rows[i] >> (8 * q)) & 0xffu), __float_as_int(vv[q]));
      }
    }
    if (tl < TILE_ROWS) {
      const bool ok = tl < nJ;
      const int64_t r = j0 + tl;
      normJ[tl] = ok ? a.J.norm[r] : 0.f;
      hnJ[tl] = ok ? a.J.hnorm[r] : 0.f;
      lenJ[tl] = ok ? (int32_t)(a.J.rp[r + 1] - a.J.rp[r]) : 0;
    }

    __syncthreads();

    // prefetch tile J+1 (registers; written to LDS at the top of the next pass)
    if (J + 1 < J_end) prefetch_tile(a.J, J + 1, KW, tl, pf);

    // this thread's first cursor product of tile J: its load completes behind
    // the bitmap-tier products
    int32_t pb = 0, pe = 0;
    if (nS > 0) { pb = (int32_t)bkt[J - ch_lo]; pe = (int32_t)bkt[J - ch_lo + 1]; }
    const int32_t k0 = pb + tl;
    int2 hp = make_int2(0, 0);
    if (k0 < pe) hp = cprod[k0];

    // ---- 2. bitmap tier: every band entry meets its feature's J segment -----
    // Two passes of I_PER_THREAD/2 entries: the LDS lookups of a pass are
    // issued together, then its segments are walked.
#pragma unroll
    for (int h = 0; h < I_PER_THREAD; h += I_PER_THREAD / 2) {
      int y0[I_PER_THREAD / 2], y1[I_PER_THREAD / 2];
#pragma unroll
      for (int q = 0; q < I_PER_THREAD / 2; ++q) {
        int rf = e_rf[h + q];
        asm volatile("" : "+v"(rf));  // no hoisting of per-entry values out of the J loop
        y0[q] = 0; y1[q] = 0;
        if (rf >= 0) {
          const int f = rf & 0xfff;  // bitmap index (bits 0..11)
          const int w = f >> 6, bit = f & 63;
          const uint64_t wj = bJ[w];
          if ((wj >> bit) & 1ull) {
            const int rk = wpJ[w] + __popcll(wj & ((1ull << bit) - 1ull));
            y0[q] = cJ[rk];
            y1[q] = cJ[rk + 1];
          }
        }
      }
#pragma unroll
      for (int q = 0; q < I_PER_THREAD / 2; ++q) {
        if (y1[q] <= y0[q]) continue;
        my_visits += 1u + (((e_rf[h + q] >> 26) & 1) ? (unsigned)(y1[q] - y0[q]) : 0u);
        const int ib = (e_rf[h + q] >> 12) & 0x3fff;
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
        my_products += (unsigned)(y1[q] - y0[q]);
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
    __syncthreads();
    // ---- 3. head: prefix-norm filter + forward-list dot -------------------
    if (a.head_mode) {
#pragma unroll 1
      for (int c = tl; c < BAND_ROWS * TILE_ROWS; c += SCAN_THREADS) {
        const int il = c / TILE_ROWS, jl = c % TILE_ROWS;
        if (il >= nI || jl >= nJ || (a.self && i0 + il >= j0 + jl)) continue;
        int32_t* cell = &tile[il * TSTRIDE + jl];
        const int32_t part = *cell;
        // pruned when even the Cauchy-Schwarz head bound cannot lift it to t
        const double bound = (double)part * a.qinv + (double)hnI[il] * (double)hnJ[jl];
        if (bound + cell_error_bound(lenI[il] < lenJ[jl] ? lenI[il] : lenJ[jl], normI[il], normJ[jl],
                                     a.qinv, 0.0) < t_band)
          continue;
        const int64_t ri = i0 + il, rj = j0 + jl;
        const int64_t ai = a.I.rp[ri], aj = a.J.rp[rj];
        const float h = head_dot(a.I.col + ai, a.I.val + ai, a.I.rp[ri + 1] - ai, a.J.col + aj,
                                 a.J.val + aj, a.J.rp[rj + 1] - aj, a.p);
        *cell = part + __float2int_rn(h * a.qscale);
        ++my_head;
      }
      __syncthreads();
    }

    // ---- 4. threshold, compaction, reset -------------------------------------
    // Pass 1 reads this thread's cells together and marks the ones that can
    // reach t; pass 2 (rolled) settles candidates and clears every cell.
    // cell c = tl + m*SCAN_THREADS: column tl % TILE_ROWS is fixed per thread,
    // row advances by SCAN_THREADS / TILE_ROWS; a cell is valid when its row
    // is below both the band height and (self-joins) the diagonal limit
    constexpr int CELLS_PER_THREAD = BAND_ROWS * TILE_ROWS / SCAN_THREADS;
    constexpr int ROW_STEP = SCAN_THREADS / TILE_ROWS;
    static_assert(SCAN_THREADS % TILE_ROWS == 0, "fixed column per thread");
    const int jl_c = tl % TILE_ROWS, il_0 = tl / TILE_ROWS;
    int row_lim = jl_c < nJ ? nI : 0;
    if (a.self) {  // i0 + il < j0 + jl  <=>  il < (j0 - i0) + jl
      const int64_t dlim = j0 - i0 + jl_c;
      row_lim = (int)(dlim < row_lim ? (dlim < 0 ? 0 : dlim) : row_lim);
    }
    unsigned cand = 0;
#pragma unroll
    for (int m = 0; m < CELLS_PER_THREAD; ++m) {
      const int il = il_0 + m * ROW_STEP;
      const int32_t q = tile[il * TSTRIDE + jl_c];
      const bool valid = il < row_lim;
      my_cells += valid ? 1u : 0u;
      cand |= (valid && q >= q_lo) ? (1u << m) : 0u;
    }
#pragma unroll 1
    for (int m = 0; m < CELLS_PER_THREAD; ++m) {
      const int c = tl + m * SCAN_THREADS;
      const int il = c / TILE_ROWS, jl = c % TILE_ROWS;
      int32_t* cell = &tile[il * TSTRIDE + jl];
      const bool is_cand = (cand >> m) & 1u;
      if (!__ballot(is_cand)) { *cell = 0; continue; }
      const int32_t q = *cell;
      *cell = 0;
      bool emit = false, to_band = false;
      double sd = 0.0;
      if (is_cand) {
        const int k = lenI[il] < lenJ[jl] ? lenI[il] : lenJ[jl];
        const double hb = a.head_mode ? gamma32(k + 1) * (double)hnI[il] * (double)hnJ[jl] : 0.0;
        const double eps = cell_error_bound(k, normI[il], normJ[jl], a.qinv, hb);
        sd = (double)q * a.qinv;
        if (sd - eps >= t_band) {
          // certain hit; its score is final only if the bound meets the relative tolerance
          if (a.rescore_all || eps > a.score_rel * (sd - eps)) to_band = true; else emit = true;
        } else if (sd + eps >= t_band) {
          to_band = true;
        }
      }
      // emitted pairs go to the workgroup's LDS buffer (one LDS atomic per
      // wave); a wave that finds the buffer full reserves in HBM directly
      const unsigned long long me = __ballot(emit);
      if (me) {
        const int tot = __popcll(me);
        const int leader = __ffsll((long long)me) - 1;
        unsigned int lb = 0;
        if (lane_id() == leader) lb = atomicAdd(&obuf_n, (unsigned)tot);
        lb = __shfl(lb, leader);
        apss_pair pr;
        pr.i = (uint32_t)(i0 + il + a.i_off);
        pr.j = (uint32_t)(j0 + jl + a.j_off);
        pr.score = (float)sd;
        // every buffer slot below OBUF_CAP is written by the wave that reserved
        // it; the part of a reservation beyond the buffer goes straight to HBM
        const unsigned int pos = lb + (unsigned)lane_rank(me);
        if (emit && pos < OBUF_CAP) obuf[pos] = pr;
        if (lb + tot > OBUF_CAP) {
          int t2;
          const bool spill = emit && pos >= OBUF_CAP;
          long long slot = wave_reserve(a.out_cnt, spill, &t2);
          if (spill && slot < a.out_cap) a.out[slot] = pr;
        }
      }
      int tot;
      long long bslot = wave_reserve(a.band_cnt, to_band, &tot);
      if (to_band && bslot < a.band_cap)
        a.band[bslot] = ((unsigned long long)(i0 + il) << 32) | (unsigned long long)(j0 + jl);
    }
    __syncthreads();
    // flush the LDS output buffer when half full, and after the last tile
    const unsigned int nbuf = obuf_n;
    if (nbuf >= OBUF_CAP / 2 || (J + 1 == J_end && nbuf > 0)) {
      const unsigned int nf = nbuf < OBUF_CAP ? nbuf : OBUF_CAP;
      if (tl == 0) obuf_base = atomicAdd(a.out_cnt, (unsigned long long)nf);
      __syncthreads();
      const unsigned long long ob = obuf_base;
      for (unsigned int k = tl; k < nf; k += SCAN_THREADS)
        if (ob + k < (unsigned long long)a.out_cap) a.out[ob + k] = obuf[k];
      __syncthreads();
      if (tl == 0) obuf_n = 0;
      __syncthreads();
    }
  }
  __syncthreads();
  }  // band queue

  atomicAdd(&acc_products, (unsigned long long)my_products + my_cursor);
  atomicAdd(&acc_head, (unsigned long long)my_head);
  atomicAdd(&acc_cells, (unsigned long long)my_cells);
  atomicAdd(&acc_visits, (unsigned long long)my_visits);
  atomicAdd(&acc_cursor, (unsigned long long)my_cursor);
  __syncthreads();
  if (tid0 == 0) {
    atomicAdd(&a.stats[0], acc_products);
    atomicAdd(&a.stats[1], acc_head);
    atomicAdd(&a.stats[2], acc_cells);
    atomicAdd(&a.stats[3], acc_visits);
    atomicAdd(&a.stats[4], acc_cursor);
  }
}

// FP64 rescore of the band: exact products summed in ascending feature id,
// the same order as the FP64 oracle, so the decision s >= t is settled exactly.
__global__ __launch_bounds__(256) void k_rescore(const unsigned long long* __restrict__ band,
                                                 int64_t nband, const int64_t* __restrict__ rpI,
                                                 const int32_t* __restrict__ colI,
                                                 const float* __restrict__ valI,
                                                 const int64_t* __restrict__ rpJ,
                                                 const int32_t* __restrict__ colJ,
                                                 const float* __restrict__ valJ, double t,
                                                 int64_t i_off, int64_t j_off, apss_pair* out,
                                                 unsigned long long* out_cnt, int64_t out_cap) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  bool emit = false;
  uint32_t i = 0, j = 0;
  double s = 0.0;
  if (k < nband) {
    unsigned long long pr = band[k];
    i = (uint32_t)(pr >> 32);
    j = (uint32_t)(pr & 0xffffffffull);
    int64_t x = rpI[i], xe = rpI[i + 1], y = rpJ[j], ye = rpJ[j + 1];
    while (x < xe && y < ye) {
      int32_t cx = colI[x], cy = colJ[y];
      if (cx < cy) ++x;
      else if (cx > cy) ++y;
      else { s += (double)valI[x] * (double)valJ[y]; ++x; ++y; }
    }
    emit = s >= t;
  }
  int tot;
  long long slot = wave_reserve(out_cnt, emit, &tot);
  if (emit && slot < out_cap) {
    apss_pair p;
    p.i = (uint32_t)(i + i_off);
    p.j = (uint32_t)(j + j_off);
    p.score = (float)s;
    out[slot] = p;
  }
}

}  // namespace apss
