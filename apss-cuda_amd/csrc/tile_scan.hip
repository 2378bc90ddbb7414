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

// Copy one tile block (bitmap words, word prefixes, compact starts, entries)
// from HBM into LDS.  Block sizes are padded to multiples of 4 entries so the
// entries move as 32-bit / 128-bit words.
__device__ __forceinline__ void stage_block(const TileSide& S, int32_t t, int KW, uint64_t* bmp,
                                            uint16_t* wpre, uint16_t* cst, uint8_t* rows,
                                            float* vals, int tid) {
  for (int w = tid; w < KW; w += SCAN_THREADS) {
    bmp[w] = S.bmp[(int64_t)t * KW + w];
    wpre[w] = S.wpre[(int64_t)t * KW + w];
  }
  const int64_t c0 = S.cbase[t];
  const int ncs = (int)(S.cbase[t + 1] - c0);  // present features + 1
  for (int k = tid; k < ncs; k += SCAN_THREADS) cst[k] = (uint16_t)S.cst[c0 + k];
  const int64_t d0 = S.dbase[t];
  const int nd4 = (int)((S.dbase[t + 1] - d0) >> 2);
  const uint32_t* r4 = reinterpret_cast<const uint32_t*>(S.drow + d0);
  const float4* v4 = reinterpret_cast<const float4*>(S.dval + d0);
  for (int k = tid; k < nd4; k += SCAN_THREADS) {
    reinterpret_cast<uint32_t*>(rows)[k] = r4[k];
    reinterpret_cast<float4*>(vals)[k] = v4[k];
  }
}

__global__ __launch_bounds__(SCAN_THREADS) void k_band_scan(ScanArgs a) {
  __shared__ float tile[BAND_ROWS * TSTRIDE];
  __shared__ uint64_t bI[KW_MAX], bJ[KW_MAX];
  __shared__ uint16_t wpI[KW_MAX], wpJ[KW_MAX];
  __shared__ uint16_t cI[KD_MAX + 1], cJ[KD_MAX + 1];
  __shared__ __align__(16) uint8_t rI[CAP_I];
  __shared__ __align__(16) uint8_t rJ[CAP_J];
  __shared__ __align__(16) float vI[CAP_I];
  __shared__ __align__(16) float vJ[CAP_J];
  __shared__ int32_t s_nxt[CAP_SI];
  __shared__ float normI[BAND_ROWS], hnI[BAND_ROWS], normJ[TILE_ROWS], hnJ[TILE_ROWS];
  __shared__ int32_t lenI[BAND_ROWS], lenJ[TILE_ROWS];
  __shared__ int32_t J_first_s;
  __shared__ unsigned long long acc_products, acc_head, acc_cells;

  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int wave = tid / 64;
  constexpr int NW = SCAN_THREADS / 64;
  const int KW = a.KW;

  const int32_t I = a.band_first + blockIdx.x;
  if (I >= a.I.T) return;
  const int64_t i0 = a.I.tstart[I];
  const int nI = (int)(a.I.tstart[I + 1] - i0);

  // ---- band set-up -------------------------------------------------------
  stage_block(a.I, I, KW, bI, wpI, cI, rI, vI, tid);
  for (int c = tid; c < BAND_ROWS * TSTRIDE; c += SCAN_THREADS) tile[c] = 0.f;
  if (tid < BAND_ROWS) {
    bool ok = tid < nI;
    int64_t r = i0 + tid;
    normI[tid] = ok ? a.I.norm[r] : 0.f;
    hnI[tid] = ok ? a.I.hnorm[r] : 0.f;
    lenI[tid] = ok ? (int32_t)(a.I.rp[r + 1] - a.I.rp[r]) : 0;
  }
  if (tid == 0) {
    acc_products = 0; acc_head = 0; acc_cells = 0;
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
  const int64_t sb = a.I.sbase[I];
  const int nS = (int)(a.I.sbase[I + 1] - sb);
  for (int k = tid; k < nS; k += SCAN_THREADS) {
    int4 r = a.I.srec[sb + k];
    int32_t nx = r.z < r.w ? a.S.ent[r.z].x : INT32_MAX;
    a.scur[sb + k] = r.z;
    if (k < CAP_SI) s_nxt[k] = nx; else a.snext_spill[sb + k] = nx;
  }
  unsigned long long my_products = 0, my_head = 0, my_cells = 0;
  __syncthreads();
  const int32_t J_first = J_first_s;

  for (int32_t J = J_first; J < a.J.T; ++J) {
    const int64_t j0 = a.J.tstart[J];
    const int nJ = (int)(a.J.tstart[J + 1] - j0);
    const int64_t j1 = j0 + nJ;

    // ---- stage tile J ------------------------------------------------------
    stage_block(a.J, J, KW, bJ, wpJ, cJ, rJ, vJ, tid);
    if (tid < TILE_ROWS) {
      bool ok = tid < nJ;
      int64_t r = j0 + tid;
      normJ[tid] = ok ? a.J.norm[r] : 0.f;
      hnJ[tid] = ok ? a.J.hnorm[r] : 0.f;
      lenJ[tid] = ok ? (int32_t)(a.J.rp[r + 1] - a.J.rp[r]) : 0;
    }

    // ---- 1. cursor tier (score tile only, no staging needed) ---------------
    for (int k = tid; k < nS; k += SCAN_THREADS) {
      int32_t nx = k < CAP_SI ? s_nxt[k] : a.snext_spill[sb + k];
      if (nx >= j1) continue;
      const int4 r = a.I.srec[sb + k];
      int32_t cur = a.scur[sb + k];
      const float x = __int_as_float(r.y);
      const int rowbase = r.x * TSTRIDE - (int)j0;
      do {
        const int2 e = a.S.ent[cur];
        atomicAdd(&tile[rowbase + nx], x * __int_as_float(e.y));
        ++my_products;
        ++cur;
        nx = cur < r.w ? a.S.ent[cur].x : INT32_MAX;
      } while (nx < j1);
      a.scur[sb + k] = cur;
      if (k < CAP_SI) s_nxt[k] = nx; else a.snext_spill[sb + k] = nx;
    }
    __syncthreads();

    // ---- 2. bitmap tier: one bitmap word per wave and pass -------------------
    for (int w = wave; w < KW; w += NW) {
      const uint64_t wI = bI[w], wJ = bJ[w];
      const uint64_t m = wI & wJ;
      if (!m) continue;
      const bool has = (m >> lane) & 1ull;
      int aI = 0, ca = 0, aJ = 0, cb = 0;
      if (has) {
        const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
        const int rkI = wpI[w] + __popcll(wI & below);
        const int rkJ = wpJ[w] + __popcll(wJ & below);
        aI = cI[rkI];
        ca = cI[rkI + 1] - aI;
        aJ = cJ[rkJ];
        cb = cJ[rkJ + 1] - aJ;
      }
      const bool big = has && ca * cb >= BIG_PRODUCTS;
      if (has && !big) {  // small segments: one lane per feature
        for (int x = 0; x < ca; ++x) {
          const int rowbase = rI[aI + x] * TSTRIDE;
          const float u = vI[aI + x];
          for (int y = 0; y < cb; ++y) atomicAdd(&tile[rowbase + rJ[aJ + y]], u * vJ[aJ + y]);
        }
        my_products += (unsigned long long)(ca * cb);
      }
      uint64_t mb = __ballot(big);
      while (mb) {  // large segments: the wave walks the flattened products
        const int b = __ffsll((long long)mb) - 1;
        mb &= mb - 1ull;
        const int fa = __shfl(aI, b), fca = __shfl(ca, b);
        const int fb = __shfl(aJ, b), fcb = __shfl(cb, b);
        const int P = fca * fcb;
        const int dx = 64 / fcb, dy = 64 - dx * fcb;
        int x = lane / fcb, y = lane - (lane / fcb) * fcb;
        for (int q = lane; q < P; q += 64) {
          atomicAdd(&tile[rI[fa + x] * TSTRIDE + rJ[fb + y]], vI[fa + x] * vJ[fb + y]);
          x += dx;
          y += dy;
          if (y >= fcb) { y -= fcb; ++x; }
        }
        if (lane == 0) my_products += (unsigned long long)P;
      }
    }
    __syncthreads();

    // ---- 3. head: prefix-norm filter + forward-list dot -------------------
    if (a.head_mode) {
      for (int c = tid; c < BAND_ROWS * TILE_ROWS; c += SCAN_THREADS) {
        const int il = c / TILE_ROWS, jl = c % TILE_ROWS;
        if (il >= nI || jl >= nJ || (a.self && i0 + il >= j0 + jl)) continue;
        float* cell = &tile[il * TSTRIDE + jl];
        const float part = *cell;
        if (part + hnI[il] * hnJ[jl] < a.t_lo) continue;  // pruned: head cannot reach t
        const int64_t ri = i0 + il, rj = j0 + jl;
        const int64_t ai = a.I.rp[ri], aj = a.J.rp[rj];
        *cell = part + head_dot(a.I.col + ai, a.I.val + ai, a.I.rp[ri + 1] - ai, a.J.col + aj,
                                a.J.val + aj, a.J.rp[rj + 1] - aj, a.p);
        ++my_head;
      }
      __syncthreads();
    }

    // ---- 4. threshold, compaction, reset -------------------------------------
    for (int c = tid; c < BAND_ROWS * TILE_ROWS; c += SCAN_THREADS) {
      const int il = c / TILE_ROWS, jl = c % TILE_ROWS;
      float* cell = &tile[il * TSTRIDE + jl];
      const float s = *cell;
      *cell = 0.f;
      const bool valid = il < nI && jl < nJ && !(a.self && i0 + il >= j0 + jl);
      my_cells += valid ? 1 : 0;
      bool emit = false, to_band = false;
      if (valid && s >= a.t_lo) {
        const int k = (lenI[il] < lenJ[jl] ? lenI[il] : lenJ[jl]) + 2;
        const double eps = gamma32(k) * (double)normI[il] * (double)normJ[jl] * (1.0 + 1e-6) +
                           (double)k * 1.401298464324817e-45;
        const double sd = (double)s;
        if (sd - eps >= a.t) {
          if (a.rescore_all) to_band = true; else emit = true;
        } else if (sd + eps >= a.t) {
          to_band = true;
        }
      }
      int tot;
      long long slot = wave_reserve(a.out_cnt, emit, &tot);
      if (emit && slot < a.out_cap) {
        apss_pair pr;
        pr.i = (uint32_t)(i0 + il + a.i_off);
        pr.j = (uint32_t)(j0 + jl + a.j_off);
        pr.score = s;
        a.out[slot] = pr;
      }
      long long bslot = wave_reserve(a.band_cnt, to_band, &tot);
      if (to_band && bslot < a.band_cap)
        a.band[bslot] = ((unsigned long long)(i0 + il) << 32) | (unsigned long long)(j0 + jl);
    }
    __syncthreads();
  }

  atomicAdd(&acc_products, my_products);
  atomicAdd(&acc_head, my_head);
  atomicAdd(&acc_cells, my_cells);
  __syncthreads();
  if (tid == 0) {
    atomicAdd(&a.stats[0], acc_products);
    atomicAdd(&a.stats[1], acc_head);
    atomicAdd(&a.stats[2], acc_cells);
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
