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

constexpr int GROUP_FEATURES = 1024;     // bitmap features intersected per pass
constexpr int N_GROUPS = KD_MAX / GROUP_FEATURES;
constexpr int LIST_CAP = GROUP_FEATURES; // records per list and pass
constexpr int S_CAP = 4096;              // cursor heads kept in LDS

struct Rec {  // one common bitmap-tier feature of a tile pair
  int32_t a;  // first I entry
  int32_t b;  // first J entry
  int32_t ab; // ca | cb << 16
};

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

__global__ __launch_bounds__(SCAN_THREADS) void k_band_scan(ScanArgs a) {
  __shared__ float tile[TILE * TSTRIDE];
  __shared__ uint64_t bmpI[KW_MAX];
  __shared__ Rec big[LIST_CAP];
  __shared__ Rec small[LIST_CAP];
  __shared__ int32_t s_cur[S_CAP];
  __shared__ int32_t s_nxt[S_CAP];
  __shared__ float normI[TILE], normJ[TILE], hnI[TILE], hnJ[TILE];
  __shared__ int32_t lenI[TILE], lenJ[TILE];
  __shared__ unsigned int n_big[N_GROUPS], n_small[N_GROUPS];  // per pass, zeroed per tile pair
  __shared__ unsigned long long acc_products, acc_head, acc_cells;

  const int tid = threadIdx.x;
  const int lane = lane_id();
  const int wave = tid / 64;
  constexpr int NW = SCAN_THREADS / 64;

  const int32_t I = a.band_first + blockIdx.x;
  if (I >= a.I.T) return;
  const int64_t i0 = (int64_t)I * TILE;
  const int nI = (int)((a.I.n - i0) < TILE ? (a.I.n - i0) : TILE);

  // ---- band set-up -------------------------------------------------------
  for (int c = tid; c < TILE * TSTRIDE; c += SCAN_THREADS) tile[c] = 0.f;
  for (int w = tid; w < a.KW; w += SCAN_THREADS) bmpI[w] = a.I.bmp[(int64_t)I * a.KW + w];
  if (tid < TILE) {
    bool ok = tid < nI;
    int64_t r = i0 + tid;
    normI[tid] = ok ? a.I.norm[r] : 0.f;
    hnI[tid] = ok ? a.I.hnorm[r] : 0.f;
    lenI[tid] = ok ? (int32_t)(a.I.rp[r + 1] - a.I.rp[r]) : 0;
  }
  if (tid == 0) { acc_products = 0; acc_head = 0; acc_cells = 0; }
  if (tid < N_GROUPS) { n_big[tid] = 0; n_small[tid] = 0; }
  const int64_t sb = a.I.sbase[I];
  const int nS = (int)(a.I.sbase[I + 1] - sb);
  for (int k = tid; k < nS; k += SCAN_THREADS) {
    int4 r = a.I.srec[sb + k];
    int32_t nx = r.z < r.w ? a.S.row[r.z] : INT32_MAX;
    if (k < S_CAP) { s_cur[k] = r.z; s_nxt[k] = nx; }
    else { a.snext_spill[2 * (sb + k)] = r.z; a.snext_spill[2 * (sb + k) + 1] = nx; }
  }
  const uint64_t* bmpJall = a.J.bmp;
  const int32_t* ptrIrow = a.I.ptr + (int64_t)I * (a.KD + 1);
  const int64_t baseI = a.I.base[I];
  unsigned long long my_products = 0, my_head = 0, my_cells = 0;
  __syncthreads();

  const int32_t J_first = a.self ? I : 0;
  for (int32_t J = J_first; J < a.J.T; ++J) {
    const int64_t j0 = (int64_t)J * TILE;
    const int nJ = (int)((a.J.n - j0) < TILE ? (a.J.n - j0) : TILE);
    const int64_t j1 = j0 + nJ;
    const bool diag = a.self && (J == I);
    if (tid < TILE) {
      bool ok = tid < nJ;
      int64_t r = j0 + tid;
      normJ[tid] = ok ? a.J.norm[r] : 0.f;
      hnJ[tid] = ok ? a.J.hnorm[r] : 0.f;
      lenJ[tid] = ok ? (int32_t)(a.J.rp[r + 1] - a.J.rp[r]) : 0;
    }

    // ---- 1. cursor tier --------------------------------------------------
    for (int k = tid; k < nS; k += SCAN_THREADS) {
      int32_t nx, cur;
      if (k < S_CAP) { nx = s_nxt[k]; }
      else { nx = a.snext_spill[2 * (sb + k) + 1]; }
      if (nx >= j1) continue;
      cur = k < S_CAP ? s_cur[k] : a.snext_spill[2 * (sb + k)];
      int4 r = a.I.srec[sb + k];
      float x = __int_as_float(r.y);
      const int rowbase = r.x * TSTRIDE;
      do {
        atomicAdd(&tile[rowbase + (int)(nx - j0)], x * a.S.val[cur]);
        ++my_products;
        ++cur;
        nx = cur < r.w ? a.S.row[cur] : INT32_MAX;
      } while (nx < j1);
      if (k < S_CAP) { s_cur[k] = cur; s_nxt[k] = nx; }
      else { a.snext_spill[2 * (sb + k)] = cur; a.snext_spill[2 * (sb + k) + 1] = nx; }
    }

    // ---- 2. bitmap tier ----------------------------------------------------
    const uint64_t* bmpJ = bmpJall + (int64_t)J * a.KW;
    const int32_t* ptrJrow = a.J.ptr + (int64_t)J * (a.KD + 1);
    const int64_t baseJ = a.J.base[J];
    for (int32_t g0 = 0; g0 < a.KD; g0 += GROUP_FEATURES) {
      const int g = g0 / GROUP_FEATURES;
      __syncthreads();  // previous pass done with big[] / small[]
      {
        const int32_t w = g0 / 64 + wave;  // one bitmap word per wave
        if (w < a.KW) {
          uint64_t m = bmpI[w] & bmpJ[w];
          if (m) {
            bool has = (m >> lane) & 1ull;
            Rec rec = {0, 0, 0};
            bool is_big = false;
            if (has) {
              int32_t f = w * 64 + lane;
              int32_t pa = ptrIrow[f], ca = ptrIrow[f + 1] - pa;
              int32_t pb = ptrJrow[f], cb = ptrJrow[f + 1] - pb;
              rec.a = (int32_t)(baseI + pa);
              rec.b = (int32_t)(baseJ + pb);
              rec.ab = ca | (cb << 16);
              is_big = ca * cb >= BIG_PRODUCTS;
            }
            unsigned long long mb = __ballot(has && is_big);
            unsigned long long ms = __ballot(has && !is_big);
            unsigned int bb = 0, bs = 0;
            if (lane == 0) {
              if (mb) bb = atomicAdd(&n_big[g], (unsigned)__popcll(mb));
              if (ms) bs = atomicAdd(&n_small[g], (unsigned)__popcll(ms));
            }
            bb = __shfl(bb, 0);
            bs = __shfl(bs, 0);
            if (has && is_big) big[bb + lane_rank(mb)] = rec;
            if (has && !is_big) small[bs + lane_rank(ms)] = rec;
          }
        }
      }
      __syncthreads();
      const int nb = (int)n_big[g], ns = (int)n_small[g];
      // large segments: one wave per feature, lanes over the flattened products
      for (int r = wave; r < nb; r += NW) {
        Rec rec = big[r];
        const int ca = rec.ab & 0xffff, cb = rec.ab >> 16;
        const int P = ca * cb;
        const int dx = 64 / cb, dy = 64 - dx * cb;
        int x = lane / cb, y = lane - x * cb;
        for (int q = lane; q < P; q += 64) {
          int il = a.I.drow[rec.a + x];
          float u = a.I.dval[rec.a + x];
          int jl = a.J.drow[rec.b + y];
          float v = a.J.dval[rec.b + y];
          atomicAdd(&tile[il * TSTRIDE + jl], u * v);
          x += dx;
          y += dy;
          if (y >= cb) { y -= cb; ++x; }
        }
        if (lane == 0) my_products += (unsigned long long)P;
      }
      // small segments: one lane per feature
      for (int r = tid; r < ns; r += SCAN_THREADS) {
        Rec rec = small[r];
        const int ca = rec.ab & 0xffff, cb = rec.ab >> 16;
        for (int x = 0; x < ca; ++x) {
          int il = a.I.drow[rec.a + x];
          float u = a.I.dval[rec.a + x];
          float* trow = tile + il * TSTRIDE;
          for (int y = 0; y < cb; ++y) atomicAdd(trow + a.J.drow[rec.b + y], u * a.J.dval[rec.b + y]);
        }
        my_products += (unsigned long long)(ca * cb);
      }
    }
    __syncthreads();
    // every pass of this tile pair has read its counters: zero them for the next
    if (tid < N_GROUPS) { n_big[tid] = 0; n_small[tid] = 0; }

    // ---- 3. head: prefix-norm filter + forward-list dot -------------------
    if (a.head_mode) {
      for (int c = tid; c < TILE * TILE; c += SCAN_THREADS) {
        const int il = c / TILE, jl = c - (c / TILE) * TILE;
        if (il >= nI || jl >= nJ || (diag && jl <= il)) continue;
        float* cell = &tile[il * TSTRIDE + jl];
        float part = *cell;
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
    for (int c = tid; c < TILE * TILE; c += SCAN_THREADS) {
      const int il = c / TILE, jl = c - (c / TILE) * TILE;
      float* cell = &tile[il * TSTRIDE + jl];
      const float s = *cell;
      *cell = 0.f;
      const bool valid = il < nI && jl < nJ && !(diag && jl <= il);
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
