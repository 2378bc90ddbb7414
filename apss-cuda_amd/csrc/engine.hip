// Host orchestration of one APSS job on one GPU.
//
// Replaces the per-job sequence of similarity_join_device
// (similarity_search.cu:1273-1343 for a cross join, :2650-2691 for a self
// join): there, every job allocates a dense nX*nY float matrix (no memset),
// launches five kernels with a host synchronisation after each and reads the
// scan grid size back over PCIe.  Here a job is: validate + norms, feature
// histogram, tier selection (one small device->host copy of the histogram),
// cursor lists, per-tile bitmap tier, one band-scan launch over all tile pairs
// and one FP64 rescore launch; the only other synchronisations read the pair
// counters.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <sys/mman.h>
#include <thread>
#include <tuple>
#include <unordered_map>

#include "engine.h"

namespace apss {

// ---------------------------------------------------------------------------
// Arena

DevicePool::~DevicePool() {
  for (auto& b : free_) (void)hipFree(b.second);
}

void* DevicePool::take(size_t bytes, size_t* got) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = free_.lower_bound(bytes);
  if (it == free_.end() || it->first > 2 * bytes + (1 << 20)) return nullptr;
  void* p = it->second;
  *got = it->first;
  idle_ -= it->first;
  free_.erase(it);
  return p;
}

void DevicePool::give(void* p, size_t bytes) {
  {
    std::lock_guard<std::mutex> lk(mu_);
    if (keep_max_ == 0 || idle_ + bytes <= keep_max_) {
      free_.emplace(bytes, p);
      idle_ += bytes;
      return;
    }
  }
  (void)hipFree(p);
}

void DevicePool::trim() {
  std::multimap<size_t, void*> f;
  {
    std::lock_guard<std::mutex> lk(mu_);
    f.swap(free_);
    idle_ = 0;
  }
  for (auto& b : f) (void)hipFree(b.second);
}

int64_t DevicePool::held() const {
  std::lock_guard<std::mutex> lk(mu_);
  return (int64_t)idle_;
}

DevicePool& run_pool(int device) {
  // created on first use and never destroyed: the blocks stay with the
  // process (its exit releases them), and no static destructor runs after
  // the HIP runtime has shut down
  static std::mutex mu;
  static std::map<int, DevicePool*>* pools = new std::map<int, DevicePool*>;
  std::lock_guard<std::mutex> lk(mu);
  DevicePool*& p = (*pools)[device];
  if (!p) p = new DevicePool(RUN_POOL_KEEP);
  return *p;
}

Arena::~Arena() {
  // after an error, work on the stream may still be running
  if ((pool_ || bump_base_) && stream_ && !blocks_.empty()) (void)hipStreamSynchronize(stream_);
  if (bump_base_) return;  // the memory belongs to the caller
  for (auto& b : blocks_) drop(b.p, b.size);
}

void Arena::drop(void* p, size_t bytes) {
  if (pool_) pool_->give(p, bytes);
  else (void)hipFree(p);
}

void* Arena::raw(size_t bytes) {
  bytes = (bytes + 255) & ~(size_t)255;
  if (bump_base_) {
    if (used_ + (int64_t)bytes > budget_)
      throw Error{APSS_E_NOMEM, "input needs more than the " + std::to_string(budget_) +
                                    "-byte scratch array (" + std::to_string(used_ + bytes) +
                                    " bytes so far)"};
    void* p = bump_base_ + used_;
    blocks_.push_back({p, bytes, bytes});
    used_ += (int64_t)bytes;
    peak_ = std::max(peak_, used_);
    return p;
  }
  if (budget_ > 0 && used_ + (int64_t)bytes > budget_)
    throw Error{APSS_E_NOMEM, "device memory budget exceeded (" + std::to_string(used_ + bytes) +
                                  " > " + std::to_string(budget_) + " bytes)"};
  void* p = nullptr;
  size_t got = bytes;
  if (pool_) p = pool_->take(bytes, &got);
  if (!p) {
    got = bytes;
    bool ok = hipMalloc(&p, bytes) == hipSuccess && p;
    if (!ok && pool_ && pool_->held() > 0) {  // idle pooled blocks may be in the way
      (void)hipGetLastError();
      pool_->trim();
      p = nullptr;
      ok = hipMalloc(&p, bytes) == hipSuccess && p;
    }
    if (!ok) {
      (void)hipGetLastError();
      throw Error{APSS_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed"};
    }
  }
  blocks_.push_back({p, got, bytes});
  used_ += (int64_t)bytes;
  peak_ = std::max(peak_, used_);
  return p;
}

void Arena::release(void* p) {
  if (bump_base_) {
    for (auto& b : blocks_)
      if (b.p == p) b.freed = true;
    while (!blocks_.empty() && blocks_.back().freed) {
      used_ -= (int64_t)blocks_.back().req;
      blocks_.pop_back();
    }
    return;
  }
  for (size_t k = 0; k < blocks_.size(); ++k)
    if (blocks_[k].p == p) {
      drop(p, blocks_[k].size);
      used_ -= (int64_t)blocks_[k].req;
      blocks_.erase(blocks_.begin() + (long)k);
      return;
    }
}

void* Arena::detach(void* p) {
  for (size_t k = 0; k < blocks_.size(); ++k)
    if (blocks_[k].p == p) {
      used_ -= (int64_t)blocks_[k].req;
      blocks_.erase(blocks_.begin() + (long)k);
      return p;
    }
  return nullptr;
}

// ---------------------------------------------------------------------------
// Timer

Timer::Timer(hipStream_t st) : s(st) {
  APSS_HIP(hipEventCreate(&a));
  APSS_HIP(hipEventCreate(&b));
}
Timer::~Timer() {
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
}
void Timer::start() { APSS_HIP(hipEventRecord(a, s)); }
double Timer::stop_ms() {
  APSS_HIP(hipEventRecord(b, s));
  APSS_HIP(hipEventSynchronize(b));
  float ms = 0.f;
  APSS_HIP(hipEventElapsedTime(&ms, a, b));
  return ms;
}

// ---------------------------------------------------------------------------
// helpers

// APSS_TRACE=1 prints the host wall time of each phase of a job to stderr.
static bool trace_on() {
  static const bool on = [] {
    const char* e = std::getenv("APSS_TRACE");
    return e && *e && *e != '0';
  }();
  return on;
}
struct PhaseTrace {
  std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
  void mark(const char* what) {
    if (!trace_on()) return;
    const auto now = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[apss] %-28s %9.2f ms\n", what,
                 std::chrono::duration<double, std::milli>(now - t).count());
    t = now;
  }
};

static void fill_i32(hipStream_t st, int32_t* p, int64_t n, int32_t v) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(n, 256, 4096)), dim3(256), 0, st, p, n, v);
  APSS_HIP(hipGetLastError());
}
static void fill_u64(hipStream_t st, unsigned long long* p, int64_t n, unsigned long long v) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_u64, dim3(blocks_for(n, 256, 4096)), dim3(256), 0, st, p, n, v);
  APSS_HIP(hipGetLastError());
}

TileSide make_view(const Side& s, const TileSet& ts) {
  TileSide v;
  v.n = s.n;
  v.T = ts.T;
  v.tstart = ts.tstart;
  v.row2tile = ts.row2tile;
  v.rp = s.rp;
  v.col = s.col;
  v.val = s.val;
  v.norm = s.norm;
  v.hnorm = s.hnorm;
  v.bmp = ts.bmp;
  v.wpre = ts.wpre;
  v.cbase = ts.cbase;
  v.cst = ts.cst;
  v.dbase = ts.dbase;
  v.drow = ts.drow;
  v.dval = ts.dval;
  v.dfeat = ts.dfeat;
  v.sbase = ts.sbase;
  v.srec = ts.srec;
  v.dseg = ts.dseg;
  v.kds = ts.kds;
  v.l1m = ts.l1m;
  v.dense = s.dense;
  v.dnorm = s.dnorm;
  return v;
}

void release_tiles(Arena& ar, TileSet& ts) {
  void* arrays[] = {ts.tstart, ts.row2tile, ts.bmp, ts.wpre, ts.cbase, ts.cst, ts.dbase,
                    ts.drow, ts.dval, ts.dfeat, ts.sbase, ts.srec, ts.dseg, ts.l1m};
  for (void* q : arrays)
    if (q) ar.release(q);
  ts = TileSet{};
}

static std::string verr_text(uint32_t e) {
  std::string s;
  if (e & VERR_ROWPTR) s += " row_ptr not monotone or ends wrong;";
  if (e & VERR_COL) s += " column id out of range;";
  if (e & VERR_ORDER) s += " column ids not strictly ascending in a row;";
  if (e & VERR_VALUE) s += " non-finite value;";
  if (e & VERR_NORM) s += " row norm too large;";
  return s;
}

// ---------------------------------------------------------------------------
// prepare_side: upload + validation + norms

void check_csr_host(const apss_csr& X, int32_t p) {
  if (X.n_rows < 0 || X.n_cols < 0 || X.nnz < 0)
    throw Error{APSS_E_INVALID, "negative matrix dimension"};
  if (X.n_rows > 0 && (!X.row_ptr || (X.nnz > 0 && (!X.col || !X.val))))
    throw Error{APSS_E_INVALID, "null CSR array"};
  if (X.n_rows >= (int64_t)0x7fffffff || X.n_cols >= (int64_t)0x7fffffff)
    throw Error{APSS_E_UNSUPPORTED, "more than 2^31-1 rows or columns"};
  if (p < 0 || p > X.n_cols) throw Error{APSS_E_INVALID, "p outside [0, n_cols]"};
  if (!X.on_device && X.n_rows > 0) {
    if (X.row_ptr[0] != 0) throw Error{APSS_E_INVALID, "row_ptr[0] != 0"};
    if (X.row_ptr[X.n_rows] != X.nnz) throw Error{APSS_E_INVALID, "row_ptr[n_rows] != nnz"};
  }
}

Side prepare_side(Arena& ar, hipStream_t st, const apss_csr& X, int32_t p) {
  check_csr_host(X, p);
  Side s;
  s.n = X.n_rows;
  s.nnz = X.nnz;
  s.d = X.n_cols;
  if (X.on_device) {
    s.rp = X.row_ptr;
    s.col = X.col;
    s.val = X.val;
  } else {
    int64_t* rp = ar.alloc<int64_t>(s.n + 1);
    int32_t* col = ar.alloc<int32_t>(s.nnz);
    float* val = ar.alloc<float>(s.nnz);
    std::vector<H2DCopy> cp;
    if (s.n >= 0) cp.push_back({rp, X.row_ptr, sizeof(int64_t) * (size_t)(s.n + 1)});
    if (s.nnz > 0) {
      cp.push_back({col, X.col, sizeof(int32_t) * (size_t)s.nnz});
      cp.push_back({val, X.val, sizeof(float) * (size_t)s.nnz});
    }
    h2d_copies(cp, st);
    s.rp = rp;
    s.col = col;
    s.val = val;
  }
  s.norm = ar.alloc<float>(s.n);
  s.hnorm = ar.alloc<float>(s.n);
  uint32_t* err = ar.alloc<uint32_t>(3);  // error bits, longest row, largest norm
  fill_i32(st, (int32_t*)err, 3, 0);
  if (s.n > 0) {
    // a grid of at most 2048 blocks: each wave takes many rows, so few waves
    // post their maxima with atomics
    hipLaunchKernelGGL(k_validate_norms, dim3(blocks_for(s.n, 4, 2048)), dim3(256), 0, st, s.n,
                       s.d, s.nnz, s.rp, s.col, s.val, p, s.norm, s.hnorm, err);
    APSS_HIP(hipGetLastError());
  }
  uint32_t herr[3] = {0, 0, 0};
  APSS_HIP(hipMemcpyAsync(herr, err, sizeof(herr), hipMemcpyDeviceToHost, st));
  APSS_HIP(hipStreamSynchronize(st));
  ar.release(err);
  if (herr[0] & ~VFLAG_NEG) throw Error{APSS_E_INVALID, "invalid CSR:" + verr_text(herr[0])};
  s.nonneg = (herr[0] & VFLAG_NEG) == 0;
  // maxima for the global fast-reject bound
  s.maxlen = (int32_t)herr[1];
  std::memcpy(&s.maxnorm, &herr[2], sizeof(float));
  return s;
}

// ---------------------------------------------------------------------------
// build_tiers: histogram, class per feature, cursor lists

int32_t dense_mode(const apss_opts& o) {
  if (o.mfma_head < 0) return DENSE_OFF;
  return o.precision == APSS_PREC_BF16_HEAD ? DENSE_BF16 : DENSE_I8;
}

// A host array for the index build's transfers, from the pinned block pool
// (pageable below the pool's minimum block): copies to and from it run at DMA
// speed, and reused blocks are already faulted in.  Every asynchronous copy
// that reads or writes it must have completed before it goes out of scope.
template <class T>
struct HostBuf {
  T* p = nullptr;
  explicit HostBuf(int64_t n) {
    const size_t bytes = sizeof(T) * (size_t)std::max<int64_t>(n, 1);
    p = reinterpret_cast<T*>(host_pairs_alloc((int64_t)((bytes + sizeof(apss_pair) - 1) / sizeof(apss_pair))));
  }
  ~HostBuf() { host_pairs_free(reinterpret_cast<apss_pair*>(p)); }
  HostBuf(const HostBuf&) = delete;
  HostBuf& operator=(const HostBuf&) = delete;
  T& operator[](int64_t i) { return p[i]; }
  const T& operator[](int64_t i) const { return p[i]; }
  T* data() { return p; }
};

Tiers build_tiers(Arena& ar, hipStream_t st, const Side& s, int32_t p, int32_t self, int32_t kd_cap,
                  int32_t dm, bool all_lists) {
  Tiers tr;
  tr.d = s.d;
  tr.p = p;
  const int64_t d = s.d;
  int32_t* cnt = ar.alloc<int32_t>(d + 1);
  fill_i32(st, cnt, d + 1, 0);
  if (s.nnz > 0) {
    const int64_t chunk = 1 << 16;
    hipLaunchKernelGGL(k_col_hist, dim3(blocks_for(s.nnz, (int)chunk, 1 << 30)), dim3(1024), 0, st,
                       s.nnz, s.col, p, chunk, cnt);
    APSS_HIP(hipGetLastError());
  }
  HostBuf<int32_t> h(d + 1);
  APSS_HIP(hipMemcpyAsync(h.data(), cnt, sizeof(int32_t) * (d + 1), hipMemcpyDeviceToHost, st));
  APSS_HIP(hipStreamSynchronize(st));
  ar.release(cnt);

  const int32_t need = self && !all_lists ? 2 : 1;
  // bitmap tier: the lists present in a sizeable share of the 64-row tiles,
  // plus every list longer than one LDS sort piece (the cursor tier would walk
  // them entry by entry), most frequent first, up to the tier's capacity; the
  // most frequent of them form the dense MFMA head when there is one
  const int64_t T = std::max<int64_t>(1, (s.n + TILE_ROWS - 1) / TILE_ROWS);
  const int64_t lmin = std::max<int64_t>(need, T / 8);
  std::vector<int32_t> cand;
  for (int64_t c = p; c < d && !all_lists; ++c)
    if (h[c] >= lmin || h[c] > S_LIST_MAX) cand.push_back((int32_t)c);
  if (all_lists) dm = DENSE_OFF;
  std::stable_sort(cand.begin(), cand.end(), [&](int32_t x, int32_t y) { return h[x] > h[y]; });
  const int32_t slots = dm == DENSE_I8 ? DENSE_I8_FEATURES : dm == DENSE_BF16 ? DENSE_BF16_FEATURES : 0;
  const int32_t n_dense = std::min<int32_t>(slots, (int32_t)cand.size());
  if (n_dense == 0) dm = DENSE_OFF;
  std::vector<int32_t> dense(cand.begin(), cand.begin() + n_dense);
  cand.erase(cand.begin(), cand.begin() + n_dense);
  const int32_t kdm = kd_max_for(dm);
  int32_t cap = kd_cap > 0 ? std::min(kd_cap, kdm) : kdm;
  if ((int64_t)cand.size() > cap) cand.resize(cap);
  tr.dm = dm;
  tr.n_dense = n_dense;
  tr.n_bitmap = (int32_t)cand.size();
  tr.KD = std::max(64, (tr.n_bitmap + 63) / 64 * 64);
  tr.KW = tr.KD / 64;
  // The bitmap index of a feature is its frequency rank, so every tile block
  // holds its segments from the most to the least frequent feature: threads
  // that own neighbouring entries see segments of similar length.
  HostBuf<int32_t> fm(d);
  for (int64_t c = 0; c < d; ++c) fm[c] = c < p ? F_HEAD : h[c] >= need ? F_CURSOR : F_IGNORE;
  for (size_t k = 0; k < cand.size(); ++k) fm[cand[k]] = (int32_t)k;
  for (size_t k = 0; k < dense.size(); ++k) fm[dense[k]] = F_DENSE0 - (int32_t)k;
  HostBuf<int64_t> sp(d + 1);
  sp[0] = 0;
  std::vector<int32_t> small_lists;
  std::vector<SortChunk> chunks;  // lists of 65.. entries, cut in LDS-sized pieces
  int64_t max_long = 0;           // longest list cut in more than one piece
  for (int64_t c = 0; c < d; ++c) {
    int64_t L = fm[c] == F_CURSOR ? h[c] : 0;
    sp[c + 1] = sp[c] + L;
    if (L >= 2 && L <= 64) small_lists.push_back((int32_t)c);
    else if (L > 64)
      for (int64_t s0 = 0; s0 < L; s0 += S_LIST_MAX)
        chunks.push_back(SortChunk{sp[c] + s0, (int32_t)std::min<int64_t>(S_LIST_MAX, L - s0), 0});
    if (L > S_LIST_MAX) max_long = std::max(max_long, L);
  }
  tr.n_cursor_entries = sp[d];
  if (tr.n_cursor_entries >= ((int64_t)1 << 31))
    throw Error{APSS_E_UNSUPPORTED, "cursor tier larger than 2^31 entries"};
  tr.fmap = ar.alloc<int32_t>(d);
  tr.sptr = ar.alloc<int64_t>(d + 1);
  tr.sent = ar.alloc<int2>(tr.n_cursor_entries);
  APSS_HIP(hipMemcpyAsync(tr.fmap, fm.data(), sizeof(int32_t) * d, hipMemcpyHostToDevice, st));
  APSS_HIP(hipMemcpyAsync(tr.sptr, sp.data(), sizeof(int64_t) * (d + 1), hipMemcpyHostToDevice, st));
  if (tr.n_cursor_entries > 0) {
    // scatter (list order within a feature is arbitrary), then sort each list by row
    unsigned long long* cur = ar.alloc<unsigned long long>(d + 1);
    hipLaunchKernelGGL(k_copy_i64, dim3(blocks_for(d + 1, 256, 4096)), dim3(256), 0, st,
                       (int64_t*)cur, tr.sptr, d + 1);
    hipLaunchKernelGGL(k_cursor_scatter, dim3(blocks_for(s.n, 4, 65536)), dim3(256), 0, st, s.n,
                       s.rp, s.col, s.val, tr.fmap, cur, tr.sent);
    APSS_HIP(hipGetLastError());
    int32_t* lists = ar.alloc<int32_t>((int64_t)small_lists.size());
    if (!small_lists.empty()) {
      APSS_HIP(hipMemcpyAsync(lists, small_lists.data(), sizeof(int32_t) * small_lists.size(),
                              hipMemcpyHostToDevice, st));
      hipLaunchKernelGGL(k_sort_small, dim3(blocks_for((int64_t)small_lists.size(), 4, 1 << 30)),
                         dim3(256), 0, st, lists, (int32_t)small_lists.size(), tr.sptr, tr.sent);
      APSS_HIP(hipGetLastError());
    }
    SortChunk* dch = ar.alloc<SortChunk>((int64_t)chunks.size());
    if (!chunks.empty()) {
      APSS_HIP(hipMemcpyAsync(dch, chunks.data(), sizeof(SortChunk) * chunks.size(),
                              hipMemcpyHostToDevice, st));
      hipLaunchKernelGGL(k_sort_chunks, dim3((unsigned)chunks.size()), dim3(256), 0, st, dch, tr.sent);
      APSS_HIP(hipGetLastError());
    }
    // lists longer than one piece: pairwise merges of sorted runs, doubling the
    // run width each pass, ping-ponging between sent and a second buffer (every
    // long list takes part in every pass, as a copy once it is one run)
    if (max_long > 0) {
      int2* other = ar.alloc<int2>(tr.n_cursor_entries);
      int2* src = tr.sent;
      int2* dst = other;
      std::vector<MergeTask> tasks;
      for (int64_t w = S_LIST_MAX; w < max_long; w *= 2) {
        tasks.clear();
        for (int64_t c = 0; c < d; ++c) {
          const int64_t L = fm[c] == F_CURSOR ? h[c] : 0;
          if (L <= S_LIST_MAX) continue;
          for (int64_t s0 = 0; s0 < L; s0 += 2 * w) {
            const int32_t la = (int32_t)std::min<int64_t>(w, L - s0);
            const int32_t lb = (int32_t)std::max<int64_t>(0, std::min<int64_t>(w, L - s0 - w));
            for (int32_t k = 0; (int64_t)k * MERGE_TILE < (int64_t)la + lb; ++k)
              tasks.push_back(MergeTask{sp[c] + s0, la, lb, k, 0});
          }
        }
        MergeTask* dt = ar.alloc<MergeTask>((int64_t)tasks.size());
        APSS_HIP(hipMemcpyAsync(dt, tasks.data(), sizeof(MergeTask) * tasks.size(),
                                hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_merge_runs, dim3((unsigned)tasks.size()), dim3(256), 0, st, dt, src, dst);
        APSS_HIP(hipGetLastError());
        APSS_HIP(hipStreamSynchronize(st));  // the task array is released next
        ar.release(dt);
        std::swap(src, dst);
      }
      // short lists were sorted in place in sent; the long ones end in src
      if (src != tr.sent) {
        for (int64_t c = 0; c < d; ++c) {
          const int64_t L = fm[c] == F_CURSOR ? h[c] : 0;
          if (L > S_LIST_MAX)
            APSS_HIP(hipMemcpyAsync(tr.sent + sp[c], src + sp[c], sizeof(int2) * L,
                                    hipMemcpyDeviceToDevice, st));
        }
      }
      APSS_HIP(hipStreamSynchronize(st));
      ar.release(other);
    }
    APSS_HIP(hipStreamSynchronize(st));
    ar.release(dch);
    ar.release(lists);
    ar.release(cur);
  }
  APSS_HIP(hipStreamSynchronize(st));  // the uploads from fm and sp are complete
  return tr;
}

// ---------------------------------------------------------------------------
// build_tiles

// Greedy cut of the rows into tiles of at most max_rows rows whose bitmap-tier
// entries (padded to a multiple of 4) stay within cap_d and whose cursor-tier
// entries stay within cap_s (a single row always forms a tile).
static std::vector<int64_t> cut_tiles(const int32_t* dc, const int32_t* sc, int64_t n,
                                      int32_t max_rows, int32_t cap_d, int32_t cap_s,
                                      const std::vector<int64_t>& bounds) {
  std::vector<int64_t> ts;
  ts.push_back(0);
  int64_t rows = 0, sd = 0, ss = 0;
  size_t nb = 0;  // next forced boundary (instance start)
  while (nb < bounds.size() && bounds[nb] <= 0) ++nb;
  for (int64_t r = 0; r < n; ++r) {
    const int64_t nd = ((sd + dc[r] + 3) / 4) * 4;
    const bool forced = nb < bounds.size() && bounds[nb] == r;
    if (forced) ++nb;
    if (rows > 0 && (forced || rows == max_rows || nd > cap_d || ss + sc[r] > cap_s)) {
      ts.push_back(r);
      rows = sd = ss = 0;
    }
    ++rows;
    sd += dc[r];
    ss += sc[r];
  }
  if (n > 0) ts.push_back(n);
  return ts;
}

TileSet build_tiles(Arena& ar, hipStream_t st, const Side& s, const Tiers& tr, int32_t self,
                    int32_t max_rows, int32_t cap_d, int32_t cap_s, bool records,
                    const std::vector<int64_t>& bounds) {
  TileSet ts;
  std::vector<int64_t> cuts(1, 0);
  if (s.n > 0) {
    int32_t* cnts = ar.alloc<int32_t>(3 * s.n);
    hipLaunchKernelGGL(k_row_counts, dim3(blocks_for(s.n, 4, 65536)), dim3(256), 0, st, s.n, s.rp,
                       s.col, tr.p, tr.fmap, cnts, cnts + s.n, cnts + 2 * s.n);
    APSS_HIP(hipGetLastError());
    // the three count arrays in one pinned block: dc, sc, hc
    HostBuf<int32_t> hcnt(3 * s.n);
    APSS_HIP(hipMemcpyAsync(hcnt.data(), cnts, sizeof(int32_t) * 3 * s.n, hipMemcpyDeviceToHost, st));
    APSS_HIP(hipStreamSynchronize(st));
    ar.release(cnts);
    int32_t* dc = hcnt.data();
    int32_t* sc = dc + s.n;
    const int32_t* hc = sc + s.n;
    if (!records) std::fill(sc, sc + s.n, 0);
    cuts = cut_tiles(dc, sc, s.n, max_rows, cap_d, cap_s, bounds);
    ts.rare.assign(cuts.size() - 1, 1);
    for (size_t k = 0; k + 1 < cuts.size(); ++k) {
      int64_t m = 0;
      for (int64_t r = cuts[k]; r < cuts[k + 1]; ++r) {
        m += sc[r];
        if (dc[r] != 0 || hc[r] != 0) ts.rare[k] = 0;
      }
      ts.max_s = std::max(ts.max_s, m);
    }
  }
  const int32_t T = (int32_t)cuts.size() - 1;
  ts.T = T;
  ts.cuts = cuts;
  ts.tstart = ar.alloc<int64_t>(T + 1);
  APSS_HIP(hipMemcpyAsync(ts.tstart, cuts.data(), sizeof(int64_t) * (T + 1), hipMemcpyHostToDevice, st));
  // cursor-list positions pack the J tile into 24 bits (tile << 8 | row)
  if (T >= (1 << 24)) throw Error{APSS_E_UNSUPPORTED, "more than 2^24 - 1 tiles"};
  ts.row2tile = ar.alloc<int32_t>(s.n);
  if (T > 0) {
    hipLaunchKernelGGL(k_row2tile, dim3(T), dim3(256), 0, st, ts.tstart, ts.row2tile);
    APSS_HIP(hipGetLastError());
  }
  ts.bmp = ar.alloc<uint64_t>((int64_t)T * tr.KW);
  ts.wpre = ar.alloc<uint16_t>((int64_t)T * tr.KW);
  ts.cbase = ar.alloc<int64_t>((int64_t)T + 1);
  ts.dbase = ar.alloc<int64_t>((int64_t)T + 1);
  ts.sbase = ar.alloc<int64_t>((int64_t)T + 1);
  int32_t* full = ar.alloc<int32_t>((int64_t)T * (tr.KD + 1));
  fill_i32(st, full, (int64_t)T * (tr.KD + 1), 0);
  BuildSide b{};
  b.n = s.n;
  b.T = T;
  b.KD = tr.KD;
  b.KW = tr.KW;
  b.p = tr.p;
  b.self = self;
  b.tstart = ts.tstart;
  b.rp = s.rp;
  b.col = s.col;
  b.val = s.val;
  b.fmap = tr.fmap;
  b.full = full;
  b.bmp = ts.bmp;
  b.wpre = ts.wpre;
  b.ctot = ts.cbase;
  b.dtot = ts.dbase;
  b.stot = ts.sbase;
  b.sptr = tr.sptr;
  b.sent = tr.sent;
  if (T > 0) {
    hipLaunchKernelGGL(k_tile_count, dim3(T), dim3(256), 0, st, b);
    hipLaunchKernelGGL(k_tile_ptr_scan, dim3(T), dim3(1024), 0, st, b);
    hipLaunchKernelGGL(k_scan_i64, dim3(1), dim3(1024), 0, st, ts.dbase, (int64_t)T);
    hipLaunchKernelGGL(k_scan_i64, dim3(1), dim3(1024), 0, st, ts.sbase, (int64_t)T);
    hipLaunchKernelGGL(k_scan_i64, dim3(1), dim3(1024), 0, st, ts.cbase, (int64_t)T);
    APSS_HIP(hipGetLastError());
  } else {
    fill_u64(st, (unsigned long long*)ts.dbase, 1, 0);
    fill_u64(st, (unsigned long long*)ts.sbase, 1, 0);
    fill_u64(st, (unsigned long long*)ts.cbase, 1, 0);
  }
  int64_t tot[3] = {0, 0, 0};
  APSS_HIP(hipMemcpyAsync(&tot[0], ts.dbase + T, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  APSS_HIP(hipMemcpyAsync(&tot[1], ts.sbase + T, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  APSS_HIP(hipMemcpyAsync(&tot[2], ts.cbase + T, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  APSS_HIP(hipStreamSynchronize(st));
  ts.nD = tot[0];
  ts.nS = records ? tot[1] : 0;
  ts.drow = ar.alloc<uint8_t>(ts.nD);
  ts.dval = ar.alloc<float>(ts.nD);
  ts.dfeat = ar.alloc<uint16_t>(ts.nD);
  ts.cst = ar.alloc<int32_t>(tot[2]);
  ts.srec = ar.alloc<int4>(records ? tot[1] : 0);
  b.drow = ts.drow;
  b.dval = ts.dval;
  b.dfeat = ts.dfeat;
  b.cst = ts.cst;
  b.srec = records ? ts.srec : nullptr;
  // row stride: KD + 1 starts rounded up to 4 (8-byte rows for the staging loads)
  ts.kds = (tr.KD + 1 + 3) / 4 * 4;
  ts.dseg = ar.alloc<uint16_t>((int64_t)std::max(T, 1) * ts.kds);
  b.dseg = ts.dseg;
  b.kds = ts.kds;
  // per-tile L1 x max statistics (block bounds of sparse bands)
  ts.l1m = ar.alloc<float2>(std::max(T, 1));
  if (T > 0) {
    hipLaunchKernelGGL(k_tile_l1max, dim3(T), dim3(256), 0, st, ts.tstart, s.rp, s.val, ts.l1m);
    hipLaunchKernelGGL(k_tile_scatter, dim3(T), dim3(256), 0, st, b);
    hipLaunchKernelGGL(k_tile_compact, dim3(T), dim3(256), 0, st, b);
    if (records && ts.nS > 0)
      hipLaunchKernelGGL(k_sort_records, dim3(T), dim3(256), 0, st, ts.sbase, ts.srec);
    APSS_HIP(hipGetLastError());
  }
  APSS_HIP(hipStreamSynchronize(st));
  ar.release(full);
  return ts;
}

void set_cursor_tiles(Arena& ar, hipStream_t st, Tiers& tr, const TileSet& tiles) {
  if (tr.spos) ar.release(tr.spos);
  if (tr.sval) ar.release(tr.sval);
  tr.spos = ar.alloc<uint32_t>(tr.n_cursor_entries + WALK_PAD);
  tr.sval = ar.alloc<float>(tr.n_cursor_entries + WALK_PAD);
  if (tr.n_cursor_entries > 0) {
    hipLaunchKernelGGL(k_cursor_pos, dim3(blocks_for(tr.n_cursor_entries, 256, 65536)), dim3(256),
                       0, st, tr.n_cursor_entries, tr.sent, tiles.row2tile, tiles.tstart, tr.spos,
                       tr.sval);
    APSS_HIP(hipGetLastError());
  }
}

void build_dense(Arena& ar, hipStream_t st, Side& s, const Tiers& tr) {
  s.dense = nullptr;
  s.dnorm = nullptr;
  if (tr.dm == DENSE_OFF) return;
  // int8 head: |x| <= maxnorm, so x * 2^B stays within 2^28 (four limbs)
  // (a side whose values are all zero is exact at any scale: use norm 1)
  const double m = s.maxnorm > 0.f ? (double)s.maxnorm : 1.0;
  s.dense_B = 28 - (int)std::ceil(std::log2(m));
  s.dense = ar.alloc<uint4>(std::max<int64_t>(s.n, 1) * 16);
  if (tr.dm == DENSE_BF16) s.dnorm = ar.alloc<float>(s.n);
  if (s.n > 0) {
    hipLaunchKernelGGL(k_dense_rows, dim3(blocks_for(s.n, 4, 65536)), dim3(256), 0, st, s.n, s.rp,
                       s.col, s.val, tr.p, tr.fmap, tr.dm, s.dense_B, (uint32_t*)s.dense, s.dnorm);
    APSS_HIP(hipGetLastError());
  }
}

Built build_config(Arena& ar, hipStream_t st, Side& s, int32_t p, int32_t self, int32_t kd_cap,
                   int32_t dm, const std::vector<int64_t>& bounds, int32_t cap_s, int32_t band_rows,
                   int32_t tile_rows) {
  Built b;
  PhaseTrace ph;
  // with APSS_TRACE set, each step is synchronised so the trace shows its device time
  auto step = [&](const char* what) {
    if (!trace_on()) return;
    APSS_HIP(hipStreamSynchronize(st));
    ph.mark(what);
  };
  b.tr = build_tiers(ar, st, s, p, self, kd_cap, dm);
  step("index: tiers + cursor lists");
  b.bands = build_tiles(ar, st, s, b.tr, self, band_rows, cap_i_for(b.tr.dm), cap_s, true, bounds);
  step("index: band blocks");
  b.tiles = build_tiles(ar, st, s, b.tr, self, tile_rows, cap_j_for(b.tr.dm), INT32_MAX, false, bounds);
  step("index: tile blocks");
  set_cursor_tiles(ar, st, b.tr, b.tiles);
  build_dense(ar, st, s, b.tr);
  step("index: cursor tiles + dense");
  return b;
}

TuneChoice choice_from_opts(const apss_opts& o) {
  TuneChoice c;
  if (o.band_rows < 0 || o.band_rows == 1 || o.band_rows > BAND_ROWS)
    throw Error{APSS_E_INVALID, "band_rows outside 2..128"};
  if (o.band_rows > 0) {
    c.band_rows = o.band_rows;
    c.tile_rows = o.band_rows / 2;
  }
  return c;
}

Side with_split(Arena& ar, hipStream_t st, const Side& s, int32_t p) {
  Side r = s;
  r.norm = ar.alloc<float>(s.n);
  r.hnorm = ar.alloc<float>(s.n);
  uint32_t* err = ar.alloc<uint32_t>(3);
  fill_i32(st, (int32_t*)err, 3, 0);
  if (s.n > 0) {
    hipLaunchKernelGGL(k_validate_norms, dim3(blocks_for(s.n, 4, 2048)), dim3(256), 0, st, s.n, s.d,
                       s.nnz, s.rp, s.col, s.val, p, r.norm, r.hnorm, err);
    APSS_HIP(hipGetLastError());
  }
  return r;
}

namespace {
// The autotuner's choices by job shape (see autotune).
struct TuneKey {
  int device;
  int64_t n, nnz, d;
  int32_t maxlen;
  uint32_t maxnorm_bits;
  int32_t self;
  uint64_t bounds_hash;
  uint64_t t_bits;
  int32_t dense_features, head;
  bool operator<(const TuneKey& o) const {
    return std::tie(device, n, nnz, d, maxlen, maxnorm_bits, self, bounds_hash, t_bits,
                    dense_features, head) < std::tie(o.device, o.n, o.nnz, o.d, o.maxlen,
                                                     o.maxnorm_bits, o.self, o.bounds_hash,
                                                     o.t_bits, o.dense_features, o.head);
  }
};
struct TuneCache {
  std::mutex mu;
  std::map<TuneKey, TuneChoice> map;
};
TuneCache& tune_cache() {
  static TuneCache c;
  return c;
}
constexpr size_t TUNE_CACHE_MAX = 256;
}  // namespace

TuneChoice autotune(hipStream_t st, const apss_opts& o, Side& s, int32_t self,
                    const std::vector<int64_t>& bounds, double t,
                    const std::function<void(RunConfig&, const Built&)>& cfg) {
  const auto t0 = std::chrono::steady_clock::now();
  TuneKey key{o.device, s.n, s.nnz, s.d, s.maxlen, 0u, self, 1469598103934665603ull, 0u,
              o.dense_features, dense_mode(o)};
  std::memcpy(&key.maxnorm_bits, &s.maxnorm, sizeof(float));
  std::memcpy(&key.t_bits, &t, sizeof(double));
  for (int64_t b : bounds) key.bounds_hash = (key.bounds_hash ^ (uint64_t)b) * 1099511628211ull;
  {
    TuneCache& tcache = tune_cache();
    std::lock_guard<std::mutex> lk(tcache.mu);
    const auto it = tcache.map.find(key);
    if (it != tcache.map.end()) {
      TuneChoice c = it->second;
      c.cached = true;
      c.candidates = 0;
      c.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      return c;
    }
  }
  // Candidates are timed on a prefix of the rows (1/16, at least 16384 and at
  // most 65536 rows): the tiers keep their relative shape (the bitmap-tier
  // threshold scales with the tile count) and each candidate's build costs a
  // fraction of the full one.
  Side sub = s;
  std::vector<int64_t> sbounds;
  if (s.n > 32768) {
    // the prefix ends at an instance start where one is near (a batched run's
    // bounds), so no instance is cut short
    int64_t m = std::min<int64_t>(65536, std::max<int64_t>(16384, s.n / 16));
    const auto nb = std::lower_bound(bounds.begin(), bounds.end(), m);
    if (nb != bounds.end() && *nb <= 2 * m) m = *nb;
    int64_t e = 0;
    APSS_HIP(hipMemcpyAsync(&e, s.rp + m, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    APSS_HIP(hipStreamSynchronize(st));
    sub.n = m;
    sub.nnz = e;
    for (int64_t x : bounds)
      if (x < m) sbounds.push_back(x);
  } else {
    sbounds = bounds;
  }
  TuneChoice best;
  best.est_ms = INFINITY;
  // One candidate: build its structures on the prefix (in a pooled arena, so
  // the candidates reuse each other's blocks; candidates that differ only in
  // the schedule or in the band sample share one build) and time a count-only
  // scan; tune_bands > 0 samples that many bands (all their partner tiles), 0
  // scans the whole prefix.  Returns the time scaled to all of the prefix's
  // tile pairs and the dense-head mode the build ended with.
  struct Build {
    TuneChoice key;
    std::unique_ptr<Arena> ar;
    Side side;
    Built b;
  };
  std::vector<std::unique_ptr<Build>> builds;
  auto same = [](const TuneChoice& x, const TuneChoice& y) {
    return x.p == y.p && x.dm == y.dm && x.kd_cap == y.kd_cap && x.band_rows == y.band_rows &&
           x.tile_rows == y.tile_rows;
  };
  auto measure = [&](const TuneChoice& c, int32_t tune_bands) {
    Build* bd = nullptr;
    for (auto& x : builds)
      if (same(x->key, c)) bd = x.get();
    const auto tb = std::chrono::steady_clock::now();
    if (!bd) {
      auto nb = std::make_unique<Build>();
      nb->key = c;
      nb->ar = std::make_unique<Arena>(o.device, 0, &run_pool(o.device));
      nb->ar->set_stream(st);
      nb->side = c.p > 0 ? with_split(*nb->ar, st, sub, c.p) : sub;
      nb->b = build_config(*nb->ar, st, nb->side, c.p, self, c.kd_cap, c.dm, sbounds, CAP_S,
                           c.band_rows, c.tile_rows);
      bd = nb.get();
      builds.push_back(std::move(nb));
      if (trace_on()) APSS_HIP(hipStreamSynchronize(st));
    }
    const auto ts = std::chrono::steady_clock::now();
    RunConfig rc;
    cfg(rc, bd->b);
    rc.p = c.p;
    rc.tune_only = true;
    rc.tune_bands = tune_bands;
    rc.items_per_cu = c.items_per_cu;
    ScanOutput so = scan_and_rescore(*bd->ar, st, bd->side, bd->b.bands, bd->side, bd->b.tiles, bd->b.tr, rc);
    ++best.candidates;
    if (trace_on()) {
      const auto te = std::chrono::steady_clock::now();
      std::fprintf(stderr,
                   "[apss] tune: p %d dense %d bitmap cap %d rows %d/%d items/CU %d -> %.2f ms "
                   "(build %.1f ms, scan wall %.1f ms)\n",
                   c.p, bd->b.tr.dm, c.kd_cap, c.band_rows, c.tile_rows, c.items_per_cu, so.ms_scan,
                   std::chrono::duration<double, std::milli>(ts - tb).count(),
                   std::chrono::duration<double, std::milli>(te - ts).count());
    }
    return std::make_pair(so.ms_scan, bd->b.tr.dm);
  };
  const int32_t head = dense_mode(o);
  TuneChoice cur;
  cur.dm = head;
  cur.kd_cap = o.dense_features;
  // 1. the split p (the reference's knob): ids below p take the prefix-norm
  //    filter and the forward-list head dot.  Timed on a sample of 16 bands,
  //    since a large p can cost a hundred times the scan.  The head-dot work
  //    only grows with p, so the sweep (ascending p) stops at the first
  //    candidate slower than twice the best one.
  double best_p_ms = INFINITY;
  for (int32_t p : {0, 1024, 4096}) {
    if (p > s.d) continue;
    TuneChoice c = cur;
    c.p = p;
    const double ms = measure(c, 16).first;
    if (ms < best_p_ms) {
      best_p_ms = ms;
      cur.p = p;
    } else if (ms > 2.0 * best_p_ms) {
      break;
    }
  }
  // 2. the tiering of the tail and the band / tile heights, on the whole prefix
  std::vector<TuneChoice> cands;
  auto add = [&](int32_t dm, int32_t cap, int32_t br, int32_t tr_) {
    TuneChoice c = cur;
    c.dm = dm;
    c.kd_cap = cap;
    c.band_rows = br;
    c.tile_rows = tr_;
    cands.push_back(c);
  };
  add(head, cur.kd_cap, BAND_ROWS, TILE_ROWS);
  if (head != DENSE_OFF) add(DENSE_OFF, cur.kd_cap, BAND_ROWS, TILE_ROWS);
  add(head, 1024, BAND_ROWS, TILE_ROWS);
  add(head, cur.kd_cap, BAND_ROWS / 2, TILE_ROWS / 2);
  for (const TuneChoice& c : cands) {
    const auto r = measure(c, 0);
    if (r.first < best.est_ms) {
      const int32_t cnt = best.candidates;
      best = c;
      best.dm = r.second;
      best.est_ms = r.first;
      best.candidates = cnt;
    }
  }
  // 3. the schedule: work items per CU (the chunk of partner tiles an item
  //    covers), which trades load balance against per-item set-up
  for (int32_t ipc : {4, 16}) {
    TuneChoice c = best;
    c.items_per_cu = ipc;
    const auto r = measure(c, 0);
    if (r.first < best.est_ms) {
      const int32_t cnt = best.candidates;
      best = c;
      best.est_ms = r.first;
      best.candidates = cnt;
    }
  }
  APSS_HIP(hipStreamSynchronize(st));
  best.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  {
    TuneCache& tcache = tune_cache();
    std::lock_guard<std::mutex> lk(tcache.mu);
    if (tcache.map.size() >= TUNE_CACHE_MAX) tcache.map.clear();
    tcache.map[key] = best;
  }
  return best;
}

// ---------------------------------------------------------------------------
// scan + rescore

// The band kernel's entry for a dense-head mode, split and work counting
// (graph nodes).
template <bool ST>
static void* band_scan_fn_t(int32_t dm, bool hm) {
  if (hm)
    return dm == DENSE_I8 ? reinterpret_cast<void*>(&k_band_scan<DENSE_I8, true, false, ST>)
         : dm == DENSE_BF16 ? reinterpret_cast<void*>(&k_band_scan<DENSE_BF16, true, false, ST>)
                            : reinterpret_cast<void*>(&k_band_scan<DENSE_OFF, true, false, ST>);
  return dm == DENSE_I8 ? reinterpret_cast<void*>(&k_band_scan<DENSE_I8, false, false, ST>)
       : dm == DENSE_BF16 ? reinterpret_cast<void*>(&k_band_scan<DENSE_BF16, false, false, ST>)
                          : reinterpret_cast<void*>(&k_band_scan<DENSE_OFF, false, false, ST>);
}
static void* band_scan_fn(int32_t dm, bool hm, bool st) {
  return st ? band_scan_fn_t<true>(dm, hm) : band_scan_fn_t<false>(dm, hm);
}

namespace {
// The device job of a batched run as a graph of four kernel nodes (counter
// clear, band scan, its sparse form, FP64 rescore), kept per device and reused: a later job
// with the same band kernel updates the nodes' arguments and grids in the
// instantiated graph (hipGraphExecKernelNodeSetParams) instead of building and
// instantiating a new one.  Callers hold `mu` from the update until their job
// has finished on its stream.  No memset node: the counters are cleared by a
// fill kernel.
struct JobGraph {
  std::mutex mu;
  hipGraph_t g = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraphNode_t node[4] = {nullptr, nullptr, nullptr, nullptr};
  void* scan_fn = nullptr;  // the band kernel the graph holds
  void build(const hipKernelNodeParams* np) {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (g) (void)hipGraphDestroy(g);
    exec = nullptr;
    g = nullptr;
    scan_fn = nullptr;
    APSS_HIP(hipGraphCreate(&g, 0));
    APSS_HIP(hipGraphAddKernelNode(&node[0], g, nullptr, 0, &np[0]));
    APSS_HIP(hipGraphAddKernelNode(&node[1], g, &node[0], 1, &np[1]));
    APSS_HIP(hipGraphAddKernelNode(&node[2], g, &node[1], 1, &np[2]));
    APSS_HIP(hipGraphAddKernelNode(&node[3], g, &node[2], 1, &np[3]));
    APSS_HIP(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
    scan_fn = np[1].func;
  }
};
JobGraph& job_graph(int device) {
  // created on first use and never destroyed (no destructor runs after the
  // HIP runtime has shut down)
  static std::mutex mu;
  static std::map<int, JobGraph*>* graphs = new std::map<int, JobGraph*>;
  std::lock_guard<std::mutex> lk(mu);
  JobGraph*& j = (*graphs)[device];
  if (!j) j = new JobGraph;
  return *j;
}
}  // namespace

// The band kernel for a dense-head mode; HM: the split p > 0 (prefix-norm
// filter and forward-list head dot) is compiled in; ST: work counters.
static void launch_band_scan(int32_t dm, bool hm, bool st, int g, hipStream_t sm, const ScanArgs& a) {
  void* args[] = {const_cast<ScanArgs*>(&a)};
  APSS_HIP(hipLaunchKernel(band_scan_fn(dm, hm, st), dim3(g), dim3(SCAN_THREADS), args, 0, sm));
}

ScanOutput scan_and_rescore(Arena& ar, hipStream_t st, const Side& I, const TileSet& bands,
                            const Side& J, const TileSet& tiles, const Tiers& tr,
                            const RunConfig& rc) {
  ScanOutput so;
  so.n_dense = tr.n_dense;
  if (I.n == 0 || J.n == 0 || bands.T == 0 || tiles.T == 0) return so;  // no pair cells
  const int64_t maxlen = std::max(I.maxlen, J.maxlen);
  // (a side whose values are all zero is exact at any scale: use norm 1)
  const double mI = I.maxnorm > 0.f ? (double)I.maxnorm : 1.0;
  const double mJ = J.maxnorm > 0.f ? (double)J.maxnorm : 1.0;
  // fixed-point scale 2^S with 2^S * max|x||y| <= 2^30 (partial sums < 2^31)
  int S = 30 - (int)std::ceil(std::log2(mI * mJ));
  S = std::max(-100, std::min(S, 100));
  const double qscale = std::ldexp(1.0, S), qinv = std::ldexp(1.0, -S);
  const double head_eps = tr.p > 0 ? gamma32(maxlen + 1) * mI * mJ : 0.0;
  // dense head bounds.  int8: quantisation |x - X 2^-B| <= 2^-(B+1) per value,
  // summed with ||x_H||_1 <= 8 ||x||; the dropped limb groups (weights 2^8
  // and 2^0, <= 64 * (2 * 2^14 * 2^8 + 2^14)); one fixed-point step for the
  // truncation of the combined value.  BF16: RNE per value (u = 2^-8, so
  // 2^-7 + 2^-16 per product), FP32 accumulation of 128 exact products
  // (gamma_128, doubled as a margin for the MFMA summation order), one step.
  int32_t dsh_hi = 0, dsh_lo = 0;
  double eps_dense = 0.0, c_bf = 0.0, eps_dense_max = 0.0;
  if (tr.dm == DENSE_I8) {
    const int E = I.dense_B + J.dense_B - S;  // the int8 value v has weight 2^-(BI+BJ); the tile 2^-S
    if (E < 16 || E > 32)
      throw Error{APSS_E_UNSUPPORTED, "values too small for the int8 MFMA head (set mfma_head = -1)"};
    dsh_hi = 32 - E;
    dsh_lo = E - 16;
    const double bI = std::ldexp(1.0, -I.dense_B), bJ = std::ldexp(1.0, -J.dense_B);
    eps_dense = (8.0 * mI * 0.5 * bJ + 8.0 * mJ * 0.5 * bI + 64.0 * 0.25 * bI * bJ +
                 64.0 * (std::ldexp(1.0, 23) + std::ldexp(1.0, 14)) * bI * bJ + qinv) * (1.0 + 1e-9);
    eps_dense_max = eps_dense;
  } else if (tr.dm == DENSE_BF16) {
    c_bf = (std::ldexp(1.0, -7) + std::ldexp(1.0, -16)) +
           2.0 * gamma32(DENSE_BF16_FEATURES) * (1.0 + std::ldexp(1.0, -7) + std::ldexp(1.0, -16));
    eps_dense_max = c_bf * mI * mJ + qinv;
  }
  const double eps_max = cell_error_bound((int)maxlen, mI, mJ, qinv, head_eps) + eps_dense_max;
  // fast-reject bound: the smallest threshold any band uses
  const double t_min = rc.band_t ? rc.t_min : rc.t;
  const double qlo = std::floor((t_min - 2.0 * eps_max) * qscale);
  const int32_t q_lo = (int32_t)std::max(-2147483647.0, std::min(2147483647.0, qlo));

  const double cells = rc.self ? 0.5 * (double)I.n * (double)(I.n - 1) : (double)I.n * (double)J.n;
  // counters: [0] output pairs [1] band pairs [2..8] and [10] kernel statistics [9] work-item
  // queue
  unsigned long long* ctr = ar.alloc<unsigned long long>(16);

  // Work items: band b against a chunk of its partner tiles [J_first, J_end)
  // (self-join: the first tile whose rows end after the band's first row).
  // Chunks are sized so the items outnumber the workgroups several times
  // over, which spreads a query batch of a few bands over the whole chip.
  int n_cu = 0, dev = 0;
  APSS_HIP(hipGetDevice(&dev));
  APSS_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev));
  n_cu = std::max(n_cu, 1);
  std::vector<int32_t> jf((size_t)std::max(bands.T, 0)), je((size_t)std::max(bands.T, 0));
  int64_t partner_tiles = 0;
  for (int32_t b = 0; b < bands.T; ++b) {
    int32_t f = 0;
    if (rc.self)
      f = (int32_t)(std::upper_bound(tiles.cuts.begin() + 1, tiles.cuts.end(), bands.cuts[b]) -
                    (tiles.cuts.begin() + 1));
    const int32_t e = rc.band_jend.empty() ? tiles.T : rc.band_jend[b];
    if (e > tiles.T || e < 0)  // the kernel would read tiles that do not exist
      throw Error{APSS_E_INVALID, "partner tile range past the tiling"};
    jf[b] = f;
    je[b] = e;
    partner_tiles += std::max(0, e - f);
  }
  // Pair cells, dense head rows and tile pairs of the launch, from the work
  // items: cells of rows [i0, i1) x [j0, j1), only i < j for a self-join.
  auto cells_of = [&](int64_t i0, int64_t i1, int64_t j0, int64_t j1) -> int64_t {
    if (i1 <= i0 || j1 <= j0) return 0;
    if (!rc.self) return (i1 - i0) * (j1 - j0);
    int64_t c = std::max<int64_t>(0, std::min(i1, j0) - i0) * (j1 - j0);  // rows below j0
    const int64_t lo = std::max(i0, j0), hi = std::min(i1, j1) - 1;         // rows inside
    if (hi >= lo) c += (hi - lo + 1) * (j1 - 1) - (lo + hi) * (hi - lo + 1) / 2;
    return c;
  };
  // items of every stride-th band; of each band's chunks, every `every`-th
  // (starting at an offset that turns with the band); returns the tile pairs
  auto cut_items = [&](int32_t chunk, int32_t stride, int32_t every, std::vector<int4>& v) {
    int64_t tp = 0;
    for (int32_t b = 0; b < bands.T; b += stride) {
      int32_t k = (b / stride) % every;
      for (int32_t j = jf[b] + k * chunk; j < je[b]; j += every * chunk) {
        v.push_back(make_int4(b, j, std::min(je[b], j + chunk), 0));
        tp += std::min(je[b], j + chunk) - j;
      }
    }
    return tp;
  };
  const int64_t want_items = (int64_t)std::max(1, rc.items_per_cu) * (int64_t)n_cu;
  const int32_t chunk = (int32_t)std::max<int64_t>(
      16, std::min<int64_t>(4096, (partner_tiles + want_items - 1) / want_items));
  std::vector<int4> hitems;
  cut_items(chunk, 1, 1, hitems);
  // Items of sparse bands (rows with cursor-tier features only; see the SP
  // form of k_band_scan) go to the end of the list and run in their own launch.
  // All thresholds must be above zero: then a cell without products holds no
  // result.
  const bool sparse_ok = q_lo > 0 && !bands.rare.empty();
  auto order_items = [&](std::vector<int4>& v) -> int32_t {  // returns the dense items' count
    if (!sparse_ok) return (int32_t)v.size();
    const auto mid = std::stable_partition(v.begin(), v.end(),
                                           [&](const int4& it) { return !bands.rare[(size_t)it.x]; });
    return (int32_t)(mid - v.begin());
  };
  int32_t n_dense_items = order_items(hitems);
  const int32_t n_items = (int32_t)hitems.size();
  const int32_t n_dense_main = n_dense_items;
  int64_t host_cells = 0, host_dense = 0;
  for (const int4& it : hitems) {
    const int64_t i0 = bands.cuts[it.x], i1 = bands.cuts[it.x + 1];
    const int64_t j0 = tiles.cuts[it.y], j1 = tiles.cuts[it.z];
    host_cells += cells_of(i0, i1, j0, j1);
    if (tr.dm != DENSE_OFF) host_dense += (j1 - j0) + (i1 - i0);
  }
  int4* items = ar.alloc<int4>(std::max(n_items, 1));
  if (n_items > 0)
    APSS_HIP(hipMemcpyAsync(items, hitems.data(), sizeof(int4) * n_items, hipMemcpyHostToDevice, st));
  // settle bounds of the rows of each side (one array when both sides are the
  // same matrix, as in a self-join)
  RowBoundArgs rb{};
  rb.qinv = qinv;
  rb.qscale = qscale;
  rb.e_fixed = tr.dm == DENSE_I8 ? eps_dense : tr.dm == DENSE_BF16 ? qinv : 0.0;
  // emitted FP32 scores within 1e-6 relative: rel_k >= (1 + 2^-24) / (1e-6 - 2^-24)
  // covers the bound and the final rounding to FP32
  rb.rel_k = tr.dm == DENSE_BF16 ? 0.0 : 1.07e6;
  rb.head = tr.p > 0 ? 1 : 0;
  const bool one_side = I.rp == J.rp && I.n == J.n && I.norm == J.norm;
  int2* rqI = ar.alloc<int2>(I.n);
  int2* rqJ = one_side ? rqI : ar.alloc<int2>(J.n);
  rb.other_norm = mJ;
  hipLaunchKernelGGL(k_row_bounds, dim3(blocks_for(I.n, 256, 4096)), dim3(256), 0, st, I.n, I.rp, I.norm,
                     I.hnorm, rb, rqI);
  if (!one_side) {
    rb.other_norm = mI;
    hipLaunchKernelGGL(k_row_bounds, dim3(blocks_for(J.n, 256, 4096)), dim3(256), 0, st, J.n, J.rp,
                       J.norm, J.hnorm, rb, rqJ);
  }
  APSS_HIP(hipGetLastError());
  // persistent grid: one 1024-thread workgroup per CU pulls work items from a
  // queue; each owns a cursor-state region (records of one band) and a
  // cursor-product region large enough for every product one tile pair can
  // receive (records x tile rows)
  const int64_t scs_cap = std::max<int64_t>(bands.max_s, 1);
  const int64_t cprod_cap =
      bands.nS > 0 ? std::max<int64_t>(rc.cprod_floor, bands.max_s * TILE_ROWS) : 0;
  // error-band list of a caller-owned output (no estimate pass sizes it)
  int64_t bcap_given = 0;
  if (rc.out_given)
    bcap_given = std::max<int64_t>(1 << 16, std::min<int64_t>(1 << 22, rc.out_given_cap / 16));
  int grid = std::max(1, std::min(n_items, n_cu));
  if (ar.bump()) {
    // a fixed scratch array: as many workgroups as their regions fit beside
    // the band list (fewer workgroups only lengthen the persistent loop)
    const int64_t per_wg = scs_cap * (int64_t)sizeof(int4) + cprod_cap * (int64_t)sizeof(int2) + 512;
    const int64_t left = ar.room() - 8 * bcap_given - ((int64_t)1 << 16);
    if (left < per_wg)
      throw Error{APSS_E_NOMEM, "input needs more than the scratch array (scan regions)"};
    grid = (int)std::min<int64_t>(grid, left / per_wg);
  }
  int4* scs = ar.alloc<int4>(scs_cap * grid);
  int2* cprod = ar.alloc<int2>(cprod_cap * grid);
  // the launch's scratch goes back to the arena as soon as the scan is done
  // (a caller running several launches in one arena, such as a query over
  // several corpus segments, then holds one launch's scratch at a time)
  auto release_scratch = [&]() {
    ar.release(cprod);
    ar.release(scs);
    if (rqJ != rqI) ar.release(rqJ);
    ar.release(rqI);
    ar.release(items);
    ar.release(ctr);
  };

  ScanArgs a{};
  a.I = make_view(I, bands);
  a.J = make_view(J, tiles);
  a.I.rq = rqI;
  a.J.rq = rqJ;
  a.S.ptr = tr.sptr;
  a.S.val = tr.sval;
  a.S.pos = tr.spos;
  a.KD = tr.KD;
  a.KW = tr.KW;
  a.p = tr.p;
  a.self = rc.self;
  a.head_mode = tr.p > 0 ? 1 : 0;
  a.t = rc.t;
  a.qscale = (float)qscale;
  a.qinv_f = (float)qinv;
  a.cbf_q = std::nextafter((float)(c_bf * qscale), INFINITY);
  a.cabs32 = (uint32_t)std::ceil(std::ldexp(c_bf / (1.0 - c_bf), 32) * (1.0 + 1e-12));
  a.bf_nonneg = (I.nonneg && J.nonneg) ? 1 : 0;
  a.qinv = qinv;
  a.q_lo = q_lo;
  a.score_rel = 1e-6;  // (the kernel applies it through the rows' Q bounds)
  a.i_off = rc.i_off;
  a.j_off = rc.j_off;
  a.out_cnt = ctr + 0;
  a.band_cnt = ctr + 1;
  a.rescore_all = rc.rescore_all;
  a.band_t = rc.band_t;
  a.row_t = rc.row_t;
  a.eps_max = eps_max;
  a.stats = ctr + 2;
  a.scs = scs;
  a.scs_cap = scs_cap;
  a.cprod = cprod;
  a.cprod_cap = cprod_cap;
  a.item_next = ctr + 9;
  a.dsh_hi = dsh_hi;
  a.dsh_lo = dsh_lo;
  a.eps_dense = eps_dense;
  a.c_bf = c_bf;
  // BF16 head: emitted scores carry the stated BF16 bound (pair set exact)
  if (tr.dm == DENSE_BF16) a.score_rel = 1e300;
  // the two launches of a list of a.n_items items, the first n_dense_items of
  // them for the band kernel of the launch's mode, the rest for the sparse form
  auto split_args = [&](ScanArgs& d, ScanArgs& sp) {
    d = a;
    d.n_items = n_dense_items;
    sp = a;
    sp.items = a.items + n_dense_items;
    sp.n_items = a.n_items - n_dense_items;
    sp.item_next = ctr + 13;  // its own queue
  };
  auto launch = [&](int g) {
    fill_u64(st, ctr, 16, 0);
    if (a.n_items <= 0) return;
    ScanArgs d, sp;
    split_args(d, sp);
    if (d.n_items > 0) {
      const int gd = std::max(1, std::min(g, d.n_items));
      launch_band_scan(tr.dm, a.head_mode != 0, rc.count_work, gd, st, d);
    }
    if (sp.n_items > 0)
      hipLaunchKernelGGL((k_band_scan<DENSE_OFF, false, true, false>), dim3(std::max(1, std::min(g, sp.n_items))),
                         dim3(SCAN_THREADS), 0, st, sp);
    APSS_HIP(hipGetLastError());
  };
  PhaseTrace ph;

  // Output capacity.  A small pair space gets room for every cell.  Otherwise
  // a count-only launch (no output room, so the kernel only counts) over every
  // stride-th band against all its partner tiles, cut in short chunks so the
  // sample fills the chip, estimates the result and band-list sizes, scaled
  // by tile pairs; a launch that still overflows is rerun at the exact size.
  size_t free_b = 0, total_b = 0;
  APSS_HIP(hipMemGetInfo(&free_b, &total_b));
  const double room = (double)free_b * 0.4 / sizeof(apss_pair);
  int64_t cap, bcap;
  if (rc.out_given) {
    cap = rc.out_given_cap;
    bcap = bcap_given;
  } else if (rc.out_cap_hint > 0) {
    cap = (int64_t)std::min<double>(cells, (double)rc.out_cap_hint);
    bcap = std::max<int64_t>(1 << 20, cap / 64);
  } else if (rc.tune_only) {
    // the autotuner's measurement: a count-only launch (no output room) over
    // a sample of tune_bands bands with all their partner tiles, or over the
    // launch's own work items; the faster of two runs, scaled by tile pairs
    std::vector<int4> sitems;
    int64_t stp = partner_tiles;
    if (rc.tune_bands > 0) stp = cut_items(64, std::max<int32_t>(1, bands.T / rc.tune_bands), 1, sitems);
    else sitems = hitems;
    n_dense_items = order_items(sitems);
    int4* sdev = ar.alloc<int4>((int64_t)std::max<size_t>(sitems.size(), 1));
    APSS_HIP(hipMemcpyAsync(sdev, sitems.data(), sizeof(int4) * sitems.size(), hipMemcpyHostToDevice, st));
    a.items = sdev;
    a.n_items = (int32_t)sitems.size();
    a.out = nullptr;
    a.out_cap = 0;
    a.band = nullptr;
    a.band_cap = 0;
    double best_ms = INFINITY;
    for (int rep = 0; rep < 2; ++rep) {
      Timer ts(st);
      ts.start();
      launch(std::max(1, std::min(a.n_items, grid)));
      best_ms = std::min(best_ms, ts.stop_ms());
    }
    ar.release(sdev);
    so.ms_scan = best_ms * (stp > 0 ? (double)partner_tiles / (double)stp : 1.0);
    so.tile_pairs = partner_tiles;
    release_scratch();
    return so;
  } else if (cells <= (double)(1 << 24) || bands.T < 64) {
    cap = (int64_t)std::min(cells, room);
    bcap = std::max<int64_t>(1 << 20, std::min<int64_t>(cap, 1 << 24));
  } else {
    // count-only launch over a sample of the bands: every stride-th band
    // against every eighth 64-tile chunk of its partner tiles
    const int32_t stride = std::max<int32_t>(1, bands.T / 128);
    std::vector<int4> sitems;
    const int64_t stp = cut_items(64, stride, 8, sitems);
    n_dense_items = order_items(sitems);
    int4* sdev = ar.alloc<int4>((int64_t)std::max<size_t>(sitems.size(), 1));
    APSS_HIP(hipMemcpyAsync(sdev, sitems.data(), sizeof(int4) * sitems.size(), hipMemcpyHostToDevice, st));
    a.items = sdev;
    a.n_items = (int32_t)sitems.size();
    a.out = nullptr;
    a.out_cap = 0;
    a.band = nullptr;
    a.band_cap = 0;
    Timer ts(st);
    ts.start();
    launch(std::max(1, std::min(a.n_items, n_cu)));
    const double sample_ms = ts.stop_ms();
    unsigned long long h2[2];
    APSS_HIP(hipMemcpyAsync(h2, ctr, sizeof(h2), hipMemcpyDeviceToHost, st));
    APSS_HIP(hipStreamSynchronize(st));
    ar.release(sdev);
    (void)sample_ms;
    const double scale = stp > 0 ? (double)partner_tiles / (double)stp : 1.0;
    cap = (int64_t)std::min(cells, std::min(room, (double)h2[0] * scale * 1.25 + (double)(1 << 22)));
    bcap = (int64_t)((double)h2[1] * scale * 1.25) + (1 << 20);
    if (trace_on())
      std::fprintf(stderr, "[apss] estimate: %.4g pairs, %.4g band pairs (sample of %lld tile pairs)\n",
                   (double)h2[0] * scale, (double)h2[1] * scale, (long long)stp);
    ph.mark("output estimate");
  }
  if (!rc.out_given) cap = std::max<int64_t>(cap, 1 << 16);
  a.items = items;
  a.n_items = n_items;
  n_dense_items = n_dense_main;

  ph.mark("scan set-up");
  // The device job: clear the counters, band scan, FP64 rescore of the band
  // list (sized on the device).  Eagerly with events around the two kernels,
  // or captured once as a hipGraph and launched as one unit.
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  struct EvGuard {
    hipEvent_t* e;
    ~EvGuard() { for (int k = 0; k < 3; ++k) if (e[k]) (void)hipEventDestroy(e[k]); }
  } evg{ev};
  for (int k = 0; k < 3; ++k) APSS_HIP(hipEventCreate(&ev[k]));
  const int rs_grid = std::max(1, std::min(4 * n_cu * 4, 1 << 14));
  auto enqueue = [&](apss_pair* out, int64_t ocap, unsigned long long* band, int64_t bc,
                     bool timed) {
    if (timed) APSS_HIP(hipEventRecord(ev[0], st));
    launch(grid);
    if (timed) APSS_HIP(hipEventRecord(ev[1], st));
    hipLaunchKernelGGL(k_rescore, dim3(rs_grid), dim3(256), 0, st, band, ctr + 1, bc, I.rp, I.col,
                       I.val, I.nnz, J.rp, J.col, J.val, J.nnz, rc.t, rc.row_t, rc.i_off, rc.j_off,
                       out, ctr + 0, ocap);
    APSS_HIP(hipGetLastError());
    if (timed) APSS_HIP(hipEventRecord(ev[2], st));
  };
  for (int attempt = 0; attempt < 3; ++attempt) {
    apss_pair* out = rc.out_given ? rc.out_given : ar.alloc<apss_pair>(cap);
    unsigned long long* band = ar.alloc<unsigned long long>(bcap);
    a.out = out;
    a.out_cap = cap;
    a.band = band;
    a.band_cap = bcap;
    if (rc.use_graph) {
      // the job as one graph launch: clear the counters, scan, rescore
      unsigned long long* c_p = ctr;
      int64_t c_n = 16;
      unsigned long long c_v = 0;
      void* fill_args[] = {&c_p, &c_n, &c_v};
      ScanArgs d, sp;
      split_args(d, sp);
      void* scan_args[] = {&d};
      void* sparse_args[] = {&sp};
      const unsigned long long* r_band = band;
      const unsigned long long* r_cnt = ctr + 1;
      int64_t r_bcap = bcap;
      const int64_t* r_rpI = I.rp;
      const int32_t* r_colI = I.col;
      const float* r_valI = I.val;
      int64_t r_nnzI = I.nnz;
      const int64_t* r_rpJ = J.rp;
      const int32_t* r_colJ = J.col;
      const float* r_valJ = J.val;
      int64_t r_nnzJ = J.nnz;
      double r_t = rc.t;
      const double* r_rowt = rc.row_t;
      int64_t r_ioff = rc.i_off, r_joff = rc.j_off;
      apss_pair* r_out = out;
      unsigned long long* r_ocnt = ctr + 0;
      int64_t r_ocap = cap;
      void* rs_args[] = {&r_band, &r_cnt,  &r_bcap, &r_rpI, &r_colI, &r_valI,
                         &r_nnzI, &r_rpJ,  &r_colJ, &r_valJ, &r_nnzJ, &r_t,
                         &r_rowt, &r_ioff, &r_joff, &r_out,  &r_ocnt, &r_ocap};
      // (a scan node with no items exits at once)
      hipKernelNodeParams np[4] = {};
      np[0].func = reinterpret_cast<void*>(&k_fill_u64);
      np[0].gridDim = dim3(1);
      np[0].blockDim = dim3(256);
      np[0].kernelParams = fill_args;
      np[1].func = band_scan_fn(tr.dm, a.head_mode != 0, rc.count_work);
      np[1].gridDim = dim3(std::max(1, std::min(grid, d.n_items)));
      np[1].blockDim = dim3(SCAN_THREADS);
      np[1].kernelParams = scan_args;
      np[2].func = reinterpret_cast<void*>(&k_band_scan<DENSE_OFF, false, true, false>);
      np[2].gridDim = dim3(std::max(1, std::min(grid, sp.n_items)));
      np[2].blockDim = dim3(SCAN_THREADS);
      np[2].kernelParams = sparse_args;
      np[3].func = reinterpret_cast<void*>(&k_rescore);
      np[3].gridDim = dim3(rs_grid);
      np[3].blockDim = dim3(256);
      np[3].kernelParams = rs_args;
      JobGraph& jg = job_graph(dev);
      std::lock_guard<std::mutex> lk(jg.mu);  // held until the job has finished
      so.graph_reused = jg.exec && jg.scan_fn == np[1].func;
      for (int k = 0; k < 4 && so.graph_reused; ++k)
        if (hipGraphExecKernelNodeSetParams(jg.exec, jg.node[k], &np[k]) != hipSuccess) {
          (void)hipGetLastError();
          so.graph_reused = false;
        }
      if (!so.graph_reused) jg.build(np);
      APSS_HIP(hipEventRecord(ev[0], st));
      APSS_HIP(hipGraphLaunch(jg.exec, st));
      APSS_HIP(hipEventRecord(ev[2], st));
      APSS_HIP(hipEventSynchronize(ev[2]));
      float ms = 0.f;
      APSS_HIP(hipEventElapsedTime(&ms, ev[0], ev[2]));
      so.ms_scan = ms;  // the whole graph (scan + rescore)
      so.ms_rescore = 0.0;
    } else {
      enqueue(out, cap, band, bcap, true);
      APSS_HIP(hipEventSynchronize(ev[2]));
      float m1 = 0.f, m2 = 0.f;
      APSS_HIP(hipEventElapsedTime(&m1, ev[0], ev[1]));
      APSS_HIP(hipEventElapsedTime(&m2, ev[1], ev[2]));
      so.ms_scan = m1;
      so.ms_rescore = m2;
    }
    ph.mark("band scan + rescore");
    unsigned long long h[16];
    APSS_HIP(hipMemcpyAsync(h, ctr, sizeof(h), hipMemcpyDeviceToHost, st));
    APSS_HIP(hipStreamSynchronize(st));
    const int64_t nout = (int64_t)h[0], nband = (int64_t)h[1];
    if (trace_on())
      std::fprintf(stderr, "[apss] scan: %lld pairs, %lld band pairs (room %lld, %lld)\n",
                   (long long)nout, (long long)nband, (long long)cap, (long long)bcap);
    if (rc.out_given && nout > cap && nband <= bcap) {
      // a caller-owned buffer is never regrown: the count tells the caller
      // how much room the result needs (the buffer holds `cap` of its pairs)
      so.count = nout;
      so.pairs = out;
      ar.release(band);
      release_scratch();
      return so;
    }
    if (nout > cap || nband > bcap) {
      // capacity too small: every count is exact except the rescore's share of
      // a truncated band list; grow and run again
      if (!rc.out_given) ar.release(out);
      ar.release(band);
      if (!rc.out_given) cap = std::max(cap, nout + std::max<int64_t>(0, nband - bcap) + 1024);
      bcap = std::max(bcap, nband + 1024);
      continue;
    }
    // without count_work the bitmap-tier visits and products are not counted
    so.products = rc.count_work ? (int64_t)h[2] : -1;
    so.head_pairs = (int64_t)h[3];
    so.cells = host_cells;
    so.visits = rc.count_work ? (int64_t)h[5] : -1;
    so.cursor = (int64_t)h[6];
    so.dense_rows = host_dense;
    so.tile_pairs = partner_tiles;
    so.band_entries = (int64_t)h[4];
    so.band_records = (int64_t)h[7];
    so.walk_entries = (int64_t)h[8];
    so.walk_scatter = (int64_t)h[10];
    so.seg_row_bytes = 2 * (int64_t)tiles.kds;
    so.skip_no_common = (int64_t)h[2 + STAT_SKIP_NO_COMMON];
    so.skip_bound = (int64_t)h[2 + STAT_SKIP_BOUND];
    so.band = nband;
    so.n_dense = tr.n_dense;
    so.count = nout;
    so.pairs = out;
    ar.release(band);
    release_scratch();  // the counters were read back above: the stream is idle
    return so;
  }
  throw Error{APSS_E_NOMEM, "output capacity could not be sized"};
}

// ---------------------------------------------------------------------------
// run_self

void copy_model_counters(const ScanOutput& so, apss_result* out) {
  out->band_entries += so.band_entries;
  out->band_records += so.band_records;
  out->walk_entries += so.walk_entries;
  out->walk_scatter_entries += so.walk_scatter;
  out->seg_row_bytes = so.seg_row_bytes;
  out->tile_pairs_skipped += so.skip_no_common;
  out->tile_pairs_pruned += so.skip_bound;
}

void sort_pairs(apss_pair* p, int64_t n) {
  std::sort(p, p + n, [](const apss_pair& x, const apss_pair& y) {
    return x.i != y.i ? x.i < y.i : x.j < y.j;
  });
}

// ---------------------------------------------------------------------------
// host result blocks

namespace {
constexpr size_t PINNED_MIN = (size_t)1 << 20;      // smaller results use malloc
constexpr size_t PINNED_MAX = (size_t)2 << 30;      // larger ones too (pinning is slow)
constexpr size_t PINNED_POOL_MAX = (size_t)8 << 30;  // pinned bytes kept for reuse

struct PinnedPool {
  std::mutex mu;
  std::unordered_map<void*, size_t> live;  // pinned blocks handed out, by size
  std::multimap<size_t, void*> idle;        // pinned blocks ready for reuse
  size_t idle_bytes = 0;
};
PinnedPool& pinned_pool() {
  static PinnedPool* pp = new PinnedPool;  // outlives every static destructor
  return *pp;
}
// Sizes are rounded up to 1/8 steps of a power of two, so blocks get reused.
size_t pinned_size(size_t bytes) {
  size_t pw = PINNED_MIN;
  while (pw * 2 <= bytes) pw *= 2;
  const size_t step = pw / 8;
  return (bytes + step - 1) / step * step;
}
}  // namespace

apss_pair* host_pairs_alloc(int64_t count) {
  const size_t bytes = sizeof(apss_pair) * (size_t)std::max<int64_t>(count, 1);
  if (bytes >= PINNED_MIN && bytes <= PINNED_MAX) {
    PinnedPool& pp = pinned_pool();
    const size_t want = pinned_size(bytes);
    {
      std::lock_guard<std::mutex> lk(pp.mu);
      auto it = pp.idle.lower_bound(want);
      if (it != pp.idle.end() && it->first <= 2 * want) {
        void* q = it->second;
        pp.live[q] = it->first;
        pp.idle_bytes -= it->first;
        pp.idle.erase(it);
        return (apss_pair*)q;
      }
    }
    void* q = nullptr;
    if (hipHostMalloc(&q, want, hipHostMallocDefault) == hipSuccess && q) {
      std::lock_guard<std::mutex> lk(pp.mu);
      pp.live[q] = want;
      return (apss_pair*)q;
    }
    (void)hipGetLastError();  // no pinned memory: fall back to pageable
  }
  apss_pair* q = nullptr;
  if (bytes > PINNED_MAX) {
    // 2 MiB-aligned and backed by huge pages where the kernel allows it: the
    // first touch of a multi-GB result then faults once per 2 MiB, not per 4 KiB
    const size_t huge = (size_t)2 << 20;
    q = (apss_pair*)std::aligned_alloc(huge, (bytes + huge - 1) / huge * huge);
    if (q) (void)madvise(q, (bytes + huge - 1) / huge * huge, MADV_HUGEPAGE);
  } else {
    q = (apss_pair*)std::malloc(bytes);
  }
  if (!q) throw Error{APSS_E_NOMEM, "host allocation of the result failed"};
  return q;
}

namespace {
bool host_pairs_pinned(const void* p) {
  PinnedPool& pp = pinned_pool();
  std::lock_guard<std::mutex> lk(pp.mu);
  return pp.live.count(const_cast<void*>(p)) != 0;
}

// memcpy of `pieces` (block offset, source, bytes) into `blk`, the bytes split
// evenly over up to eight host threads
struct Piece {
  size_t at;
  const char* src;
  size_t bytes;
};
template <class Work>
void on_threads(size_t total, Work work) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int nt = total < ((size_t)4 << 20) ? 1 : (int)std::min(8u, hw);
  auto part = [&](int t) {
    work(total * (size_t)t / (size_t)nt, total * (size_t)(t + 1) / (size_t)nt);
  };
  std::vector<std::thread> th;
  int t = 1;
  try {
    for (; t < nt; ++t) th.emplace_back(part, t);
  } catch (...) {  // no thread to spare: this thread copies the rest
    for (; t < nt; ++t) part(t);
  }
  part(0);
  for (auto& x : th) x.join();
}
void fill_block(char* blk, const std::vector<Piece>& pieces, size_t total) {
  on_threads(total, [&](size_t lo, size_t hi) {
    for (const Piece& q : pieces) {
      const size_t a = std::max(lo, q.at), b = std::min(hi, q.at + q.bytes);
      if (a < b) std::memcpy(blk + a, q.src + (a - q.at), b - a);
    }
  });
}

constexpr size_t STAGED_MIN = (size_t)16 << 20;  // smaller copies go direct
constexpr size_t CHUNK = (size_t)32 << 20;       // bytes per staging block

// two pinned staging blocks and their DMA events
struct Stage {
  apss_pair* blk[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool ok = false;  // both blocks pinned
  Stage() {
    const int64_t slots = (int64_t)(CHUNK / sizeof(apss_pair)) + 1;
    for (int b = 0; b < 2; ++b) blk[b] = host_pairs_alloc(slots);
    ok = host_pairs_pinned(blk[0]) && host_pairs_pinned(blk[1]);
    for (int b = 0; b < 2 && ok; ++b)
      if (hipEventCreateWithFlags(&ev[b], hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        ev[b] = nullptr;
        ok = false;  // the copies go direct
      }
  }
  ~Stage() {
    for (int b = 0; b < 2; ++b) {
      if (ev[b]) {
        (void)hipEventSynchronize(ev[b]);  // no DMA still uses the block
        (void)hipEventDestroy(ev[b]);
      }
      host_pairs_free(blk[b]);
    }
  }
};
}  // namespace

void h2d_copies(const std::vector<H2DCopy>& list, hipStream_t st) {
  size_t total = 0;
  for (const H2DCopy& c : list) total += c.bytes;
  auto direct = [&]() {
    for (const H2DCopy& c : list)
      if (c.bytes) APSS_HIP(hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyHostToDevice, st));
  };
  if (total < STAGED_MIN) return direct();
  Stage sg;
  if (!sg.ok) return direct();
  // the list as one byte stream, cut in CHUNK-sized blocks
  size_t li = 0, lo = 0;  // current copy and offset inside it
  for (int k = 0; li < list.size(); ++k) {
    const int b = k & 1;
    if (k >= 2) APSS_HIP(hipEventSynchronize(sg.ev[b]));  // block b's last DMA is done
    std::vector<Piece> pieces;
    std::vector<H2DCopy> dmas;
    size_t fill = 0;
    while (li < list.size() && fill < CHUNK) {
      const H2DCopy& c = list[li];
      const size_t n = std::min(c.bytes - lo, CHUNK - fill);
      if (n) {
        pieces.push_back({fill, (const char*)c.src + lo, n});
        dmas.push_back({(char*)c.dst + lo, (const char*)sg.blk[b] + fill, n});
      }
      fill += n;
      lo += n;
      if (lo == c.bytes) { ++li; lo = 0; }
    }
    fill_block((char*)sg.blk[b], pieces, fill);
    for (const H2DCopy& d : dmas)
      APSS_HIP(hipMemcpyAsync(d.dst, d.src, d.bytes, hipMemcpyHostToDevice, st));
    APSS_HIP(hipEventRecord(sg.ev[b], st));
  }
}

void d2h_copy(void* dst, const void* src, size_t bytes, hipStream_t st) {
  if (bytes < STAGED_MIN || host_pairs_pinned(dst)) {
    if (bytes) APSS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
    APSS_HIP(hipStreamSynchronize(st));
    return;
  }
  Stage sg;
  if (!sg.ok) {
    APSS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
    APSS_HIP(hipStreamSynchronize(st));
    return;
  }
  const size_t nch = (bytes + CHUNK - 1) / CHUNK;
  auto issue = [&](size_t k) {  // chunk k into block k % 2
    const size_t off = k * CHUNK, len = std::min(CHUNK, bytes - off);
    APSS_HIP(hipMemcpyAsync(sg.blk[k & 1], (const char*)src + off, len, hipMemcpyDeviceToHost, st));
    APSS_HIP(hipEventRecord(sg.ev[k & 1], st));
  };
  issue(0);
  for (size_t k = 0; k < nch; ++k) {
    // block (k+1) % 2 was emptied in the previous step
    if (k + 1 < nch) issue(k + 1);
    APSS_HIP(hipEventSynchronize(sg.ev[k & 1]));
    const size_t off = k * CHUNK, len = std::min(CHUNK, bytes - off);
    const char* from = (const char*)sg.blk[k & 1];
    char* to = (char*)dst + off;
    on_threads(len, [&](size_t lo, size_t hi) { std::memcpy(to + lo, from + lo, hi - lo); });
  }
}

apss_pair* host_pairs_grow(apss_pair* p, int64_t used, int64_t count) {
  if (!p) return host_pairs_alloc(count);
  size_t have = 0;
  {
    PinnedPool& pp = pinned_pool();
    std::lock_guard<std::mutex> lk(pp.mu);
    auto it = pp.live.find(p);
    if (it != pp.live.end()) have = it->second;
  }
  const size_t bytes = sizeof(apss_pair) * (size_t)std::max<int64_t>(count, 1);
  if (have >= bytes) return p;
  apss_pair* q = host_pairs_alloc(count);
  if (used > 0) std::memcpy(q, p, sizeof(apss_pair) * (size_t)used);
  host_pairs_free(p);
  return q;
}

void host_pairs_free(apss_pair* p) {
  if (!p) return;
  PinnedPool& pp = pinned_pool();
  std::unique_lock<std::mutex> lk(pp.mu);
  auto it = pp.live.find(p);
  if (it == pp.live.end()) {
    lk.unlock();
    std::free(p);
    return;
  }
  const size_t sz = it->second;
  pp.live.erase(it);
  if (pp.idle_bytes + sz <= PINNED_POOL_MAX) {
    pp.idle.emplace(sz, (void*)p);
    pp.idle_bytes += sz;
    return;
  }
  lk.unlock();
  (void)hipHostFree(p);
}

int run_self(const apss_csr* X, double t, int32_t p, const apss_opts* o, apss_result* out) {
  apss_opts def;
  apss_opts_init(&def);
  if (!o) o = &def;
  if (!X || !out) throw Error{APSS_E_INVALID, "null argument"};
  if (!(t == t) || std::isinf(t)) throw Error{APSS_E_INVALID, "t must be finite"};
  std::memset(out, 0, sizeof(*out));
  check_csr_host(*X, p < 0 ? 0 : p);  // argument errors are reported before any device call
  APSS_HIP(hipSetDevice(o->device));
  hipStream_t st = (hipStream_t)o->stream;
  bool own_stream = false;
  if (!st) {
    APSS_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    own_stream = true;
  }
  struct StreamGuard {
    hipStream_t s;
    bool own;
    ~StreamGuard() { if (own) (void)hipStreamDestroy(s); }
  } sg{st, own_stream};
  auto t0 = std::chrono::steady_clock::now();
  Arena ar(o->device, o->mem_budget, &run_pool(o->device));
  ar.set_stream(st);
  // p < 0: autotune.  The split is p = 0 (every feature on the tiered
  // inverted-list / dense-head path; p > 0 adds the reference's forward-list
  // head dot, which loses on Zipf data); the tuner picks the tiering.
  const bool tune = p < 0;
  if (p < 0) p = 0;
  Timer tm(st);
  PhaseTrace ph;
  tm.start();
  Side s = prepare_side(ar, st, *X, p);
  out->ms_upload = tm.stop_ms();
  ph.mark("upload + validate");
  int32_t dm = dense_mode(*o), kd_cap = o->dense_features;
  TuneChoice tc = choice_from_opts(*o);  // when not tuning
  if (tune) {
    tc = autotune(st, *o, s, 1, {}, t, [&](RunConfig& r, const Built&) {
      r.t = t;
      r.self = 1;
    });
    p = tc.p;
    dm = tc.dm;
    kd_cap = tc.kd_cap;
    if (p > 0) s = with_split(ar, st, s, p);
    out->ms_tune = tc.ms;
    ph.mark("autotune");
  }
  tm.start();
  Built bc = build_config(ar, st, s, p, 1, kd_cap, dm, {}, CAP_S, tc.band_rows, tc.tile_rows);
  Tiers& tr = bc.tr;
  TileSet& bands = bc.bands;
  TileSet& tiles = bc.tiles;
  out->bitmap_features = tr.n_bitmap;
  out->ms_index = tm.stop_ms();
  ph.mark("index");
  RunConfig rc;
  rc.t = t;
  rc.p = p;
  rc.self = 1;
  rc.rescore_all = o->rescore_all;
  rc.items_per_cu = tc.items_per_cu;
  rc.count_work = o->count_work != 0;
  ScanOutput so = scan_and_rescore(ar, st, s, bands, s, tiles, tr, rc);
  out->ms_scan = so.ms_scan;
  out->ms_rescore = so.ms_rescore;
  out->count = so.count;
  out->p_used = p;
  out->cells = so.cells;
  out->band = so.band;
  out->products = so.products;
  out->head_pairs = so.head_pairs;
  out->list_visits = so.visits;
  out->cursor_products = so.cursor;
  out->mfma_features = so.n_dense;
  out->dense_rows = so.dense_rows;
  out->tile_pairs = so.tile_pairs;
  copy_model_counters(so, out);
  out->band_rows = tc.band_rows;
  out->tile_rows = tc.tile_rows;
  out->items_per_cu = tc.items_per_cu;
  tm.start();
  if (o->output_on_device && !o->sort_output) {
    out->pairs = (apss_pair*)ar.detach(so.pairs);
    out->on_device = 1;
  } else {
    apss_pair* h = host_pairs_alloc(so.count);
    try {
      d2h_copy(h, so.pairs, sizeof(apss_pair) * (size_t)so.count, st);
    } catch (...) {
      host_pairs_free(h);
      throw;
    }
    if (o->sort_output) sort_pairs(h, so.count);
    if (o->output_on_device) {
      apss_pair* d = so.pairs;
      APSS_HIP(hipMemcpyAsync(d, h, sizeof(apss_pair) * so.count, hipMemcpyHostToDevice, st));
      APSS_HIP(hipStreamSynchronize(st));
      host_pairs_free(h);
      out->pairs = (apss_pair*)ar.detach(d);
      out->on_device = 1;
    } else {
      out->pairs = h;
      out->on_device = 0;
    }
  }
  out->ms_download = tm.stop_ms();
  ph.mark("result");
  APSS_HIP(hipStreamSynchronize(st));
  out->ms_total =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return APSS_OK;
}

}  // namespace apss

namespace apss {

StreamScope::StreamScope(void* given, int device) {
  APSS_HIP(hipSetDevice(device));
  s = (hipStream_t)given;
  if (!s) {
    APSS_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    own = true;
  }
}
StreamScope::~StreamScope() {
  if (own) (void)hipStreamDestroy(s);
}

void finish_result(hipStream_t st, const apss_pair* dpairs, int64_t count, bool sort,
                   apss_result* out) {
  apss_pair* h = host_pairs_alloc(count);
  try {
    d2h_copy(h, dpairs, sizeof(apss_pair) * (size_t)count, st);
  } catch (...) {
    host_pairs_free(h);
    throw;
  }
  if (sort) sort_pairs(h, count);
  out->pairs = h;
  out->count = count;
  out->on_device = 0;
}

// ---------------------------------------------------------------------------
// batched: every instance's rows are concatenated with tile boundaries forced
// at the instance starts; each band only meets the tiles of its own instance
// (per-band partner range) and uses its instance's threshold, so one band-scan
// launch covers the whole batch.

int run_batched(const apss_csr* const* X, int32_t count, const double* t, int32_t p,
                const apss_opts* o, apss_result* out) {
  apss_opts def;
  apss_opts_init(&def);
  if (!o) o = &def;
  if (count < 0 || (count > 0 && (!X || !t || !out))) throw Error{APSS_E_INVALID, "null argument"};
  const bool tune = p < 0;  // autotune (see run_self)
  if (p < 0) p = 0;
  int64_t n_all = 0, nnz_all = 0, d = 0;
  double t_min = INFINITY;
  for (int32_t k = 0; k < count; ++k) {
    if (!X[k]) throw Error{APSS_E_INVALID, "null instance"};
    if (X[k]->on_device) throw Error{APSS_E_UNSUPPORTED, "batched instances must be host CSR"};
    check_csr_host(*X[k], p);
    if (!(t[k] == t[k]) || std::isinf(t[k])) throw Error{APSS_E_INVALID, "t must be finite"};
    n_all += X[k]->n_rows;
    nnz_all += X[k]->nnz;
    d = std::max(d, X[k]->n_cols);
    t_min = std::min(t_min, t[k]);
    std::memset(&out[k], 0, sizeof(apss_result));
  }
  if (count == 0) return APSS_OK;
  auto t0 = std::chrono::steady_clock::now();
  // One device CSR holds every instance: the rebased row pointers are built on
  // the host, the column ids and values go straight from each instance's
  // arrays into their slice of the device arrays.
  std::vector<int64_t> off((size_t)count + 1, 0), rp((size_t)n_all + 1, 0);
  int64_t r0 = 0, e0 = 0;
  for (int32_t k = 0; k < count; ++k) {
    off[k] = r0;
    const apss_csr& A = *X[k];
    for (int64_t i = 0; i < A.n_rows; ++i) rp[r0 + i + 1] = e0 + A.row_ptr[i + 1];
    r0 += A.n_rows;
    e0 += A.nnz;
  }
  off[count] = r0;
  StreamScope ss(o->stream, o->device);
  hipStream_t st = ss.s;
  Arena ar(o->device, o->mem_budget, &run_pool(o->device));
  ar.set_stream(st);
  Timer tm(st);
  tm.start();
  int64_t* drp = ar.alloc<int64_t>(n_all + 1);
  int32_t* dcol = ar.alloc<int32_t>(nnz_all);
  float* dval = ar.alloc<float>(nnz_all);
  std::vector<H2DCopy> cp;
  cp.push_back({drp, rp.data(), sizeof(int64_t) * (size_t)(n_all + 1)});
  e0 = 0;
  for (int32_t k = 0; k < count; ++k) {
    const apss_csr& A = *X[k];
    if (A.nnz) {
      cp.push_back({dcol + e0, A.col, sizeof(int32_t) * (size_t)A.nnz});
      cp.push_back({dval + e0, A.val, sizeof(float) * (size_t)A.nnz});
    }
    e0 += A.nnz;
  }
  h2d_copies(cp, st);
  apss_csr all{n_all, d, nnz_all, drp, dcol, dval, 1, 0};
  Side s = prepare_side(ar, st, all, p);
  const double ms_upload = tm.stop_ms();
  std::vector<int64_t> bounds(off.begin() + 1, off.end() - 1);
  std::sort(bounds.begin(), bounds.end());
  bounds.erase(std::unique(bounds.begin(), bounds.end()), bounds.end());
  // per band: its instance's partner range end and threshold
  auto band_ranges = [&](const TileSet& bands, const TileSet& tiles, std::vector<int32_t>& jend,
                         std::vector<double>* bt) {
    jend.assign((size_t)std::max(bands.T, 0), 0);
    if (bt) bt->assign((size_t)std::max(bands.T, 1), t_min);
    for (int32_t b = 0; b < bands.T; ++b) {
      const int64_t r = bands.cuts[b];
      const int32_t k = (int32_t)(std::upper_bound(off.begin(), off.end(), r) - off.begin()) - 1;
      const int64_t end_row = off[k + 1];
      // the tiling may cover only a prefix of the rows (the autotuner's), whose
      // last instance it cuts short: its bands end at the last tile
      jend[b] = (int32_t)std::min<int64_t>(
          tiles.T, std::lower_bound(tiles.cuts.begin(), tiles.cuts.end(), end_row) - tiles.cuts.begin());
      if (bt) (*bt)[b] = t[k];
    }
  };
  int32_t dm = dense_mode(*o), kd_cap = o->dense_features;
  double ms_tune = 0.0;
  TuneChoice tc = choice_from_opts(*o);  // when not tuning
  if (tune) {  // the tuner's candidates decide at the lowest threshold of the batch
    tc = autotune(st, *o, s, 1, bounds, t_min, [&](RunConfig& r, const Built& b) {
      r.t = t_min;
      r.self = 1;
      band_ranges(b.bands, b.tiles, r.band_jend, nullptr);
    });
    p = tc.p;
    dm = tc.dm;
    kd_cap = tc.kd_cap;
    if (p > 0) s = with_split(ar, st, s, p);
    ms_tune = tc.ms;
  }
  tm.start();
  Built bc = build_config(ar, st, s, p, 1, kd_cap, dm, bounds, CAP_S, tc.band_rows, tc.tile_rows);
  Tiers& tr = bc.tr;
  TileSet& bands = bc.bands;
  TileSet& tiles = bc.tiles;
  std::vector<int32_t> hjend;
  std::vector<double> ht;
  band_ranges(bands, tiles, hjend, &ht);
  double* dt = ar.alloc<double>(bands.T);
  if (bands.T > 0)
    APSS_HIP(hipMemcpyAsync(dt, ht.data(), sizeof(double) * bands.T, hipMemcpyHostToDevice, st));
  // per row: its instance's threshold (the band kernel and the FP64 rescore
  // decide every pair against the threshold of its own instance)
  std::vector<double> hrt((size_t)std::max<int64_t>(n_all, 1));
  for (int32_t k = 0; k < count; ++k)
    for (int64_t r = off[k]; r < off[k + 1]; ++r) hrt[(size_t)r] = t[k];
  double* drt = ar.alloc<double>(n_all);
  if (n_all > 0)
    APSS_HIP(hipMemcpyAsync(drt, hrt.data(), sizeof(double) * n_all, hipMemcpyHostToDevice, st));
  const double ms_index = tm.stop_ms();
  RunConfig rc;
  rc.t = t_min;
  rc.p = p;
  rc.self = 1;
  rc.rescore_all = o->rescore_all;
  rc.band_jend = std::move(hjend);
  rc.band_t = dt;
  rc.row_t = drt;
  rc.t_min = t_min;
  rc.use_graph = true;  // the batch's device job is one graph launch
  rc.items_per_cu = tc.items_per_cu;
  rc.count_work = o->count_work != 0;
  ScanOutput so = scan_and_rescore(ar, st, s, bands, s, tiles, tr, rc);
  // split the pair list by instance
  std::vector<apss_pair> h((size_t)so.count);
  if (so.count)
    APSS_HIP(hipMemcpyAsync(h.data(), so.pairs, sizeof(apss_pair) * so.count, hipMemcpyDeviceToHost, st));
  APSS_HIP(hipStreamSynchronize(st));
  std::vector<int64_t> cnt((size_t)count, 0);
  std::vector<int32_t> inst((size_t)so.count);
  for (int64_t q = 0; q < so.count; ++q) {
    const int32_t k = (int32_t)(std::upper_bound(off.begin(), off.end(), (int64_t)h[q].i) - off.begin()) - 1;
    inst[q] = k;
    ++cnt[k];
  }
  for (int32_t k = 0; k < count; ++k) {
    out[k].pairs = host_pairs_alloc(cnt[k]);
    out[k].count = 0;
  }
  for (int64_t q = 0; q < so.count; ++q) {
    apss_result& R = out[inst[q]];
    apss_pair pr = h[q];
    pr.i -= (uint32_t)off[inst[q]];
    pr.j -= (uint32_t)off[inst[q]];
    R.pairs[R.count++] = pr;
  }
  const double ms_total =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  for (int32_t k = 0; k < count; ++k) {
    if (o->sort_output) sort_pairs(out[k].pairs, out[k].count);
    out[k].p_used = p;
    out[k].cells = X[k]->n_rows * (X[k]->n_rows - 1) / 2;
    out[k].band = so.band;
    out[k].products = so.products;
    out[k].list_visits = so.visits;
    out[k].cursor_products = so.cursor;
    out[k].dense_rows = so.dense_rows;
    out[k].tile_pairs = so.tile_pairs;
    copy_model_counters(so, &out[k]);
    out[k].ms_upload = ms_upload;
    out[k].ms_index = ms_index;
    out[k].ms_scan = so.ms_scan;
    out[k].ms_rescore = so.ms_rescore;
    out[k].ms_total = ms_total;
    out[k].ms_tune = ms_tune;
    out[k].band_rows = tc.band_rows;
    out[k].tile_rows = tc.tile_rows;
    out[k].items_per_cu = tc.items_per_cu;
    out[k].graph_reused = so.graph_reused ? 1 : 0;
    out[k].bitmap_features = tr.n_bitmap;
    out[k].mfma_features = tr.n_dense;
  }
  return APSS_OK;
}

}  // namespace apss
