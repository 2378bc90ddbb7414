// Host-side engine of libapss.so: device memory, index build, scan launches.
#pragma once
#include <functional>
#include <map>
#include <mutex>
#include <memory>

#include "apss_internal.h"

namespace apss {

// ---- kernels (defined in index_build.hip / tile_scan.hip) ----
__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v);
__global__ void k_fill_u64(unsigned long long* p, int64_t n, unsigned long long v);
__global__ void k_copy_i64(int64_t* dst, const int64_t* src, int64_t n);
__global__ void k_validate_norms(int64_t n, int64_t d, int64_t nnz, const int64_t* rp,
                                 const int32_t* col, const float* val, int32_t p, float* norm,
                                 float* hnorm, uint32_t* err);
__global__ void k_col_hist(int64_t nnz, const int32_t* col, int32_t p, int64_t chunk, int32_t* cnt);
__global__ void k_cursor_scatter(int64_t n, const int64_t* rp, const int32_t* col, const float* val,
                                 const int32_t* fmap, unsigned long long* cur, int2* sent);
__global__ void k_sort_small(const int32_t* lists, int32_t nl, const int64_t* sptr, int2* sent);
__global__ void k_sort_chunks(const SortChunk* chunks, int2* sent);
__global__ void k_merge_runs(const MergeTask* tasks, const int2* src, int2* dst);
__global__ void k_row_counts(int64_t n, const int64_t* rp, const int32_t* col, int32_t p,
                             const int32_t* fmap, int32_t* dcnt, int32_t* scnt, int32_t* hcnt);
__global__ void k_tile_count(BuildSide s);
__global__ void k_scan_i64(int64_t* a, int64_t n);
__global__ void k_tile_ptr_scan(BuildSide s);
__global__ void k_tile_scatter(BuildSide s);
__global__ void k_tile_compact(BuildSide s);
__global__ void k_sort_records(const int64_t* sbase, int4* srec);
__global__ void k_row2tile(const int64_t* tstart, int32_t* row2tile);
__global__ void k_dense_rows(int64_t n, const int64_t* rp, const int32_t* col, const float* val,
                             int32_t p, const int32_t* fmap, int32_t dm, int32_t B, uint32_t* dense,
                             float* dnorm);
__global__ void k_cursor_pos(int64_t n, const int2* sent, const int32_t* row2tile,
                             const int64_t* tstart, uint32_t* spos, float* sval);
template <int DM, bool HM, bool SP, bool ST>
__global__ void k_band_scan(ScanArgs a);
__global__ void k_tile_l1max(const int64_t* tstart, const int64_t* rp, const float* val, float2* out);
__global__ void k_row_bounds(int64_t n, const int64_t* rp, const float* norm, const float* hnorm,
                             RowBoundArgs b, int2* out);
__global__ void k_rescore(const unsigned long long* band, const unsigned long long* band_cnt,
                          int64_t band_cap, const int64_t* rpI, const int32_t* colI,
                          const float* valI, int64_t nnzI, const int64_t* rpJ,
                          const int32_t* colJ, const float* valJ, int64_t nnzJ, double t,
                          const double* row_t, int64_t i_off, int64_t j_off, apss_pair* out,
                          unsigned long long* out_cnt, int64_t out_cap);

// ---- device memory --------------------------------------------------------

// Device blocks kept for reuse: blocks freed by an arena that draws on the
// pool go back to it instead of to hipFree, so repeated calls stop calling
// hipMalloc / hipFree (which synchronise the device) once the pool has grown
// to their working set.  Each serving slot has a pool of its own; apss_run and
// apss_run_batched share one per device (run_pool), which keeps at most
// RUN_POOL_KEEP idle bytes.  Thread-safe.
class DevicePool {
 public:
  explicit DevicePool(size_t keep_max = 0) : keep_max_(keep_max) {}  // 0: no limit
  ~DevicePool();
  DevicePool(const DevicePool&) = delete;
  DevicePool& operator=(const DevicePool&) = delete;
  // A free block of at least `bytes` (and at most twice that), or nullptr.
  void* take(size_t bytes, size_t* got);
  // Keeps the block, or frees it when the pool would exceed its limit.
  void give(void* p, size_t bytes);
  // Frees every idle block (before retrying a failed allocation).
  void trim();
  int64_t held() const;

 private:
  mutable std::mutex mu_;
  std::multimap<size_t, void*> free_;
  size_t idle_ = 0;
  size_t keep_max_;
};
constexpr size_t RUN_POOL_KEEP = (size_t)8 << 30;
DevicePool& run_pool(int device);

// Owns every device allocation of one call (or one index); enforces a byte
// budget and frees everything on destruction (or returns it to a pool).
class Arena {
 public:
  Arena(int device, int64_t budget, DevicePool* pool = nullptr)
      : device_(device), budget_(budget), pool_(pool) {}
  // Bump mode: every allocation is carved out of [base, base + bytes), memory
  // the caller owns; nothing is allocated or freed.  release() gives memory
  // back only from the top (a block released below the top stays taken until
  // the blocks above it are released too).
  Arena(int device, void* base, int64_t bytes)
      : device_(device), budget_(bytes), bump_base_(static_cast<char*>(base)) {}
  int device() const { return device_; }
  ~Arena();
  Arena(const Arena&) = delete;
  Arena& operator=(const Arena&) = delete;
  template <class T>
  T* alloc(int64_t count) {
    return static_cast<T*>(raw(count < 1 ? sizeof(T) : (size_t)count * sizeof(T)));
  }
  // The stream whose work may still use the arena's blocks: the destructor
  // waits for it before the blocks go back to a pool that other threads share
  // (hipFree would wait by itself).
  void set_stream(hipStream_t s) { stream_ = s; }
  void release(void* p);  // free one allocation early (its users must be done)
  void* detach(void* p);  // hand one allocation to the caller (not freed here)
  int64_t used() const { return used_; }
  int64_t peak() const { return peak_; }
  // Bytes that can still be allocated under the budget (INT64_MAX: no budget).
  int64_t room() const { return budget_ > 0 ? budget_ - used_ : INT64_MAX; }
  bool bump() const { return bump_base_ != nullptr; }

 private:
  void* raw(size_t bytes);
  void drop(void* p, size_t bytes);
  int device_;
  int64_t budget_;
  DevicePool* pool_;
  hipStream_t stream_ = nullptr;
  int64_t used_ = 0, peak_ = 0;
  char* bump_base_ = nullptr;
  struct Blk {
    void* p;
    size_t size;  // bytes of the block (a pooled block may exceed the request)
    size_t req;   // bytes counted against the budget
    bool freed = false;  // bump mode: released, waiting for the blocks above it
  };
  std::vector<Blk> blocks_;
};

// ---- index pieces -----------------------------------------------------------

// Feature classes and the cursor-tier lists of a corpus matrix.
struct Tiers {
  int64_t d = 0;
  int32_t p = 0;
  int32_t KD = 64, KW = 1;
  int32_t n_bitmap = 0;          // features in the bitmap tier
  int32_t dm = DENSE_OFF;        // dense MFMA head mode
  int32_t n_dense = 0;           // features in the dense head
  int64_t n_cursor_entries = 0;
  int32_t* fmap = nullptr;       // d
  int64_t* sptr = nullptr;       // d+1
  int2* sent = nullptr;          // {row, value bits}
  uint32_t* spos = nullptr;      // J tile << 8 | row in tile, per entry (set per J tiling)
  float* sval = nullptr;         // value of each entry (set with spos)
};

// One matrix side: CSR on the device and its norms.
struct Side {
  int64_t n = 0, nnz = 0, d = 0;
  const int64_t* rp = nullptr;
  const int32_t* col = nullptr;
  const float* val = nullptr;
  float* norm = nullptr;
  float* hnorm = nullptr;
  int32_t maxlen = 0;
  float maxnorm = 0.f;
  bool nonneg = false;           // no negative value (set by the validation)
  // dense head records (n x 256 bytes) for one tiering of the corpus
  uint4* dense = nullptr;
  float* dnorm = nullptr;        // BF16 head: norm of each row's head part
  int32_t dense_B = 0;           // int8 head: values are scaled by 2^dense_B
};

// Per-tile structures of one tiling of a side.
struct TileSet {
  int32_t T = 0;
  int64_t* tstart = nullptr;
  int32_t* row2tile = nullptr;
  uint64_t* bmp = nullptr;
  uint16_t* wpre = nullptr;
  int64_t* cbase = nullptr;
  int32_t* cst = nullptr;
  int64_t* dbase = nullptr;
  uint8_t* drow = nullptr;
  float* dval = nullptr;
  uint16_t* dfeat = nullptr;
  int64_t* sbase = nullptr;
  int4* srec = nullptr;
  uint16_t* dseg = nullptr;      // T x kds segment starts by bitmap index
  float2* l1m = nullptr;         // T: largest row L1 norm and |value| of each tile
  int32_t kds = 0;
  int64_t nD = 0, nS = 0;
  int64_t max_s = 0;             // most cursor records in one tile
  std::vector<int64_t> cuts;     // host copy of tstart
  std::vector<uint8_t> rare;     // per tile: its rows hold cursor-tier features only (no
                                 // bitmap-tier, dense-head or head feature)
};

TileSide make_view(const Side& s, const TileSet& ts);
// Returns every device array of a tile set to its arena (its users must be done).
void release_tiles(Arena& ar, TileSet& ts);

struct Timer {  // hipEvent pair on a stream
  hipEvent_t a = nullptr, b = nullptr;
  hipStream_t s;
  explicit Timer(hipStream_t st);
  ~Timer();
  void start();
  double stop_ms();  // synchronises
};

struct RunConfig {
  double t = 0.5;
  int32_t p = 0;
  int32_t self = 1;
  int32_t rescore_all = 0;
  int32_t kd_cap = 0;
  int64_t i_off = 0, j_off = 0;
  std::vector<int32_t> band_jend;      // per band: one past its last partner tile (empty: all)
  const double* band_t = nullptr;      // device, per band: lowest row threshold (batched runs,
                                       // packed queries)
  const double* row_t = nullptr;       // device, per I row: threshold (NULL: t)
  double t_min = 0.0;                  // smallest per-band threshold
  int64_t out_cap_hint = 0;            // initial output capacity in pairs (0: from free memory)
  bool use_graph = false;              // run the device job (counters, scan, rescore) as one
                                       // captured hipGraph
  bool tune_only = false;              // only time a count-only launch (ms_scan = the
                                       // estimated full scan time; see tune_bands)
  int32_t tune_bands = 0;              // tune_only: sample this many bands (0: all work items)
  int32_t items_per_cu = 8;            // work items per CU (sets the partner-tile chunk)
  bool count_work = false;             // count the bitmap-tier list visits and products (a
                                       // slower form of the band kernel)
  apss_pair* out_given = nullptr;      // output buffer owned by the caller (device): no size
  int64_t out_given_cap = 0;           // estimate and no regrowth; a larger result is
                                       // counted (ScanOutput::count) but not all stored
  int64_t cprod_floor = (int64_t)1 << 19;  // least cursor-product records per workgroup
};

// Configuration chosen by the autotuner (p < 0).
struct TuneChoice {
  int32_t p = 0;           // head / tail split
  int32_t dm = DENSE_OFF;  // dense MFMA head mode
  int32_t kd_cap = 0;      // cap of the bitmap tier (0: the mode's maximum)
  int32_t band_rows = BAND_ROWS, tile_rows = TILE_ROWS;  // height caps of bands and tiles
  int32_t items_per_cu = 8;  // work items per CU of the scan's queue
  int32_t candidates = 0;  // configurations timed
  double est_ms = 0.0;     // estimated scan time of the choice
  double ms = 0.0;         // time spent tuning
  bool cached = false;     // taken from an earlier job of the same shape
};

// Everything a configuration builds on top of a prepared side.
struct Built {
  Tiers tr;
  TileSet bands, tiles;
};
// Tiers, band and tile tilings, cursor tiles and dense records of one configuration.
Built build_config(Arena& ar, hipStream_t st, Side& s, int32_t p, int32_t self, int32_t kd_cap,
                   int32_t dm, const std::vector<int64_t>& bounds, int32_t cap_s = CAP_S,
                   int32_t band_rows = BAND_ROWS, int32_t tile_rows = TILE_ROWS);
// The configuration the options ask for (band / tile height caps), without tuning.
TuneChoice choice_from_opts(const apss_opts& o);
// The side's norms and head norms recomputed for the split p (new arrays).
Side with_split(Arena& ar, hipStream_t st, const Side& s, int32_t p);
// Autotuner: on a prefix of the rows, times count-only scans of candidate
// configurations in three stages (the split p on a band sample; the tiering --
// dense head on / off, bitmap tier at its maximum / 1024 features -- and the
// band / tile heights; the work items per CU) and returns the fastest.  `cfg`
// fills the run configuration of a candidate's tiling; `t` is the threshold
// the candidates decide at.  A choice is kept for the life of the process per
// job shape (device, rows, nonzeros, columns, longest row, largest norm,
// instance starts, threshold, head options): a later job of the same shape
// takes it without timing (TuneChoice::cached).  The choice changes the
// speed only, never the pair set.
TuneChoice autotune(hipStream_t st, const apss_opts& o, Side& s, int32_t self,
                    const std::vector<int64_t>& bounds, double t,
                    const std::function<void(RunConfig&, const Built&)>& cfg);

struct ScanOutput {
  apss_pair* pairs = nullptr;   // device, owned by the arena
  int64_t count = 0;
  int64_t band = 0, products = 0, head_pairs = 0, cells = 0, visits = 0, cursor = 0;
  int64_t dense_rows = 0, tile_pairs = 0;
  int64_t band_entries = 0, band_records = 0, walk_entries = 0, walk_scatter = 0;
  int64_t seg_row_bytes = 0;
  int64_t skip_no_common = 0, skip_bound = 0;  // tile pairs of sparse bands skipped
  int32_t n_dense = 0;
  bool graph_reused = false;    // the job ran in the kept graph with updated arguments
  double ms_scan = 0, ms_rescore = 0;
};

// Host-side argument checks (no device access).
void check_csr_host(const apss_csr& X, int32_t p);
// Upload (if needed) a CSR matrix and validate it; computes norms.
Side prepare_side(Arena& ar, hipStream_t st, const apss_csr& X, int32_t p);
// Histogram + tier selection + cursor lists of a corpus side.
// all_lists: every feature >= p with a nonempty list goes to the row-sorted
// cursor lists (no bitmap tier, no dense head): the plain CSR -> CSC transpose.
Tiers build_tiers(Arena& ar, hipStream_t st, const Side& s, int32_t p, int32_t self, int32_t kd_cap,
                  int32_t dm, bool all_lists = false);
// The transpose of X's features >= p into caller-owned host arrays.
int run_transpose(const apss_csr* X, int32_t p, const apss_opts* o, int64_t* col_ptr, int32_t* row,
                  float* val);
// Dense-head records of a side for the tiers' head features (no-op without a head).
void build_dense(Arena& ar, hipStream_t st, Side& s, const Tiers& tr);
// Dense-head mode asked for by the options.
int32_t dense_mode(const apss_opts& o);
// Per-tile bitmap tier (+ cursor records when `records`) of a side against
// corpus tiers, for tiles of at most max_rows rows and cap_d bitmap entries.
TileSet build_tiles(Arena& ar, hipStream_t st, const Side& s, const Tiers& tr, int32_t self,
                    int32_t max_rows, int32_t cap_d, int32_t cap_s, bool records,
                    const std::vector<int64_t>& bounds);
// J tile and row in the tile of every cursor-list entry for the J tiling `tiles`.
void set_cursor_tiles(Arena& ar, hipStream_t st, Tiers& tr, const TileSet& tiles);
// Band scan + rescore; output pairs stay on the device.
ScanOutput scan_and_rescore(Arena& ar, hipStream_t st, const Side& I, const TileSet& bands,
                            const Side& J, const TileSet& tiles, const Tiers& tr,
                            const RunConfig& rc);

// Full self-join into a result (host or device pairs per opts).
int run_self(const apss_csr* X, double t, int32_t p, const apss_opts* o, apss_result* out);
// Self-join without device allocations (scratch array + caller's output buffer).
int run_noalloc(const apss_csr* X, double t, int32_t p, const apss_opts* o, apss_pair* pairs,
                int64_t capacity, int64_t* count, apss_result* out);
// Many independent self-joins in one band-scan launch.
int run_batched(const apss_csr* const* X, int32_t count, const double* t, int32_t p,
                const apss_opts* o, apss_result* out);
// Copy device pairs into a host result (optionally sorted).
void finish_result(hipStream_t st, const apss_pair* dpairs, int64_t count, bool sort, apss_result* out);
// Stream owned for the duration of a call when the caller gives none.
struct StreamScope {
  hipStream_t s = nullptr;
  bool own = false;
  StreamScope(void* given, int device);
  ~StreamScope();
};

// Adds a launch's byte-model counters (band entries, records, walk reads) to a
// result; the segment-row size is the launch's.
void copy_model_counters(const ScanOutput& so, apss_result* out);
// Sorting and copying helpers.
void sort_pairs(apss_pair* p, int64_t n);

// Host result blocks.  Results of 1 MiB to 2 GiB live in pinned memory from a
// process-wide pool (blocks are reused across results, so device-to-host
// copies run at DMA speed into pages that are already resident); the others
// come from malloc.  Every host result (apss_result.pairs with on_device = 0)
// is allocated and released through these.
apss_pair* host_pairs_alloc(int64_t count);                       // room for max(count, 1) pairs
apss_pair* host_pairs_grow(apss_pair* p, int64_t used, int64_t count);  // keeps the first `used`
void host_pairs_free(apss_pair* p);

// Host-to-device copies of pageable host arrays.  A large batch streams
// through two pinned staging blocks: host threads fill one block while the
// DMA engine reads the other (pageable arrays cannot be pinned in place:
// hipHostRegister is refused on this platform).  Small batches, or no pinned
// memory, go straight to hipMemcpyAsync.  The copies are complete on `st`
// in stream order; the call returns once the staging blocks are free.
struct H2DCopy {
  void* dst;
  const void* src;
  size_t bytes;
};
void h2d_copies(const std::vector<H2DCopy>& list, hipStream_t st);
// Device-to-host copy into `dst`, complete on return.  A large copy into
// pageable memory runs through two pinned staging blocks: the DMA engine
// fills one while host threads empty the other into `dst`.
void d2h_copy(void* dst, const void* src, size_t bytes, hipStream_t st);
inline int blocks_for(int64_t n, int threads, int cap = 1 << 20) {
  int64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (int)b;
}

}  // namespace apss
