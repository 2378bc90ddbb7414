// apss_run_noalloc: the self-join of apss_run without any device allocation,
// for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// Every device structure of the call (norms, feature tiers, cursor lists, band
// and tile blocks, dense head records, the scan's per-workgroup regions and
// counters, the error-band list) is carved out of `scratch`, a file-scope
// device array of SCRATCH_BYTES; the pairs go to the caller's device buffer.
// Each GPU holds its own copy of the array (for as long as the process runs);
// its device address is looked up once per device with hipGetSymbolAddress,
// with that device current, and only that address is used from the host.
// Calls on one GPU share the array: the library serialises them with a
// per-device lock held for the whole call (which ends with the stream
// synchronised), so a second call waits for the first.  Everything that must
// start at zero is cleared by fill kernels at the start of the call.
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>

#include "engine.h"

namespace apss {

constexpr int64_t SCRATCH_BYTES = (int64_t)APSS_NOALLOC_SCRATCH_MIB << 20;
// band cut for the cursor tier: at most this many records per band, so each
// workgroup's cursor-product region (records x tile rows) stays small
constexpr int32_t NOALLOC_CAP_S = 1024;

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {
struct ScratchSlot {
  void* ptr = nullptr;  // device address of this GPU's copy of `scratch`
  std::mutex run;       // held by the call that uses the array
};

// The slot of `device`; the device must be current.
ScratchSlot& scratch_of(int device) {
  static std::mutex mu;
  static std::map<int, ScratchSlot*>* slots = new std::map<int, ScratchSlot*>;  // never freed
  std::lock_guard<std::mutex> lk(mu);
  ScratchSlot*& s = (*slots)[device];
  if (!s) {
    auto fresh = std::make_unique<ScratchSlot>();
    APSS_HIP(hipGetSymbolAddress(&fresh->ptr, HIP_SYMBOL(scratch)));
    if (!fresh->ptr) throw Error{APSS_E_DEVICE, "the scratch array has no device address"};
    s = fresh.release();
  }
  return *s;
}
}  // namespace

int run_noalloc(const apss_csr* X, double t, int32_t p, const apss_opts* o, apss_pair* pairs,
                int64_t capacity, int64_t* count, apss_result* out) {
  apss_opts def;
  apss_opts_init(&def);
  if (!o) o = &def;
  if (!X || !pairs || !count) throw Error{APSS_E_INVALID, "null argument"};
  if (!(t == t) || std::isinf(t)) throw Error{APSS_E_INVALID, "t must be finite"};
  if (capacity < 0) throw Error{APSS_E_INVALID, "negative capacity"};
  if (!X->on_device) throw Error{APSS_E_INVALID, "apss_run_noalloc takes a device CSR (on_device = 1)"};
  if (p < 0) throw Error{APSS_E_INVALID, "apss_run_noalloc does not autotune (p < 0)"};
  if (o->sort_output) throw Error{APSS_E_INVALID, "apss_run_noalloc does not sort its output"};
  check_csr_host(*X, p);
  // every nonzero lands in at least 8 bytes of tier structures, every row in
  // 8 bytes of norms, every feature in 12 bytes of class and list pointers
  const int64_t least = 8 * X->nnz + 8 * X->n_rows + 12 * (X->n_cols + 1);
  if (least > SCRATCH_BYTES)
    throw Error{APSS_E_NOMEM, "input needs at least " + std::to_string(least) + " bytes, beyond the " +
                                  std::to_string(SCRATCH_BYTES) + "-byte scratch array"};
  *count = 0;
  if (out) std::memset(out, 0, sizeof(*out));
  APSS_HIP(hipSetDevice(o->device));
  ScratchSlot& slot = scratch_of(o->device);
  std::lock_guard<std::mutex> busy(slot.run);
  StreamScope ss(o->stream, o->device);
  hipStream_t st = ss.s;
  const auto t0 = std::chrono::steady_clock::now();
  Arena ar(o->device, slot.ptr, SCRATCH_BYTES);
  ar.set_stream(st);
  Side s = prepare_side(ar, st, *X, p);
  Built bc = build_config(ar, st, s, p, 1, o->dense_features, dense_mode(*o), {}, NOALLOC_CAP_S);
  RunConfig rc;
  rc.t = t;
  rc.p = p;
  rc.self = 1;
  rc.rescore_all = o->rescore_all;
  rc.out_given = pairs;
  rc.out_given_cap = capacity;
  rc.cprod_floor = (int64_t)1 << 16;
  rc.count_work = o->count_work != 0;
  ScanOutput so = scan_and_rescore(ar, st, s, bc.bands, s, bc.tiles, bc.tr, rc);
  APSS_HIP(hipStreamSynchronize(st));
  *count = so.count;
  if (out) {
    out->count = so.count;
    out->on_device = 1;
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
    out->bitmap_features = bc.tr.n_bitmap;
    out->band_rows = BAND_ROWS;
    out->tile_rows = TILE_ROWS;
    out->items_per_cu = rc.items_per_cu;
    copy_model_counters(so, out);
    out->ms_scan = so.ms_scan;
    out->ms_rescore = so.ms_rescore;
    out->ms_total =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  if (so.count > capacity)
    throw Error{APSS_E_NOMEM, "the result has " + std::to_string(so.count) + " pairs, the buffer room for " +
                                  std::to_string(capacity)};
  return APSS_OK;
}

}  // namespace apss
