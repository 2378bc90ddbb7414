// Serving: a corpus index resident in device memory, queried by batches of
// query rows, plus a continuous-batching server over it.
//
// The reference has no serving path (one batch job, similarity_search.cu:934);
// the index keeps what a query needs resident: corpus CSR + norms, the feature
// tiers and cursor lists, and the 64-row tile blocks of the bitmap tier.  A
// query builds only its own 128-row bands (against the corpus tiers) and runs
// one band-scan launch per corpus segment in cross mode (every query row meets
// every corpus row).
//
// Inserts append a delta segment (its own tiers and tiles, row ids continuing
// the corpus), built outside the query lock; when more than MAX_SEGMENTS
// segments exist the delta segments are merged into one (and into the base
// segment once they hold a quarter of its rows).  Queries hold a shared lock;
// an insert takes the exclusive lock only to swap segments in.  Each serving
// slot keeps its device blocks in a pool, so steady-state requests allocate
// nothing with hipMalloc.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <thread>

#include "engine.h"

namespace apss {

constexpr size_t MAX_SEGMENTS = 5;  // the base segment and up to four deltas

struct HostCSR {
  std::vector<int64_t> rp{0};
  std::vector<int32_t> col;
  std::vector<float> val;
  int64_t d = 0;
  int64_t n() const { return (int64_t)rp.size() - 1; }
  apss_csr view() const {
    return apss_csr{n(), d, (int64_t)col.size(), rp.data(), col.data(), val.data(), 0, 0};
  }
  void append(const apss_csr& X) {
    const int64_t e0 = (int64_t)col.size();
    for (int64_t i = 0; i < X.n_rows; ++i) rp.push_back(e0 + X.row_ptr[i + 1]);
    col.insert(col.end(), X.col, X.col + X.nnz);
    val.insert(val.end(), X.val, X.val + X.nnz);
  }
};

// Copy a CSR (host or device) into host vectors.
static HostCSR to_host(const apss_csr& X, int device) {
  HostCSR h;
  h.d = X.n_cols;
  if (!X.on_device) {
    h.append(X);
    return h;
  }
  APSS_HIP(hipSetDevice(device));
  h.rp.resize((size_t)X.n_rows + 1);
  h.col.resize((size_t)X.nnz);
  h.val.resize((size_t)X.nnz);
  APSS_HIP(hipMemcpy(h.rp.data(), X.row_ptr, sizeof(int64_t) * (X.n_rows + 1), hipMemcpyDeviceToHost));
  if (X.nnz) {
    APSS_HIP(hipMemcpy(h.col.data(), X.col, sizeof(int32_t) * X.nnz, hipMemcpyDeviceToHost));
    APSS_HIP(hipMemcpy(h.val.data(), X.val, sizeof(float) * X.nnz, hipMemcpyDeviceToHost));
  }
  return h;
}

struct Segment {
  std::unique_ptr<Arena> ar;
  HostCSR host;
  Side side;
  Tiers tr;
  TileSet tiles;
  int64_t row_off = 0;
};

struct Index {
  int device = 0;
  int32_t p = 0;
  int64_t d = 0;
  apss_opts opts{};
  int64_t n_rows = 0;
  int64_t scratch = 0;  // device bytes of one query launch's scan scratch
  std::vector<std::unique_ptr<Segment>> segs;
  mutable std::shared_mutex mu;
  std::mutex insert_mu;  // serialises inserts (segment builds run outside mu)
};

// Upper bound of the scan scratch one query launch allocates: per workgroup
// of the persistent grid, cursor state for a band's records and the cursor
// product region (both sized by the band cut CAP_S).
static int64_t scan_scratch_bytes(int device) {
  int n_cu = 256;
  (void)hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device);
  const int64_t per_wg = (int64_t)CAP_S * (int64_t)sizeof(int4) +
                         std::max<int64_t>((int64_t)1 << 19, (int64_t)CAP_S * TILE_ROWS) * (int64_t)sizeof(int2);
  return (int64_t)std::max(n_cu, 1) * per_wg;
}

// Output room a query batch starts with (grown and rerun if exceeded).
static int64_t query_out_hint(int64_t q_rows, int64_t corpus_rows) {
  return std::min<int64_t>(q_rows * std::max<int64_t>(corpus_rows, 1),
                           std::max<int64_t>((int64_t)1 << 16, q_rows * 16384));
}

static std::unique_ptr<Segment> build_segment(HostCSR&& h, int32_t p, const apss_opts& o,
                                              int64_t row_off) {
  auto sg = std::make_unique<Segment>();
  sg->ar = std::make_unique<Arena>(o.device, 0);
  sg->host = std::move(h);
  sg->row_off = row_off;
  StreamScope ss(nullptr, o.device);
  apss_csr v = sg->host.view();
  sg->side = prepare_side(*sg->ar, ss.s, v, p);
  sg->tr = build_tiers(*sg->ar, ss.s, sg->side, p, 0, o.dense_features, dense_mode(o));
  sg->tiles = build_tiles(*sg->ar, ss.s, sg->side, sg->tr, 0, TILE_ROWS, cap_j_for(sg->tr.dm),
                          INT32_MAX, false, {});
  set_cursor_tiles(*sg->ar, ss.s, sg->tr, sg->tiles);
  build_dense(*sg->ar, ss.s, sg->side, sg->tr);
  APSS_HIP(hipStreamSynchronize(ss.s));
  return sg;
}

// Growable host pair buffer (a host result block, see host_pairs_alloc);
// release() hands the block to a result, which apss_result_free releases.
struct PairBuf {
  apss_pair* p = nullptr;
  int64_t n = 0, cap = 0;
  PairBuf() = default;
  PairBuf(const PairBuf&) = delete;
  PairBuf& operator=(const PairBuf&) = delete;
  ~PairBuf() { host_pairs_free(p); }
  apss_pair* grow(int64_t extra) {  // room for `extra` more pairs; returns where they go
    if (n + extra > cap) {
      const int64_t c = std::max<int64_t>(n + extra, std::max<int64_t>(1, cap + cap / 2));
      p = host_pairs_grow(p, n, c);
      cap = c;
    }
    apss_pair* at = p + n;
    n += extra;
    return at;
  }
  apss_pair* release() {
    if (!p) p = host_pairs_alloc(1);
    apss_pair* q = p;
    p = nullptr;
    n = cap = 0;
    return q;
  }
};

// Query rows (host CSR) against every segment; `bounds` are the first rows of
// the requests packed into Q (ascending, excluding 0) and t_req their
// thresholds (one per request).  Pairs (query row, corpus row) go to `pairs`.
static void query_index(const Index& ix, const HostCSR& Q, const std::vector<int64_t>& bounds,
                        const std::vector<double>& t_req, hipStream_t st, int64_t budget,
                        int rescore_all, PairBuf& pairs, apss_result* stats,
                        DevicePool* pool) {
  std::shared_lock<std::shared_mutex> lk(ix.mu);
  if (Q.d > ix.d) throw Error{APSS_E_INVALID, "query has more columns than the corpus"};
  Arena ar(ix.device, budget, pool);
  apss_csr qv = Q.view();
  qv.n_cols = ix.d;
  Timer tm(st);
  tm.start();
  Side qs = prepare_side(ar, st, qv, ix.p);
  const double ms_up = tm.stop_ms();
  // Every query row carries its request's threshold, so requests packed in
  // one launch share bands (a band may hold rows of several requests).
  std::vector<int64_t> starts(1, 0);
  starts.insert(starts.end(), bounds.begin(), bounds.end());
  double t_min = INFINITY;
  for (double x : t_req) t_min = std::min(t_min, x);
  std::vector<double> hrt((size_t)std::max<int64_t>(Q.n(), 1), t_min);
  for (size_t k = 0; k < starts.size() && k < t_req.size(); ++k) {
    const int64_t e = k + 1 < starts.size() ? starts[k + 1] : Q.n();
    for (int64_t r = starts[k]; r < e; ++r) hrt[(size_t)r] = t_req[k];
  }
  double* drt = ar.alloc<double>(Q.n());
  if (Q.n() > 0)
    APSS_HIP(hipMemcpyAsync(drt, hrt.data(), sizeof(double) * Q.n(), hipMemcpyHostToDevice, st));
  for (const auto& sg : ix.segs) {
    tm.start();
    TileSet bands = build_tiles(ar, st, qs, sg->tr, 0, BAND_ROWS, cap_i_for(sg->tr.dm), CAP_S, true, {});
    Side qd = qs;  // the query's head records follow this segment's head features
    build_dense(ar, st, qd, sg->tr);
    std::vector<double> ht((size_t)std::max(bands.T, 1), t_min);
    for (int32_t b = 0; b < bands.T; ++b) {  // the lowest threshold of the band's rows
      double m = INFINITY;
      for (int64_t r = bands.cuts[b]; r < bands.cuts[b + 1]; ++r) m = std::min(m, hrt[(size_t)r]);
      ht[b] = m;
    }
    double* dt = ar.alloc<double>(bands.T);
    if (bands.T > 0)
      APSS_HIP(hipMemcpyAsync(dt, ht.data(), sizeof(double) * bands.T, hipMemcpyHostToDevice, st));
    const double ms_idx = tm.stop_ms();
    RunConfig rc;
    rc.t = t_min;
    rc.p = ix.p;
    rc.self = 0;
    rc.rescore_all = rescore_all;
    rc.j_off = sg->row_off;
    rc.band_t = dt;
    rc.row_t = drt;
    rc.t_min = t_min;
    rc.out_cap_hint = query_out_hint(Q.n(), sg->side.n);
    rc.count_work = ix.opts.count_work != 0;
    ScanOutput so = scan_and_rescore(ar, st, qd, bands, sg->side, sg->tiles, sg->tr, rc);
    apss_pair* at = pairs.grow(so.count);
    if (so.count)
      APSS_HIP(hipMemcpyAsync(at, so.pairs, sizeof(apss_pair) * so.count, hipMemcpyDeviceToHost, st));
    APSS_HIP(hipStreamSynchronize(st));
    if (stats) {
      stats->cells += so.cells;
      stats->band += so.band;
      // (-1: the visits and products were not counted)
      stats->products = (stats->products < 0 || so.products < 0) ? -1 : stats->products + so.products;
      stats->list_visits = (stats->list_visits < 0 || so.visits < 0) ? -1 : stats->list_visits + so.visits;
      stats->cursor_products += so.cursor;
      stats->dense_rows += so.dense_rows;
      stats->tile_pairs += so.tile_pairs;
      copy_model_counters(so, stats);
      stats->ms_index += ms_idx;
      stats->ms_scan += so.ms_scan;
      stats->ms_rescore += so.ms_rescore;
    }
    // this segment's query-side structures go before the next segment's are built
    ar.release(so.pairs);
    ar.release(dt);
    if (qd.dense) ar.release(qd.dense);
    if (qd.dnorm) ar.release(qd.dnorm);
    release_tiles(ar, bands);
    if (stats) stats->mfma_features = so.n_dense;
  }
  if (stats) stats->ms_upload = ms_up;
}

// ---------------------------------------------------------------------------
// continuous-batching server

struct Request {
  int64_t ticket = 0;
  HostCSR q;
  double t = 0.5;
  int64_t est = 0;          // device bytes reserved at admission
  enum State { QUEUED, RUNNING, DONE } state = QUEUED;
  int rc = APSS_OK;
  std::string err;          // detail of a failed request (apss_last_error after the wait)
  apss_result res{};
  std::chrono::steady_clock::time_point t_submit, t_done;
};

struct Server {
  apss_index* ix = nullptr;
  Index* idx = nullptr;
  int32_t slots = 1;
  int64_t budget = 0;
  int32_t max_queue = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> queue;
  std::map<int64_t, std::unique_ptr<Request>> reqs;
  int64_t in_use = 0;
  int64_t next_ticket = 0;
  bool stop = false;
  std::vector<std::thread> workers;
};

// Rows packed into one launch when several small requests wait.
constexpr int64_t PACK_ROWS = 4096;

static int64_t request_bytes(const HostCSR& q, int64_t corpus_rows, int64_t scratch) {
  // query-side structures, the output room the launch starts with (+ the band
  // list and a regrow margin) and the scan scratch of one launch
  const int64_t out = query_out_hint(q.n(), corpus_rows) * (int64_t)sizeof(apss_pair) * 5 / 4;
  return ((int64_t)64 << 20) + (int64_t)q.col.size() * 96 + q.n() * 512 + out + scratch;
}

static void worker_loop(Server* s, int slot) {
  (void)slot;
  hipStream_t st = nullptr;
  if (hipSetDevice(s->idx->device) != hipSuccess ||
      hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess)
    st = nullptr;
  // this slot's reusable device blocks, capped at its share of the budget so
  // idle blocks of one slot never hold memory another slot's admitted work needs
  const int64_t keep = s->budget > 0 ? std::max<int64_t>(s->budget / std::max(s->slots, 1), (int64_t)1 << 30)
                                     : (int64_t)RUN_POOL_KEEP;
  DevicePool pool((size_t)keep);
  for (;;) {
    std::vector<Request*> batch;
    int64_t est = 0;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] {
        if (s->stop) return true;
        if (s->queue.empty()) return false;
        const int64_t need = s->queue.front()->est;
        return s->budget <= 0 || s->in_use + need <= s->budget || s->in_use == 0;
      });
      if (s->stop && s->queue.empty()) break;
      if (s->queue.empty()) continue;
      // admit the head, then pack more small requests behind it while rows
      // and memory allow (FIFO order is kept)
      int64_t rows = 0;
      while (!s->queue.empty()) {
        Request* r = s->queue.front();
        const bool fits_mem = s->budget <= 0 || s->in_use + est + r->est <= s->budget || batch.empty();
        if (!batch.empty() && (rows + r->q.n() > PACK_ROWS || !fits_mem)) break;
        s->queue.pop_front();
        r->state = Request::RUNNING;
        batch.push_back(r);
        rows += r->q.n();
        est += r->est;
      }
      s->in_use += est;
    }
    // one launch sequence for the packed batch
    // (requests without rows take no part and complete with no pairs)
    HostCSR Q;
    Q.d = s->idx->d;
    std::vector<int64_t> bounds, lstart;  // interior request starts; start of each live request
    std::vector<double> ts;
    std::vector<size_t> live;             // batch positions of the requests with rows
    for (size_t k = 0; k < batch.size(); ++k) {
      Request* r = batch[k];
      if (r->q.n() == 0) continue;
      if (Q.n() > 0) bounds.push_back(Q.n());
      lstart.push_back(Q.n());
      Q.append(r->q.view());
      ts.push_back(r->t);
      live.push_back(k);
    }
    PairBuf pairs;
    apss_result stats{};
    int rc = APSS_OK;
    std::string err;
    const auto t_start = std::chrono::steady_clock::now();
    try {
      if (!st) throw Error{APSS_E_DEVICE, "stream creation failed"};
      if (!live.empty()) query_index(*s->idx, Q, bounds, ts, st, est, 0, pairs, &stats, &pool);
    } catch (const Error& e) {
      rc = e.code;
      err = e.what;
    } catch (...) {
      rc = APSS_E_DEVICE;
      err = "unexpected exception";
    }
    const auto t_scanned = std::chrono::steady_clock::now();
    // split by request: pairs carry packed query rows; one request takes the
    // whole buffer, several are split by a counting pass
    std::vector<apss_pair*> outp(batch.size(), nullptr);
    std::vector<int64_t> outn(batch.size(), 0);
    try {
      if (rc == APSS_OK && live.size() == 1) {
        outn[live[0]] = pairs.n;
        outp[live[0]] = pairs.release();
      } else if (rc == APSS_OK && live.size() > 1) {
        std::vector<int32_t> req_of((size_t)Q.n());  // live index of every packed row
        for (size_t m = 0; m < live.size(); ++m) {
          const int64_t e = m + 1 < live.size() ? lstart[m + 1] : Q.n();
          for (int64_t r = lstart[m]; r < e; ++r) req_of[(size_t)r] = (int32_t)m;
        }
        std::vector<int64_t> cnt(live.size(), 0), fill(live.size(), 0);
        for (int64_t q = 0; q < pairs.n; ++q) ++cnt[(size_t)req_of[pairs.p[q].i]];
        for (size_t m = 0; m < live.size(); ++m) {
          outn[live[m]] = cnt[m];
          outp[live[m]] = host_pairs_alloc(cnt[m]);
        }
        for (int64_t q = 0; q < pairs.n; ++q) {
          apss_pair x = pairs.p[q];
          const size_t m = (size_t)req_of[x.i];
          x.i -= (uint32_t)lstart[m];
          outp[live[m]][fill[m]++] = x;
        }
      }
      for (size_t k = 0; k < batch.size(); ++k)  // requests without rows (or pairs)
        if (rc == APSS_OK && !outp[k]) {
          outp[k] = host_pairs_alloc(1);
        }
    } catch (const Error& e) {
      rc = e.code;
      err = e.what;
    }
    {
      std::lock_guard<std::mutex> lk(s->mu);
      const auto now = std::chrono::steady_clock::now();
      for (size_t k = 0; k < batch.size(); ++k) {
        Request* r = batch[k];
        r->rc = rc;
        r->err = err;
        r->res = stats;
        r->res.ms_queue = std::chrono::duration<double, std::milli>(t_start - r->t_submit).count();
        r->res.ms_handover = std::chrono::duration<double, std::milli>(now - t_scanned).count();
        r->res.ms_total = std::chrono::duration<double, std::milli>(now - r->t_submit).count();
        r->res.count = rc == APSS_OK ? outn[k] : 0;
        r->res.pairs = rc == APSS_OK ? outp[k] : nullptr;
        if (rc != APSS_OK && outp[k]) host_pairs_free(outp[k]);
        r->res.on_device = 0;
        r->res.p_used = s->idx->p;
        r->state = Request::DONE;
        r->t_done = now;
      }
      s->in_use -= est;
    }
    s->cv_done.notify_all();
    s->cv_work.notify_all();
  }
  if (st) (void)hipStreamDestroy(st);
}

}  // namespace apss

using namespace apss;

struct apss_index {
  Index impl;
};
struct apss_server {
  Server impl;
};

namespace apss {
extern thread_local std::string g_last_error;
template <class F>
static int guard(F&& f) {
  try {
    return f();
  } catch (const Error& e) {
    g_last_error = e.what;
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return APSS_E_NOMEM;
  } catch (...) {
    g_last_error = "unexpected exception";
    return APSS_E_DEVICE;
  }
}
}  // namespace apss

extern "C" {

int apss_index_create(const apss_csr* corpus, int32_t p, const apss_opts* o, apss_index** out) {
  return guard([&] {
    if (!corpus || !out) throw Error{APSS_E_INVALID, "null argument"};
    *out = nullptr;
    apss_opts def;
    apss_opts_init(&def);
    if (!o) o = &def;
    if (p < 0) p = 0;
    check_csr_host(*corpus, p);
    auto ix = std::make_unique<apss_index>();
    ix->impl.device = o->device;
    ix->impl.p = p;
    ix->impl.d = corpus->n_cols;
    ix->impl.opts = *o;
    ix->impl.n_rows = corpus->n_rows;
    ix->impl.scratch = scan_scratch_bytes(o->device);
    ix->impl.segs.push_back(build_segment(to_host(*corpus, o->device), p, *o, 0));
    *out = ix.release();
    return APSS_OK;
  });
}

int apss_index_insert(apss_index* ix, const apss_csr* rows) {
  return guard([&] {
    if (!ix || !rows) throw Error{APSS_E_INVALID, "null argument"};
    Index& I = ix->impl;
    if (rows->n_cols > I.d) throw Error{APSS_E_INVALID, "inserted rows have more columns than the corpus"};
    check_csr_host(*rows, std::min<int64_t>(I.p, rows->n_cols));
    HostCSR h = to_host(*rows, I.device);
    h.d = I.d;
    // only inserts change segs and n_rows, and they hold insert_mu: reading
    // them here needs no query lock; queries keep running while segments build
    std::lock_guard<std::mutex> ins(I.insert_mu);
    auto sg = build_segment(std::move(h), I.p, I.opts, I.n_rows);
    {
      std::unique_lock<std::shared_mutex> lk(I.mu);
      I.segs.push_back(std::move(sg));
      I.n_rows += rows->n_rows;
    }
    if (I.segs.size() > MAX_SEGMENTS) {
      // merge the delta segments into one; fold them into the base segment
      // once they hold a quarter of its rows
      int64_t delta_rows = 0;
      for (size_t k = 1; k < I.segs.size(); ++k) delta_rows += I.segs[k]->host.n();
      const size_t first = 4 * delta_rows >= I.segs[0]->host.n() ? 0 : 1;
      HostCSR all;
      all.d = I.d;
      for (size_t k = first; k < I.segs.size(); ++k) all.append(I.segs[k]->host.view());
      auto merged = build_segment(std::move(all), I.p, I.opts, I.segs[first]->row_off);
      std::vector<std::unique_ptr<Segment>> old;
      {
        std::unique_lock<std::shared_mutex> lk(I.mu);
        for (size_t k = first; k < I.segs.size(); ++k) old.push_back(std::move(I.segs[k]));
        I.segs.resize(first);
        I.segs.push_back(std::move(merged));
      }
    }  // the replaced segments are freed here, outside the lock
    return APSS_OK;
  });
}

int apss_index_query(apss_index* ix, const apss_csr* queries, double t, apss_result* out) {
  return guard([&] {
    if (!ix || !queries || !out) throw Error{APSS_E_INVALID, "null argument"};
    if (!(t == t) || std::isinf(t)) throw Error{APSS_E_INVALID, "t must be finite"};
    std::memset(out, 0, sizeof(*out));
    Index& I = ix->impl;
    check_csr_host(*queries, std::min<int64_t>(I.p, queries->n_cols));
    auto t0 = std::chrono::steady_clock::now();
    HostCSR q = to_host(*queries, I.device);
    StreamScope ss(I.opts.stream, I.device);
    PairBuf pairs;
    query_index(I, q, {}, {t}, ss.s, I.opts.mem_budget, I.opts.rescore_all, pairs, out, nullptr);
    if (I.opts.sort_output) sort_pairs(pairs.p, pairs.n);
    out->count = pairs.n;
    out->pairs = pairs.release();
    out->on_device = 0;
    out->p_used = I.p;
    out->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return APSS_OK;
  });
}

int64_t apss_index_rows(const apss_index* ix) {
  if (!ix) return -1;
  std::shared_lock<std::shared_mutex> lk(ix->impl.mu);
  return ix->impl.n_rows;
}

int apss_index_destroy(apss_index* ix) {
  return guard([&] {
    if (ix) {
      (void)hipSetDevice(ix->impl.device);
      delete ix;
    }
    return APSS_OK;
  });
}

int apss_server_create(apss_index* ix, int32_t slots, int64_t mem_budget, int32_t max_queue,
                       apss_server** out) {
  return guard([&] {
    if (!ix || !out || slots < 1 || slots > 64) throw Error{APSS_E_INVALID, "bad server arguments"};
    auto s = std::make_unique<apss_server>();
    Server& S = s->impl;
    S.ix = ix;
    S.idx = &ix->impl;
    S.slots = slots;
    S.budget = mem_budget;
    S.max_queue = max_queue;
    for (int32_t k = 0; k < slots; ++k) S.workers.emplace_back(worker_loop, &S, k);
    *out = s.release();
    return APSS_OK;
  });
}

int64_t apss_server_submit(apss_server* s, const apss_csr* queries, double t) {
  if (!s || !queries) return APSS_E_INVALID;
  if (!(t == t) || std::isinf(t)) return APSS_E_INVALID;
  Server& S = s->impl;
  try {
    check_csr_host(*queries, std::min<int64_t>(S.idx->p, queries->n_cols));
    if (queries->n_cols > S.idx->d) return APSS_E_INVALID;
    auto r = std::make_unique<Request>();
    r->q = to_host(*queries, S.idx->device);
    r->q.d = S.idx->d;
    r->t = t;
    r->est = request_bytes(r->q, S.idx->n_rows, S.idx->scratch);
    r->t_submit = std::chrono::steady_clock::now();
    std::lock_guard<std::mutex> lk(S.mu);
    if (S.stop) return APSS_E_CANCELLED;
    if (S.max_queue > 0 && (int32_t)S.queue.size() >= S.max_queue) return APSS_E_BUSY;
    if (S.budget > 0 && r->est > S.budget) return APSS_E_NOMEM;
    r->ticket = S.next_ticket++;
    const int64_t tk = r->ticket;
    S.queue.push_back(r.get());
    S.reqs[tk] = std::move(r);
    S.cv_work.notify_one();
    return tk;
  } catch (const Error& e) {
    g_last_error = e.what;
    return e.code;
  } catch (...) {
    return APSS_E_NOMEM;
  }
}

int apss_server_wait(apss_server* s, int64_t ticket, apss_result* out, double* latency_ms) {
  if (!s || !out) return APSS_E_INVALID;
  Server& S = s->impl;
  std::unique_lock<std::mutex> lk(S.mu);
  auto it = S.reqs.find(ticket);
  if (it == S.reqs.end()) return APSS_E_INVALID;
  Request* r = it->second.get();
  S.cv_done.wait(lk, [&] { return r->state == Request::DONE; });
  *out = r->res;
  if (latency_ms)
    *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
  const int rc = r->rc;
  if (rc != APSS_OK) g_last_error = r->err;
  if (rc != APSS_OK && out->pairs) {
    host_pairs_free(out->pairs);
    out->pairs = nullptr;
    out->count = 0;
  }
  S.reqs.erase(it);
  return rc;
}

int apss_server_cancel(apss_server* s, int64_t ticket) {
  if (!s) return APSS_E_INVALID;
  Server& S = s->impl;
  std::lock_guard<std::mutex> lk(S.mu);
  auto it = S.reqs.find(ticket);
  if (it == S.reqs.end()) return APSS_E_INVALID;
  Request* r = it->second.get();
  if (r->state != Request::QUEUED) return APSS_E_BUSY;  // already on the device
  S.queue.erase(std::remove(S.queue.begin(), S.queue.end(), r), S.queue.end());
  r->state = Request::DONE;
  r->rc = APSS_E_CANCELLED;
  r->t_done = std::chrono::steady_clock::now();
  S.cv_done.notify_all();
  return APSS_OK;
}

int apss_server_destroy(apss_server* s) {
  if (!s) return APSS_OK;
  Server& S = s->impl;
  {
    std::lock_guard<std::mutex> lk(S.mu);
    S.stop = true;
    for (Request* r : S.queue) {
      r->state = Request::DONE;
      r->rc = APSS_E_CANCELLED;
    }
    S.queue.clear();
  }
  S.cv_work.notify_all();
  S.cv_done.notify_all();
  for (auto& w : S.workers) w.join();
  for (auto& kv : S.reqs)
    if (kv.second->res.pairs) host_pairs_free(kv.second->res.pairs);
  delete s;
  return APSS_OK;
}

}  // extern "C"
