// C-ABI entry points of libapss.so (declared in include/apss.h).
#include <cstdlib>
#include <cstring>

#include "engine.h"

namespace apss {
thread_local std::string g_last_error;

template <class F>
static int guarded(F&& f) {
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

using namespace apss;

extern "C" {

const char* apss_version(void) { return "apss-mi355x 0.1.0 (gfx950)"; }

const char* apss_strerror(int code) {
  switch (code) {
    case APSS_OK: return "ok";
    case APSS_E_INVALID: return "invalid argument or malformed CSR matrix";
    case APSS_E_NOMEM: return "out of memory or over the memory budget";
    case APSS_E_DEVICE: return "HIP runtime error";
    case APSS_E_UNSUPPORTED: return "input outside the supported envelope";
    case APSS_E_CANCELLED: return "request cancelled";
    case APSS_E_BUSY: return "server busy (admission refused)";
    default: return "unknown error";
  }
}

const char* apss_last_error(void) { return g_last_error.c_str(); }

void apss_opts_init(apss_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
}

int apss_run(const apss_csr* X, double t, int32_t p, const apss_opts* o, apss_result* out) {
  return guarded([&] { return run_self(X, t, p, o, out); });
}

int apss_run_cb(const apss_csr* X, double t, int32_t p, const apss_opts* o, apss_emit_fn emit,
                void* user, apss_result* out) {
  if (!emit) {
    g_last_error = "null emit callback";
    return APSS_E_INVALID;
  }
  apss_opts local;
  apss_opts_init(&local);
  if (o) local = *o;
  // the pairs stay on the device and reach the host in chunks through one
  // pinned staging block, so the host never holds the whole result
  local.output_on_device = 1;
  apss_result r;
  int rc = apss_run(X, t, p, &local, &r);
  if (rc != APSS_OK) return rc;
  return guarded([&] {
    const int64_t chunk = (int64_t)1 << 20;
    struct Release {  // the device result and the staging block, on every exit
      apss_pair* s;
      apss_result* r;
      ~Release() {
        host_pairs_free(s);
        apss_result_free(r);
      }
    } rel{nullptr, &r};
    apss_pair* stage = host_pairs_alloc(std::min<int64_t>(chunk, std::max<int64_t>(r.count, 1)));
    rel.s = stage;
    for (int64_t k = 0; k < r.count; k += chunk) {
      const int64_t m = std::min<int64_t>(chunk, r.count - k);
      APSS_HIP(hipMemcpy(stage, r.pairs + k, sizeof(apss_pair) * (size_t)m, hipMemcpyDeviceToHost));
      emit(stage, (size_t)m, user);
    }
    if (out) {
      *out = r;
      out->pairs = nullptr;
      out->on_device = 0;
    }
    return APSS_OK;
  });
}

int apss_transpose(const apss_csr* X, int32_t p, const apss_opts* o, int64_t* col_ptr, int32_t* row,
                   float* val) {
  return guarded([&] { return run_transpose(X, p, o, col_ptr, row, val); });
}

int apss_run_noalloc(const apss_csr* X, double t, int32_t p, const apss_opts* o, apss_pair* pairs,
                     int64_t capacity, int64_t* count, apss_result* out) {
  return guarded([&] { return run_noalloc(X, t, p, o, pairs, capacity, count, out); });
}

int apss_run_batched(const apss_csr* const* X, int32_t count, const double* t, int32_t p,
                     const apss_opts* o, apss_result* out) {
  return guarded([&] { return run_batched(X, count, t, p, o, out); });
}

void apss_result_free(apss_result* r) {
  if (!r) return;
  if (r->pairs) {
    if (r->on_device) (void)hipFree(r->pairs);
    else apss::host_pairs_free(r->pairs);
  }
  r->pairs = nullptr;
  r->count = 0;
}

}  // extern "C"
