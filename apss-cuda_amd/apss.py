"""ctypes binding of libapss.so (C-ABI declared in include/apss.h) and of the
synthetic CSR generator in libapss_host.so.

This module is the Python-side mirror of the C entry points; it holds no
compute of its own.  If libapss.so is missing it raises at load time -- there
is no fallback path.

Workload definitions (WORKLOADS) follow the five configurations the reference
method is benchmarked on: row lengths a truncated power law, Zipf-ranked
feature ids (id 0 most frequent), uniform(0,1] or lognormal(0,1) values,
L2-normalised rows.
"""
import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_float, c_int32, c_int64, c_uint32, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libapss.so")
HOST_LIB_PATH = os.path.join(HERE, "libapss_host.so")

APSS_OK, APSS_E_INVALID, APSS_E_NOMEM, APSS_E_DEVICE, APSS_E_UNSUPPORTED, APSS_E_CANCELLED, APSS_E_BUSY = (
    0, -1, -2, -3, -4, -5, -6)
APSS_PREC_FP32, APSS_PREC_BF16_HEAD = 0, 1
PREC_FP32, PREC_BF16_HEAD = 0, 1

PAIR_DTYPE = np.dtype([("i", "<u4"), ("j", "<u4"), ("score", "<f4")])


class CSR(ctypes.Structure):
    _fields_ = [("n_rows", c_int64), ("n_cols", c_int64), ("nnz", c_int64),
                ("row_ptr", c_void_p), ("col", c_void_p), ("val", c_void_p),
                ("on_device", c_int32), ("reserved_", c_int32)]


class Opts(ctypes.Structure):
    _fields_ = [("device", c_int32), ("precision", c_int32), ("sort_output", c_int32),
                ("rescore_all", c_int32), ("output_on_device", c_int32),
                ("dense_features", c_int32), ("mem_budget", c_int64), ("stream", c_void_p),
                ("mfma_head", c_int32), ("band_rows", c_int32), ("count_work", c_int32),
                ("reserved32_", c_int32), ("reserved_", c_int64 * 2)]


class Result(ctypes.Structure):
    _fields_ = [("count", c_int64), ("pairs", c_void_p), ("on_device", c_int32),
                ("p_used", c_int32), ("cells", c_int64), ("band", c_int64),
                ("products", c_int64), ("head_pairs", c_int64), ("list_visits", c_int64),
                ("cursor_products", c_int64), ("mfma_features", c_int64),
                ("dense_rows", c_int64), ("tile_pairs", c_int64), ("bitmap_features", c_int64),
                ("band_entries", c_int64), ("band_records", c_int64), ("walk_entries", c_int64),
                ("walk_scatter_entries", c_int64), ("seg_row_bytes", c_int64),
                ("tile_pairs_skipped", c_int64), ("tile_pairs_pruned", c_int64),
                ("ms_upload", c_double), ("ms_index", c_double), ("ms_scan", c_double),
                ("ms_rescore", c_double), ("ms_download", c_double), ("ms_total", c_double),
                ("ms_tune", c_double), ("band_rows", c_int32), ("tile_rows", c_int32),
                ("items_per_cu", c_int32), ("graph_reused", c_int32), ("ms_queue", c_double), ("ms_handover", c_double),
                ("internal_", c_void_p)]


class GenSpec(ctypes.Structure):
    _fields_ = [("n", c_int64), ("row_offset", c_int64), ("d", c_int32), ("kmin", c_int32),
                ("kmax", c_int32), ("alpha", c_double), ("zipf_s", c_double),
                ("values", c_int32), ("pad_", c_int32), ("seed", c_uint64)]


# Every symbol declared in include/apss.h.
ABI_SYMBOLS = [
    "apss_version", "apss_strerror", "apss_last_error", "apss_opts_init", "apss_run",
    "apss_run_cb", "apss_run_noalloc", "apss_transpose", "apss_run_batched", "apss_result_free", "apss_index_create",
    "apss_index_insert", "apss_index_query", "apss_index_rows", "apss_index_destroy",
    "apss_server_create", "apss_server_submit", "apss_server_wait", "apss_server_cancel",
    "apss_server_destroy", "similarity_join_device", "apss_compat_last_count",
    "apss_load_text", "apss_save_text", "apss_save_binary", "apss_load_binary", "apss_csr_free",
    "apss_write_pairs", "apss_compare_pair_files",
]

_lib = None
_host = None


def lib():
    """Load libapss.so (raises OSError if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        L.apss_version.restype = c_char_p
        L.apss_strerror.restype = c_char_p
        L.apss_strerror.argtypes = [c_int32]
        L.apss_last_error.restype = c_char_p
        L.apss_opts_init.argtypes = [POINTER(Opts)]
        L.apss_run.argtypes = [POINTER(CSR), c_double, c_int32, POINTER(Opts), POINTER(Result)]
        L.apss_run.restype = c_int32
        L.apss_result_free.argtypes = [POINTER(Result)]
        _lib = L
    return _lib


def host_lib():
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB_PATH):
            raise OSError(f"{HOST_LIB_PATH} not built")
        H = ctypes.CDLL(HOST_LIB_PATH)
        H.apss_gen_alpha.restype = c_double
        H.apss_gen_alpha.argtypes = [c_int32, c_int32, c_double]
        H.apss_gen_rowptr.restype = c_int64
        H.apss_gen_rowptr.argtypes = [POINTER(GenSpec), c_void_p]
        H.apss_gen_fill.restype = c_int32
        H.apss_gen_fill.argtypes = [POINTER(GenSpec), c_void_p, c_void_p, c_void_p]
        _host = H
    return _host


class ApssError(RuntimeError):
    def __init__(self, code, detail=""):
        self.code = code
        msg = lib().apss_strerror(code).decode()
        super().__init__(f"apss error {code}: {msg}" + (f" ({detail})" if detail else ""))


# ---------------------------------------------------------------------------
# synthetic workloads

SEED_BASE = 0x41505353

WORKLOADS = {
    # name: n, d, mean nnz/row, kmin, kmax, zipf s, values, thresholds
    "small": dict(n=20_000, d=50_000, mean=40, kmin=5, kmax=500, zipf=1.1, values="uniform", t=(0.7,), index=0),
    "large": dict(n=1_000_000, d=1_000_000, mean=100, kmin=5, kmax=5000, zipf=1.05, values="lognormal",
                  t=(0.5, 0.7, 0.9), index=1),
    "batched": dict(n=None, d=100_000, mean=None, kmin=5, kmax=5000, zipf=1.1, values="uniform", t=(0.8,), index=2),
    "mixed": dict(n=500_000, d=500_000, mean=80, kmin=5, kmax=5000, zipf=1.2, values="uniform", t=(0.5,), index=3),
    "serving": dict(n=2_000_000, d=1_000_000, mean=60, kmin=5, kmax=5000, zipf=1.1, values="lognormal",
                    t=(0.6, 0.8), index=4),
}


def generate(n, d, mean, kmin=5, kmax=5000, zipf=1.1, values="uniform", seed=SEED_BASE, row_offset=0,
             alpha=None):
    """Deterministic CSR (row_ptr int64, col int32, val float32) from libapss_host.so."""
    H = host_lib()
    kmax = min(kmax, d)
    kmin = min(kmin, kmax)
    if alpha is None:
        alpha = H.apss_gen_alpha(kmin, kmax, float(mean))
    spec = GenSpec(n=n, row_offset=row_offset, d=d, kmin=kmin, kmax=kmax, alpha=alpha, zipf_s=zipf,
                   values=1 if values == "lognormal" else 0, pad_=0, seed=seed)
    rp = np.empty(n + 1, dtype=np.int64)
    nnz = H.apss_gen_rowptr(ctypes.byref(spec), rp.ctypes.data)
    if nnz < 0:
        raise ValueError("bad generator spec")
    col = np.empty(max(nnz, 1), dtype=np.int32)[:nnz]
    val = np.empty(max(nnz, 1), dtype=np.float32)[:nnz]
    if H.apss_gen_fill(ctypes.byref(spec), rp.ctypes.data, col.ctypes.data, val.ctypes.data) != 0:
        raise MemoryError("generator failed")
    return rp, col, val


def workload(name, n=None, seed_offset=0):
    """Matrix of a named workload, optionally with fewer rows (same distribution)."""
    w = WORKLOADS[name]
    n = n if n is not None else w["n"]
    return generate(n, w["d"], w["mean"], w["kmin"], w["kmax"], w["zipf"], w["values"],
                    seed=SEED_BASE + w["index"] + seed_offset) + (w["d"],)


def batched_instances(count=256, n_lo=500, n_hi=50_000, scale=1.0):
    """The batched workload: `count` independent instances, n log-uniform in
    [n_lo, n_hi] (times `scale`), mean nnz/row uniform in [20, 200], d = 1e5,
    Zipf(1.1) ids, uniform(0,1] values.  Instance k uses seed base + 2 + k
    (instance parameters from a generator seeded with base + 2)."""
    w = WORKLOADS["batched"]
    rng = np.random.default_rng(SEED_BASE + w["index"])
    ns = np.exp(rng.uniform(np.log(n_lo), np.log(n_hi), size=count))
    means = rng.uniform(20.0, 200.0, size=count)
    out = []
    for k in range(count):
        n = max(1, int(round(ns[k] * scale)))
        out.append(generate(n, w["d"], float(means[k]), w["kmin"], w["kmax"], w["zipf"], w["values"],
                            seed=SEED_BASE + w["index"] + k) + (w["d"],))
    return out


# ---------------------------------------------------------------------------
# calls

def make_csr(rp, col, val, d, on_device=False):
    return CSR(n_rows=len(rp) - 1, n_cols=d, nnz=len(col), row_ptr=rp.ctypes.data,
               col=col.ctypes.data if len(col) else None, val=val.ctypes.data if len(val) else None,
               on_device=1 if on_device else 0)


def make_opts(**kw):
    o = Opts()
    lib().apss_opts_init(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def result_pairs(res):
    """Copy a host-resident result into a numpy structured array."""
    n = res.count
    if n == 0 or not res.pairs:
        return np.zeros(0, dtype=PAIR_DTYPE)
    buf = (ctypes.c_char * (n * 12)).from_address(res.pairs)
    return np.frombuffer(buf, dtype=PAIR_DTYPE, count=n).copy()


def run(rp, col, val, d, t, p=0, **opts):
    """Self-join APSS through apss_run; returns (pairs, info)."""
    L = lib()
    X = make_csr(rp, col, val, d)
    o = make_opts(**opts)
    r = Result()
    rc = L.apss_run(ctypes.byref(X), float(t), int(p), ctypes.byref(o), ctypes.byref(r))
    if rc != APSS_OK:
        raise ApssError(rc, L.apss_last_error().decode())
    try:
        pairs = result_pairs(r) if not r.on_device else None
        info = {f: getattr(r, f) for f, _ in Result._fields_ if f not in ("pairs", "internal_")}
    finally:
        L.apss_result_free(ctypes.byref(r))
    return pairs, info


EMIT_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p)


def run_cb(rp, col, val, d, t, p, emit, **opts):
    """Streaming self-join through apss_run_cb: emit(pairs) receives each chunk
    as a PAIR_DTYPE view that is valid only during the call.  Returns info."""
    L = lib()
    L.apss_run_cb.argtypes = [POINTER(CSR), c_double, c_int32, POINTER(Opts), EMIT_FN, c_void_p,
                              POINTER(Result)]
    L.apss_run_cb.restype = c_int32
    X = make_csr(rp, col, val, d)
    o = make_opts(**opts)
    r = Result()
    errors = []

    def _emit(ptr, count, _user):
        try:
            buf = (ctypes.c_char * (count * 12)).from_address(ptr)
            emit(np.frombuffer(buf, dtype=PAIR_DTYPE, count=count))
        except Exception as e:  # noqa: BLE001 -- re-raised after the call
            errors.append(e)

    cb = EMIT_FN(_emit)
    rc = L.apss_run_cb(ctypes.byref(X), float(t), int(p), ctypes.byref(o), cb, None, ctypes.byref(r))
    if rc != APSS_OK:
        raise ApssError(rc, L.apss_last_error().decode())
    if errors:
        raise errors[0]
    return {f: getattr(r, f) for f, _ in Result._fields_ if f not in ("pairs", "internal_")}


def sort_pairs(pairs):
    return pairs[np.lexsort((pairs["j"], pairs["i"]))]


# ---------------------------------------------------------------------------
# minimal HIP runtime access (device buffers, streams, events) for callers that
# hand device-resident CSR to the library

class Hip:
    """ctypes view of the few libamdhip64 calls the benchmarks need."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipGetErrorString.restype = ctypes.c_char_p

    def _ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed: {self.L.hipGetErrorString(rc).decode()}")

    def malloc(self, nbytes):
        p = c_void_p()
        self._ok(self.L.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(int(nbytes), 1))), "hipMalloc")
        return p.value

    def free(self, p):
        self._ok(self.L.hipFree(c_void_p(p)), "hipFree")

    def h2d(self, dst, arr):
        a = np.ascontiguousarray(arr)
        self._ok(self.L.hipMemcpy(c_void_p(dst), c_void_p(a.ctypes.data), ctypes.c_size_t(a.nbytes), 1),
                 "hipMemcpy H2D")

    def d2h(self, arr, src):
        self._ok(self.L.hipMemcpy(c_void_p(arr.ctypes.data), c_void_p(src), ctypes.c_size_t(arr.nbytes), 2),
                 "hipMemcpy D2H")

    def d2d(self, dst, src, nbytes):
        self._ok(self.L.hipMemcpy(c_void_p(dst), c_void_p(src), ctypes.c_size_t(int(nbytes)), 3),
                 "hipMemcpy D2D")

    def sync(self):
        self._ok(self.L.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def stream(self):
        s = c_void_p()
        self._ok(self.L.hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        return s.value

    def event(self):
        e = c_void_p()
        self._ok(self.L.hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        return e.value

    def record(self, ev, stream):
        self._ok(self.L.hipEventRecord(c_void_p(ev), c_void_p(stream)), "hipEventRecord")

    def elapsed_ms(self, e0, e1):
        self._ok(self.L.hipEventSynchronize(c_void_p(e1)), "hipEventSynchronize")
        ms = ctypes.c_float()
        self._ok(self.L.hipEventElapsedTime(ctypes.byref(ms), c_void_p(e0), c_void_p(e1)), "hipEventElapsedTime")
        return float(ms.value)


class DeviceCSR:
    """A CSR matrix copied to device memory (freed with .release())."""

    def __init__(self, hip, rp, col, val, d):
        self.hip = hip
        self.n, self.nnz, self.d = len(rp) - 1, len(col), d
        self.rp = hip.malloc(rp.nbytes)
        self.col = hip.malloc(col.nbytes)
        self.val = hip.malloc(val.nbytes)
        hip.h2d(self.rp, rp)
        if len(col):
            hip.h2d(self.col, col)
            hip.h2d(self.val, val)

    def csr(self):
        return CSR(n_rows=self.n, n_cols=self.d, nnz=self.nnz, row_ptr=self.rp, col=self.col, val=self.val,
                   on_device=1)

    def release(self):
        for p in (self.rp, self.col, self.val):
            self.hip.free(p)


def transpose(rp, col, val, d, p=0, **opts):
    """Inverted index (CSC of the ids >= p) built on the device through
    apss_transpose: (col_ptr int64[d+1], row int32, val float32)."""
    L = _bind_more(lib())
    X = make_csr(rp, col, val, d)
    o = make_opts(**opts)
    m = int(np.count_nonzero(np.asarray(col) >= p))
    cptr = np.zeros(d + 1, np.int64)
    crow = np.zeros(max(m, 1), np.int32)
    cval = np.zeros(max(m, 1), np.float32)
    _check(L.apss_transpose(ctypes.byref(X), int(p), ctypes.byref(o), cptr.ctypes.data,
                            crow.ctypes.data, cval.ctypes.data))
    return cptr, crow[:m], cval[:m]


def run_csr(X, t, p=0, **opts):
    """apss_run on a prepared CSR struct; returns the Result (caller frees it)."""
    L = lib()
    o = make_opts(**opts)
    r = Result()
    rc = L.apss_run(ctypes.byref(X), float(t), int(p), ctypes.byref(o), ctypes.byref(r))
    if rc != APSS_OK:
        raise ApssError(rc, L.apss_last_error().decode())
    return r


# ---------------------------------------------------------------------------
# batched, index, server and compat entry points

def _bind_more(L):
    if getattr(L, "_more_bound", False):
        return L
    L.apss_run_batched.argtypes = [ctypes.c_void_p, c_int32, ctypes.c_void_p, c_int32, POINTER(Opts), c_void_p]
    L.apss_run_batched.restype = c_int32
    L.apss_index_create.argtypes = [POINTER(CSR), c_int32, POINTER(Opts), POINTER(c_void_p)]
    L.apss_index_create.restype = c_int32
    L.apss_index_insert.argtypes = [c_void_p, POINTER(CSR)]
    L.apss_index_insert.restype = c_int32
    L.apss_index_query.argtypes = [c_void_p, POINTER(CSR), c_double, POINTER(Result)]
    L.apss_index_query.restype = c_int32
    L.apss_index_rows.argtypes = [c_void_p]
    L.apss_index_rows.restype = c_int64
    L.apss_index_destroy.argtypes = [c_void_p]
    L.apss_server_create.argtypes = [c_void_p, c_int32, c_int64, c_int32, POINTER(c_void_p)]
    L.apss_server_create.restype = c_int32
    L.apss_server_submit.argtypes = [c_void_p, POINTER(CSR), c_double]
    L.apss_server_submit.restype = c_int64
    L.apss_server_wait.argtypes = [c_void_p, c_int64, POINTER(Result), POINTER(c_double)]
    L.apss_server_wait.restype = c_int32
    L.apss_server_cancel.argtypes = [c_void_p, c_int64]
    L.apss_server_cancel.restype = c_int32
    L.apss_server_destroy.argtypes = [c_void_p]
    L.apss_compat_last_count.restype = c_int64
    L.apss_run_noalloc.argtypes = [POINTER(CSR), c_double, c_int32, POINTER(Opts), c_void_p, c_int64,
                                   POINTER(c_int64), POINTER(Result)]
    L.apss_run_noalloc.restype = c_int32
    L.apss_transpose.argtypes = [POINTER(CSR), c_int32, POINTER(Opts), c_void_p, c_void_p, c_void_p]
    L.apss_transpose.restype = c_int32
    L._more_bound = True
    return L


def _check(rc):
    if rc != APSS_OK:
        raise ApssError(rc, lib().apss_last_error().decode())


def _take_result(r):
    pairs = result_pairs(r)
    info = {f: getattr(r, f) for f, _ in Result._fields_ if f not in ("pairs", "internal_")}
    lib().apss_result_free(ctypes.byref(r))
    return pairs, info


def run_batched(instances, ts, p=0, **opts):
    """instances: list of (rp, col, val, d); ts: thresholds.  Returns list of (pairs, info)."""
    L = _bind_more(lib())
    k = len(instances)
    keep = [(np.ascontiguousarray(rp, np.int64), np.ascontiguousarray(c, np.int32),
             np.ascontiguousarray(v, np.float32), d) for rp, c, v, d in instances]
    csrs = (CSR * max(k, 1))(*[make_csr(*x) for x in keep])
    ptrs = (ctypes.c_void_p * max(k, 1))(*[ctypes.addressof(csrs[i]) for i in range(k)])
    tarr = (c_double * max(k, 1))(*[float(x) for x in ts])
    res = (Result * max(k, 1))()
    o = make_opts(**opts)
    _check(L.apss_run_batched(ptrs, k, tarr, int(p), ctypes.byref(o), res))
    return [_take_result(res[i]) for i in range(k)]


class Index:
    """Corpus index resident in device memory (apss_index_*)."""

    def __init__(self, rp, col, val, d, p=0, **opts):
        self.L = _bind_more(lib())
        self.d = d
        X = make_csr(rp, col, val, d)
        o = make_opts(**opts)
        h = c_void_p()
        _check(self.L.apss_index_create(ctypes.byref(X), int(p), ctypes.byref(o), ctypes.byref(h)))
        self.h = h

    def query(self, rp, col, val, t):
        X = make_csr(rp, col, val, self.d)
        r = Result()
        _check(self.L.apss_index_query(self.h, ctypes.byref(X), float(t), ctypes.byref(r)))
        return _take_result(r)

    def insert(self, rp, col, val):
        X = make_csr(rp, col, val, self.d)
        _check(self.L.apss_index_insert(self.h, ctypes.byref(X)))

    def rows(self):
        return self.L.apss_index_rows(self.h)

    def close(self):
        if self.h:
            self.L.apss_index_destroy(self.h)
            self.h = None


class Server:
    """Continuous-batching server over an Index (apss_server_*)."""

    def __init__(self, index, slots=2, mem_budget=0, max_queue=0):
        self.L = _bind_more(lib())
        self.index = index
        h = c_void_p()
        _check(self.L.apss_server_create(index.h, int(slots), int(mem_budget), int(max_queue), ctypes.byref(h)))
        self.h = h
        self._keep = {}

    def submit(self, rp, col, val, t):
        X = make_csr(rp, col, val, self.index.d)
        tk = self.L.apss_server_submit(self.h, ctypes.byref(X), float(t))
        if tk < 0:
            raise ApssError(int(tk), lib().apss_last_error().decode())
        return tk

    def wait(self, ticket):
        pairs, _, lat = self.wait_info(ticket)
        return pairs, lat

    def wait_info(self, ticket, keep_pairs=True):
        """(pairs, counters of the launch that served the request, latency ms);
        with keep_pairs=False the pairs are released uncopied (pairs = count)."""
        r = Result()
        lat = c_double()
        rc = self.L.apss_server_wait(self.h, int(ticket), ctypes.byref(r), ctypes.byref(lat))
        _check(rc)
        if keep_pairs:
            pairs, info = _take_result(r)
        else:
            info = {f: getattr(r, f) for f, _ in Result._fields_ if f not in ("pairs", "internal_")}
            pairs = r.count
            lib().apss_result_free(ctypes.byref(r))
        return pairs, info, lat.value

    def cancel(self, ticket):
        return self.L.apss_server_cancel(self.h, int(ticket))

    def close(self):
        if self.h:
            self.L.apss_server_destroy(self.h)
            self.h = None


def sell32_partition(rp, col, val):
    """SELL-32 arrays (dim, feature, tile_start) of a CSR block, as main.cu builds
    them per partition (main.cu:286-339): slices of 32 rows stored column-major,
    width = the largest row of the slice, padding id -1."""
    n = len(rp) - 1
    tiles = (n + 31) // 32
    starts = np.zeros(tiles + 1, np.int32)
    lens = np.diff(rp)
    for tt in range(tiles):
        w = int(lens[tt * 32:(tt + 1) * 32].max()) if n else 0
        starts[tt + 1] = starts[tt] + 32 * w
    dim = np.full(max(int(starts[-1]), 1), -1, np.int32)
    feat = np.zeros(max(int(starts[-1]), 1), np.float32)
    for r in range(n):
        base = starts[r // 32] + r % 32
        k = int(lens[r])
        dim[base + 32 * np.arange(k)] = col[rp[r]:rp[r + 1]]
        feat[base + 32 * np.arange(k)] = val[rp[r]:rp[r + 1]]
    return dim, feat, starts


def similarity_join_device(parts):
    """Call the reference-signature shim with 8 partitions (list of (rp, col, val))."""
    L = _bind_more(lib())
    assert len(parts) == 8
    sells = [sell32_partition(*p) for p in parts]
    args = []
    for dim, feat, starts in sells:
        args += [dim.ctypes.data_as(c_void_p), feat.ctypes.data_as(c_void_p), starts.ctypes.data_as(c_void_p)]
    norms = [np.zeros(max(len(p[0]) - 1, 1), np.float32) for p in parts]
    args += [x.ctypes.data_as(c_void_p) for x in norms]
    for (rp, col, val), (dim, feat, starts) in zip(parts, sells):
        args += [ctypes.c_int(len(rp) - 1), ctypes.c_int(len(dim)), ctypes.c_int(len(starts) - 1)]
    args += [c_void_p(None), c_void_p(None), c_void_p(None)]
    L.similarity_join_device.restype = None
    L.similarity_join_device(*args)
    return L.apss_compat_last_count()


# ---------------------------------------------------------------------------
# files (host-only entry points)

def _bind_io(L):
    if getattr(L, "_io_bound", False):
        return L
    L.apss_load_text.argtypes = [c_char_p, c_int64, POINTER(CSR)]
    L.apss_load_text.restype = c_int32
    L.apss_save_text.argtypes = [c_char_p, POINTER(CSR)]
    L.apss_save_text.restype = c_int32
    L.apss_save_binary.argtypes = [c_char_p, POINTER(CSR)]
    L.apss_save_binary.restype = c_int32
    L.apss_load_binary.argtypes = [c_char_p, POINTER(CSR)]
    L.apss_load_binary.restype = c_int32
    L.apss_csr_free.argtypes = [POINTER(CSR)]
    L.apss_csr_free.restype = None
    L.apss_write_pairs.argtypes = [c_char_p, c_void_p, c_int64]
    L.apss_write_pairs.restype = c_int32
    L.apss_compare_pair_files.argtypes = [c_char_p, c_char_p, POINTER(c_int64), POINTER(c_int64)]
    L.apss_compare_pair_files.restype = c_int64
    L._io_bound = True
    return L


def _take_csr(L, X):
    """Copy a library-owned CSR into numpy arrays and release it."""
    try:
        n, nnz = X.n_rows, X.nnz
        rp = np.ctypeslib.as_array((c_int64 * (n + 1)).from_address(X.row_ptr)).copy()
        col = (np.ctypeslib.as_array((c_int32 * nnz).from_address(X.col)).copy() if nnz
               else np.zeros(0, np.int32))
        val = (np.ctypeslib.as_array((c_float * nnz).from_address(X.val)).copy() if nnz
               else np.zeros(0, np.float32))
        return rp, col, val, int(X.n_cols)
    finally:
        L.apss_csr_free(ctypes.byref(X))


def load_text(path, n_cols=0):
    """Rows of a file in the reference's text format -> (row_ptr, col, val, d)."""
    L = _bind_io(lib())
    X = CSR()
    _check(L.apss_load_text(os.fsencode(path), int(n_cols), ctypes.byref(X)))
    return _take_csr(L, X)


def save_text(path, rp, col, val, d):
    L = _bind_io(lib())
    rp, col, val = (np.ascontiguousarray(rp, np.int64), np.ascontiguousarray(col, np.int32),
                    np.ascontiguousarray(val, np.float32))
    _check(L.apss_save_text(os.fsencode(path), ctypes.byref(make_csr(rp, col, val, d))))


def save_binary(path, rp, col, val, d):
    L = _bind_io(lib())
    rp, col, val = (np.ascontiguousarray(rp, np.int64), np.ascontiguousarray(col, np.int32),
                    np.ascontiguousarray(val, np.float32))
    _check(L.apss_save_binary(os.fsencode(path), ctypes.byref(make_csr(rp, col, val, d))))


def load_binary(path):
    L = _bind_io(lib())
    X = CSR()
    _check(L.apss_load_binary(os.fsencode(path), ctypes.byref(X)))
    return _take_csr(L, X)


def write_pairs(path, pairs):
    """Write a PAIR_DTYPE array as "i j score" lines."""
    L = _bind_io(lib())
    pairs = np.ascontiguousarray(pairs, dtype=PAIR_DTYPE)
    _check(L.apss_write_pairs(os.fsencode(path), pairs.ctypes.data if len(pairs) else None, len(pairs)))


def compare_pair_files(a, b):
    """(only_in_a, only_in_b) pair counts of two pair files."""
    L = _bind_io(lib())
    oa, ob = c_int64(0), c_int64(0)
    rc = L.apss_compare_pair_files(os.fsencode(a), os.fsencode(b), ctypes.byref(oa), ctypes.byref(ob))
    if rc < 0:
        raise ApssError(int(rc), L.apss_last_error().decode())
    return oa.value, ob.value
