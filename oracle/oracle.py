"""ctypes binding of liboracle.so -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
import this module, and only as the checker or the timed CPU baseline.
"""
import ctypes
import os
from ctypes import POINTER, c_double, c_int32, c_int64, c_uint32, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class _Pairs(ctypes.Structure):
    _fields_ = [("count", c_int64), ("i", POINTER(c_uint32)), ("j", POINTER(c_uint32)),
                ("s", POINTER(c_double))]


class _HStats(ctypes.Structure):
    _fields_ = [("candidates", c_int64), ("survivors", c_int64), ("head_gathers", c_int64),
                ("search_seconds", c_double)]


PAIR64 = np.dtype([("i", "<u4"), ("j", "<u4"), ("score", "<f8")])
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} not built (make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_apss_fp64.argtypes = [c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_double, POINTER(_Pairs)]
        L.oracle_query_fp64.argtypes = [c_int64, c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p,
                                        c_void_p, c_int64, c_double, POINTER(_Pairs)]
        L.oracle_hybrid_fp32.argtypes = [c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_double, c_int32,
                                         c_void_p, c_int64, POINTER(_Pairs), POINTER(_HStats)]
        L.oracle_transpose.argtypes = [c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_int32, c_void_p,
                                       c_void_p, c_void_p]
        L.oracle_pair_fp64.restype = c_double
        L.oracle_pair_fp64.argtypes = [c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_int64]
        L.oracle_inverted_search_fp32.argtypes = [c_int64, c_int64, c_void_p, c_void_p, c_void_p, c_double,
                                                  c_void_p, c_int64, c_double, POINTER(c_int64),
                                                  POINTER(c_int64), POINTER(c_double)]
        L.oracle_pairs_free.argtypes = [POINTER(_Pairs)]
        L.oracle_num_threads.restype = c_int32
        _lib = L
    return _lib


def _take(p):
    n = p.count
    out = np.zeros(n, dtype=PAIR64)
    if n:
        out["i"] = np.ctypeslib.as_array(p.i, shape=(n,))
        out["j"] = np.ctypeslib.as_array(p.j, shape=(n,))
        out["score"] = np.ctypeslib.as_array(p.s, shape=(n,))
    lib().oracle_pairs_free(ctypes.byref(p))
    return out


def _c(a):
    a = np.ascontiguousarray(a)
    return a, (a.ctypes.data if a.size else None)


def apss_fp64(rp, col, val, d, t):
    """Exact pairs (i<j, s>=t) sorted by (i, j) with FP64 scores."""
    rp, prp = _c(np.asarray(rp, np.int64)); col, pc = _c(np.asarray(col, np.int32)); val, pv = _c(np.asarray(val, np.float32))
    p = _Pairs()
    if lib().oracle_apss_fp64(len(rp) - 1, d, prp, pc, pv, float(t), ctypes.byref(p)) != 0:
        raise MemoryError("oracle_apss_fp64 failed")
    return _take(p)


def query_fp64(q, c, d, t):
    """Exact (query, corpus) pairs with s >= t; q and c are (rp, col, val)."""
    qa = [_c(np.asarray(x, dt)) for x, dt in zip(q, (np.int64, np.int32, np.float32))]
    ca = [_c(np.asarray(x, dt)) for x, dt in zip(c, (np.int64, np.int32, np.float32))]
    p = _Pairs()
    rc = lib().oracle_query_fp64(len(qa[0][0]) - 1, qa[0][1], qa[1][1], qa[2][1], len(ca[0][0]) - 1,
                                 ca[0][1], ca[1][1], ca[2][1], d, float(t), ctypes.byref(p))
    if rc != 0:
        raise MemoryError("oracle_query_fp64 failed")
    return _take(p)


def hybrid_fp32(rp, col, val, d, t, p, rows=None):
    """Reference hybrid on the CPU; returns (pairs, stats dict)."""
    rp, prp = _c(np.asarray(rp, np.int64)); col, pc = _c(np.asarray(col, np.int32)); val, pv = _c(np.asarray(val, np.float32))
    if rows is not None:
        rows, prows = _c(np.asarray(rows, np.int64))
        nr = len(rows)
    else:
        prows, nr = None, 0
    out = _Pairs()
    st = _HStats()
    if lib().oracle_hybrid_fp32(len(rp) - 1, d, prp, pc, pv, float(t), int(p), prows, nr,
                                ctypes.byref(out), ctypes.byref(st)) != 0:
        raise MemoryError("oracle_hybrid_fp32 failed")
    return _take(out), dict(candidates=st.candidates, survivors=st.survivors, head_gathers=st.head_gathers,
                            search_seconds=st.search_seconds)


def inverted_search_fp32(rp, col, val, d, t, rows=None, budget_s=0.0):
    """Serial inverted-index search; returns (pair count, search seconds, rows
    searched).  With budget_s > 0 the rows are taken in order until that many
    seconds have passed."""
    rp, prp = _c(np.asarray(rp, np.int64)); col, pc = _c(np.asarray(col, np.int32)); val, pv = _c(np.asarray(val, np.float32))
    if rows is not None:
        rows, prows = _c(np.asarray(rows, np.int64))
        nr = len(rows)
    else:
        prows, nr = None, 0
    cnt, sec, done = c_int64(0), c_double(0.0), c_int64(0)
    if lib().oracle_inverted_search_fp32(len(rp) - 1, d, prp, pc, pv, float(t), prows, nr, float(budget_s),
                                         ctypes.byref(done), ctypes.byref(cnt), ctypes.byref(sec)) != 0:
        raise MemoryError("oracle_inverted_search_fp32 failed")
    return cnt.value, sec.value, done.value


def transpose(rp, col, val, d, p=0):
    """Stable CSR->CSC of ids >= p: (cptr int64[d+1], crow int32, cval float32)."""
    rp, prp = _c(np.asarray(rp, np.int64)); col, pc = _c(np.asarray(col, np.int32)); val, pv = _c(np.asarray(val, np.float32))
    m = int(np.count_nonzero(col >= p))
    cptr = np.zeros(d + 1, np.int64)
    crow = np.zeros(max(m, 1), np.int32)
    cval = np.zeros(max(m, 1), np.float32)
    lib().oracle_transpose(len(rp) - 1, d, prp, pc, pv, int(p), cptr.ctypes.data, crow.ctypes.data,
                           cval.ctypes.data)
    return cptr, crow[:m], cval[:m]


def pair_fp64(rp, col, val, i, j):
    """FP64 score of one pair in ascending feature order."""
    a0, a1, b0, b1 = rp[i], rp[i + 1], rp[j], rp[j + 1]
    ac = np.ascontiguousarray(col[a0:a1]); av = np.ascontiguousarray(val[a0:a1])
    bc = np.ascontiguousarray(col[b0:b1]); bv = np.ascontiguousarray(val[b0:b1])
    return lib().oracle_pair_fp64(ac.ctypes.data, av.ctypes.data, len(ac), bc.ctypes.data, bv.ctypes.data, len(bc))


def num_threads():
    return lib().oracle_num_threads()
