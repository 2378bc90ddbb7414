"""apss_run_noalloc (no device allocation: scratch array + caller's buffer)
against apss_run, through the C-ABI, on two calls in a row (the second with a
smaller input), exact pair set and scores within 1e-6 relative."""
import ctypes

import numpy as np
import pytest

import apss
import oracle
from helpers import check_pairs

pytestmark = pytest.mark.gpu


def _noalloc(hip, X, t, capacity):
    L = apss._bind_more(apss.lib())
    buf = hip.malloc(max(capacity, 1) * apss.PAIR_DTYPE.itemsize)
    try:
        cnt = ctypes.c_int64(-1)
        r = apss.Result()
        o = apss.make_opts()
        rc = L.apss_run_noalloc(ctypes.byref(X), float(t), 0, ctypes.byref(o), buf, capacity,
                                ctypes.byref(cnt), ctypes.byref(r))
        pairs = np.zeros(min(max(cnt.value, 0), capacity), dtype=apss.PAIR_DTYPE)
        if rc == apss.APSS_OK and len(pairs):
            hip.d2h(pairs, buf)
        return rc, cnt.value, pairs, r
    finally:
        hip.free(buf)


@pytest.mark.parametrize("t", [0.5])
def test_noalloc_matches_apss_run_twice(t):
    hip = apss.Hip()
    for n in (9000, 2500):  # the second call is smaller: the shared array keeps old contents
        rp, col, val, d = apss.workload("small", n=n, seed_offset=n)
        X = apss.DeviceCSR(hip, rp, col, val, d)
        try:
            print(f"noalloc n={n} t={t}", flush=True)
            ref, _ = apss.run(rp, col, val, d, t, 0)
            rc, cnt, got, info = _noalloc(hip, X.csr(), t, len(ref) + 1000)
            assert rc == apss.APSS_OK, apss.lib().apss_last_error()
            assert cnt == len(ref) == info.count
            assert np.array_equal(apss.sort_pairs(got), apss.sort_pairs(ref))
            check_pairs(got, oracle.apss_fp64(rp, col, val, d, t))
            # a buffer too small: the count is the room the result needs
            rc, cnt2, _, _ = _noalloc(hip, X.csr(), t, len(ref) // 2)
            assert rc == apss.APSS_E_NOMEM and cnt2 == len(ref)
        finally:
            X.release()


def test_noalloc_rejects_what_it_cannot_hold():
    hip = apss.Hip()
    L = apss._bind_more(apss.lib())
    n, per_row = 1 << 20, 80  # 8.4e7 nonzeros: beyond the scratch array
    X = apss.CSR(n_rows=n, n_cols=1 << 20, nnz=n * per_row, row_ptr=1, col=1, val=1, on_device=1)
    cnt = ctypes.c_int64(-1)
    o = apss.make_opts()
    rc = L.apss_run_noalloc(ctypes.byref(X), 0.5, 0, ctypes.byref(o), 1, 16, ctypes.byref(cnt), None)
    assert rc == apss.APSS_E_NOMEM and b"scratch" in L.apss_last_error()
    del hip
