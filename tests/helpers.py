"""Shared helpers of the test suite (matrices, pair-set comparison)."""
import numpy as np

import apss
import oracle


def tiny_matrix(rows, d):
    """CSR from a list of {col: value} dicts (values as given, float32)."""
    rp = [0]
    col, val = [], []
    for r in rows:
        for c in sorted(r):
            col.append(c)
            val.append(r[c])
        rp.append(len(col))
    return (np.array(rp, np.int64), np.array(col, np.int32), np.array(val, np.float32), d)


def subset(rp, col, val, n):
    return rp[:n + 1].copy(), col[:rp[n]].copy(), val[:rp[n]].copy()


def check_pairs(got, exp, rel=1e-6):
    """got: PAIR_DTYPE (any order); exp: oracle PAIR64 sorted by (i,j).
    Exact pair-set equality and scores within `rel` relative."""
    g = apss.sort_pairs(got)
    assert len(g) == len(exp), f"pair count {len(g)} != oracle {len(exp)}"
    if len(g) == 0:
        return 0.0
    assert np.array_equal(g["i"], exp["i"]) and np.array_equal(g["j"], exp["j"]), "pair sets differ"
    err = np.abs(g["score"].astype(np.float64) - exp["score"]) / np.maximum(np.abs(exp["score"]), 1e-30)
    assert err.max() <= rel, f"max relative score error {err.max():.3e} > {rel}"
    return float(err.max())


def check_pairs_abs(got, exp, atol):
    """Exact pair-set equality and scores within `atol` absolute (BF16 head)."""
    g = apss.sort_pairs(got)
    assert len(g) == len(exp), f"pair count {len(g)} != oracle {len(exp)}"
    if len(g) == 0:
        return 0.0
    assert np.array_equal(g["i"], exp["i"]) and np.array_equal(g["j"], exp["j"]), "pair sets differ"
    err = np.abs(g["score"].astype(np.float64) - exp["score"])
    assert err.max() <= atol, f"max absolute score error {err.max():.3e} > {atol}"
    return float(err.max())
