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


# Emitted scores of the BF16 head on unit rows: the head error is at most
# c (|x_H||y_H|) <= c with c = 2^-7 + 2^-16 + 2 gamma_128 (7.87e-3), or, with no
# negative value, c / (1 - c) s <= 7.93e-3 for s <= 1, plus the tail's FP32
# and fixed-point terms (below 1e-5).
BF16_ATOL = 8.0e-3


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


K1 = np.uint64(0x9E3779B97F4A7C15)
K2 = np.uint64(0xC2B2AE3D27D4EB4F)


def sample_rows(n, k, seed):
    return np.sort(np.random.default_rng(seed).choice(n, k, replace=False)).astype(np.int64)


def oracle_pairs_touching(rp, col, val, d, sample, t):
    """Self-join pairs (i < j) touching the sampled rows, from the FP64 oracle
    (the sample rows queried against all rows): sorted keys i*2^32+j, scores."""
    q_rp = np.zeros(len(sample) + 1, np.int64)
    q_rp[1:] = np.cumsum(rp[sample + 1] - rp[sample])
    q_col = np.concatenate([col[rp[r]:rp[r + 1]] for r in sample])
    q_val = np.concatenate([val[rp[r]:rp[r + 1]] for r in sample])
    exp = oracle.query_fp64((q_rp, q_col, q_val), (rp, col, val), d, t)
    s = sample[exp["i"].astype(np.int64)]
    c = exp["j"].astype(np.int64)
    keep = s != c
    s, c, sc = s[keep], c[keep], exp["score"][keep]
    key = np.minimum(s, c) * (1 << 32) + np.maximum(s, c)
    key, first = np.unique(key, return_index=True)
    return key, sc[first]


def stream_touching(rp, col, val, d, sample, t, p=0, **opts):
    """Runs the self-join through apss_run_cb; returns (count, order-free
    checksum of all pairs, sorted keys and scores of the pairs touching the
    sample, counters)."""
    mask = np.zeros(len(rp) - 1, bool)
    mask[sample] = True
    acc = {"count": 0, "sum": 0, "keys": [], "scores": []}

    def emit(chunk):
        i, j = chunk["i"], chunk["j"]
        assert np.all(i < j)
        acc["count"] += len(chunk)
        h = (i.astype(np.uint64) * K1) ^ (j.astype(np.uint64) * K2)
        acc["sum"] = (acc["sum"] + int(h.sum(dtype=np.uint64))) % (1 << 64)
        m = mask[i] | mask[j]
        if m.any():
            acc["keys"].append(i[m].astype(np.int64) * (1 << 32) + j[m].astype(np.int64))
            acc["scores"].append(chunk["score"][m].astype(np.float64))

    info = apss.run_cb(rp, col, val, d, t, p, emit, **opts)
    assert info["count"] == acc["count"]
    keys = np.concatenate(acc["keys"]) if acc["keys"] else np.zeros(0, np.int64)
    scores = np.concatenate(acc["scores"]) if acc["scores"] else np.zeros(0)
    o = np.argsort(keys)
    return acc["count"], int(acc["sum"]), keys[o], scores[o], info


def check_touching(keys, scores, exp_keys, exp_scores, rel=None, atol=None):
    assert len(keys) == len(exp_keys), f"sample pairs {len(keys)} != oracle {len(exp_keys)}"
    assert np.array_equal(keys, exp_keys), "sample pair sets differ"
    if len(exp_keys):
        if rel is not None:
            err = np.abs(scores - exp_scores) / np.abs(exp_scores)
            assert err.max() <= rel, err.max()
        if atol is not None:
            assert np.abs(scores - exp_scores).max() <= atol
    return len(exp_keys)
