"""Parity at the full size of the large workload (n = d = 1e6, ~1e8 nonzeros,
Zipf(1.05), lognormal values): the whole pair set is streamed through
apss_run_cb, and

  - every pair that touches a random sample of rows is compared with the
    FP64 oracle's answer for those rows against all 1e6 rows (exact set,
    scores within 1e-6 relative, or the stated BF16 bound);
  - the pair count and an order-independent checksum of the whole set are
    identical for the exact int8 head, the plain sparse path and the BF16
    head (the pair set does not depend on how the scores are formed).
"""
import numpy as np
import pytest

import apss
import oracle

pytestmark = pytest.mark.gpu

N_SAMPLE = 48
K1 = np.uint64(0x9E3779B97F4A7C15)
K2 = np.uint64(0xC2B2AE3D27D4EB4F)
_sets = {}


@pytest.fixture(scope="module")
def large():
    rp, col, val, d = apss.workload("large")
    n = len(rp) - 1
    rng = np.random.default_rng(1234)
    sample = np.sort(rng.choice(n, N_SAMPLE, replace=False)).astype(np.int64)
    return rp, col, val, d, sample


_expected = {}


def expected_for_sample(large, t):
    """Self-join pairs (i < j) touching the sample, from the FP64 oracle."""
    if t in _expected:
        return _expected[t]
    rp, col, val, d, sample = large
    q_rp = np.zeros(len(sample) + 1, np.int64)
    q_rp[1:] = np.cumsum(rp[sample + 1] - rp[sample])
    q_col = np.concatenate([col[rp[r]:rp[r + 1]] for r in sample])
    q_val = np.concatenate([val[rp[r]:rp[r + 1]] for r in sample])
    exp = oracle.query_fp64((q_rp, q_col, q_val), (rp, col, val), d, t)
    s = sample[exp["i"].astype(np.int64)]
    c = exp["j"].astype(np.int64)
    keep = s != c
    s, c, sc = s[keep], c[keep], exp["score"][keep]
    i, j = np.minimum(s, c), np.maximum(s, c)
    key = i * (1 << 32) + j
    key, first = np.unique(key, return_index=True)
    _expected[t] = (key, sc[first])
    return _expected[t]


def stream(large, t, **opts):
    rp, col, val, d, sample = large
    mask = np.zeros(len(rp) - 1, bool)
    mask[sample] = True
    acc = {"count": 0, "sum": 0, "keys": [], "scores": []}

    def emit(chunk):
        i = chunk["i"]
        j = chunk["j"]
        assert np.all(i < j)
        acc["count"] += len(chunk)
        h = (i.astype(np.uint64) * K1) ^ (j.astype(np.uint64) * K2)
        acc["sum"] = (acc["sum"] + int(h.sum(dtype=np.uint64))) % (1 << 64)  # sum mod 2^64
        m = mask[i] | mask[j]
        if m.any():
            acc["keys"].append(i[m].astype(np.int64) * (1 << 32) + j[m].astype(np.int64))
            acc["scores"].append(chunk["score"][m].astype(np.float64))

    info = apss.run_cb(rp, col, val, d, t, 0, emit, **opts)
    assert info["count"] == acc["count"]
    keys = np.concatenate(acc["keys"]) if acc["keys"] else np.zeros(0, np.int64)
    scores = np.concatenate(acc["scores"]) if acc["scores"] else np.zeros(0)
    o = np.argsort(keys)
    return acc["count"], int(acc["sum"]), keys[o], scores[o], info


def check_sample(large, t, keys, scores, rel=None, atol=None):
    ek, es = expected_for_sample(large, t)
    assert len(keys) == len(ek), f"sample pairs {len(keys)} != oracle {len(ek)}"
    assert np.array_equal(keys, ek), "sample pair sets differ"
    if len(ek):
        if rel is not None:
            err = np.abs(scores - es) / np.abs(es)
            assert err.max() <= rel, err.max()
        if atol is not None:
            assert np.abs(scores - es).max() <= atol
    return len(ek)


@pytest.mark.parametrize("form", ["int8_head", "plain", "bf16_head"])
def test_large_t05_pair_set(large, form):
    opts = {"int8_head": dict(mfma_head=0), "plain": dict(mfma_head=-1),
            "bf16_head": dict(mfma_head=0, precision=apss.PREC_BF16_HEAD)}[form]
    count, csum, keys, scores, info = stream(large, 0.5, **opts)
    print(f"{form}: {count} pairs, scan {info['ms_scan']:.0f} ms, rescored {info['band']}", flush=True)
    if form == "bf16_head":
        # BF16 head bound for unit rows (at most 7.93e-3) + the FP32 tail bound
        check_sample(large, 0.5, keys, scores, atol=8.1e-3)
    else:
        check_sample(large, 0.5, keys, scores, rel=1e-6)
    if _sets:
        ref = next(iter(_sets.values()))
        assert (count, csum) == ref, f"{form} pair set differs from {next(iter(_sets))}"
    _sets[form] = (count, csum)


def test_large_t09_pair_set(large):
    count, _, keys, scores, info = stream(large, 0.9)
    assert check_sample(large, 0.9, keys, scores, rel=1e-6) >= 0
    assert count > 0


def digest(pairs):
    """Order-independent digest of a pair array: count and the sum mod 2^64
    of a hash of (i, j, score bits)."""
    i = pairs["i"].astype(np.uint64)
    j = pairs["j"].astype(np.uint64)
    s = pairs["score"].view(np.uint32).astype(np.uint64)
    h = (i * K1) ^ (j * K2) ^ (s * np.uint64(0x165667B19E3779F9))
    return len(pairs), int(h.sum(dtype=np.uint64))


def test_host_result_above_2gib_matches_device_result(large):
    # A host result above 2 GiB comes back through pinned staging blocks into
    # a huge-page-backed array (the smaller ones land in pinned blocks
    # directly); it must hold the same pairs and scores as the result left in
    # device memory.  The matrix goes in from host memory, through the staged
    # upload.
    rp, col, val, d, _ = large
    t = 0.7
    host_pairs, info = apss.run(rp, col, val, d, t, 0, mfma_head=0)
    print(f"host result: {info['count']} pairs, download {info['ms_download']:.0f} ms", flush=True)
    assert info["count"] * apss.PAIR_DTYPE.itemsize > (2 << 30)
    hip = apss.Hip()
    X = apss.make_csr(rp, col, val, d)
    r = apss.run_csr(X, t, 0, output_on_device=1, mfma_head=0)
    try:
        assert r.on_device == 1
        dev = np.zeros(r.count, apss.PAIR_DTYPE)
        hip.d2h(dev, r.pairs)
    finally:
        apss.lib().apss_result_free(apss.ctypes.byref(r))
    assert digest(host_pairs) == digest(dev)
