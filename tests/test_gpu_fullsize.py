"""Parity at the sizes the workloads are quoted on, beside the headline test
(test_gpu_large.py): the mixed-precision workload at 5e5 rows (int8 and BF16
heads), the large workload at t = 0.7 and, on a 1e5-row slice, at the split
p = 4096 (the reference's prefix-norm filter and forward-list head dot), the
256-instance batched workload with the autotuner on, the full small workload
with the default tiering, a query whose result outgrows the first output
buffer, the device-built inverted index against the oracle's transpose, and the
2e6-row serving corpus queried through the index and the server.
Pairs touching a seeded sample of rows are compared with the FP64 oracle's
answer for those rows against all rows (exact set; scores within 1e-6
relative, or the stated BF16 bound); the whole sets are compared across forms
by count and an order-free checksum."""
import numpy as np
import pytest

import apss
import oracle
from helpers import (check_pairs, check_touching, oracle_pairs_touching, sample_rows,
                     stream_touching, subset)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mixed():
    rp, col, val, d = apss.workload("mixed")
    sample = sample_rows(len(rp) - 1, 40, 4321)
    return rp, col, val, d, sample, oracle_pairs_touching(rp, col, val, d, sample, 0.5)


_mixed_sets = {}


@pytest.mark.parametrize("form", ["int8_head", "bf16_head"])
def test_mixed_workload_full_size(mixed, form):
    rp, col, val, d, sample, (ek, es) = mixed
    opts = {"int8_head": dict(mfma_head=0), "bf16_head": dict(mfma_head=0, precision=apss.PREC_BF16_HEAD)}[form]
    count, csum, keys, scores, info = stream_touching(rp, col, val, d, sample, 0.5, **opts)
    print(f"mixed {form}: {count} pairs, rescored {info['band']}", flush=True)
    if form == "bf16_head":
        check_touching(keys, scores, ek, es, atol=8.1e-3)
    else:
        check_touching(keys, scores, ek, es, rel=1e-6)
    assert len(ek) > 0
    for other, ref in _mixed_sets.items():
        assert (count, csum) == ref, f"{form} pair set differs from {other}"
    _mixed_sets[form] = (count, csum)


@pytest.fixture(scope="module")
def large():
    return apss.workload("large")


def test_large_t07_against_oracle(large):
    rp, col, val, d = large
    sample = sample_rows(len(rp) - 1, 40, 777)
    ek, es = oracle_pairs_touching(rp, col, val, d, sample, 0.7)
    count, _, keys, scores, info = stream_touching(rp, col, val, d, sample, 0.7)
    print(f"large t=0.7: {count} pairs", flush=True)
    assert check_touching(keys, scores, ek, es, rel=1e-6) > 0


def test_large_slice_split_4096(large):
    # p = 4096 routes the 4096 most frequent features through the prefix-norm
    # filter and the FP32 forward-list head dot; on 1e5 rows of the large
    # workload the pair set equals the oracle's and the p = 0 run's
    rp, col, val, d = large
    rp, col, val = subset(rp, col, val, 100_000)
    sample = sample_rows(100_000, 40, 4096)
    ek, es = oracle_pairs_touching(rp, col, val, d, sample, 0.5)
    c0, s0, _, _, _ = stream_touching(rp, col, val, d, sample, 0.5, p=0)
    c1, s1, keys, scores, info = stream_touching(rp, col, val, d, sample, 0.5, p=4096)
    print(f"large slice p=4096: {c1} pairs, {info['head_pairs']} head dots", flush=True)
    assert info["head_pairs"] > 0 and info["p_used"] == 4096
    check_touching(keys, scores, ek, es, rel=1e-6)
    assert (c1, s1) == (c0, s0)


def test_batched_workload_autotuned():
    # the 256-instance batched workload in one launch sequence with p = -1
    # (the autotuner chooses the configuration); sampled instances against
    # the oracle, every instance's count against its run alone
    inst = apss.batched_instances()
    ts = [0.8] * len(inst)
    res = apss.run_batched(inst, ts, p=-1)
    assert len(res) == 256 and res[0][1]["ms_tune"] > 0
    rng = np.random.default_rng(256)
    for k in rng.choice(len(inst), 4, replace=False):
        print(f"batched instance {k}: n={len(inst[k][0]) - 1}", flush=True)
        check_pairs(res[k][0], oracle.apss_fp64(*inst[k], ts[k]))
    for k in rng.choice(len(inst), 16, replace=False):
        alone, _ = apss.run(*inst[k], ts[k], 0)
        assert len(alone) == len(res[k][0])


def test_full_small_config_default_tiering():
    rp, col, val, d = apss.workload("small")
    got, info = apss.run(rp, col, val, d, 0.7, 0)
    assert info["mfma_features"] > 0 and info["bitmap_features"] > 0
    check_pairs(got, oracle.apss_fp64(rp, col, val, d, 0.7))


def test_query_result_outgrows_first_buffer():
    # one query row meets every row of a 100,000-row corpus: a query starts
    # with room for 65,536 pairs per row, so the scan runs again with the
    # counted size
    rng = np.random.default_rng(5)
    n, d = 100_000, 5000
    rp = np.arange(0, 2 * n + 1, 2, dtype=np.int64)
    col = np.empty(2 * n, np.int32)
    col[0::2] = 0
    col[1::2] = rng.integers(1, d, n)
    val = rng.uniform(0.1, 1.0, 2 * n).astype(np.float32)
    q = (np.array([0, 1], np.int64), np.array([0], np.int32), np.array([1.0], np.float32))
    ix = apss.Index(rp, col, val, d)
    try:
        got, info = ix.query(*q, 0.05)
    finally:
        ix.close()
    exp = oracle.query_fp64(q, (rp, col, val), d, 0.05)
    assert len(exp) == n
    check_pairs(got, exp)


@pytest.mark.parametrize("name,n,p", [("small", 20_000, 0), ("small", 20_000, 4096), ("large", 60_000, 0),
                                      ("mixed", 30_000, 100)])
def test_transpose_matches_oracle(name, n, p):
    rp, col, val, d = apss.workload(name, n=n)
    got = apss.transpose(rp, col, val, d, p)
    exp = oracle.transpose(rp, col, val, d, p)
    for g, e in zip(got, exp):
        assert np.array_equal(g, e)


def test_transpose_long_lists_and_empty():
    # lists of 30,000 entries (cut into sorted pieces and merged) and an empty matrix
    rng = np.random.default_rng(9)
    n, d = 30_000, 400
    rp = np.arange(0, 4 * n + 1, 4, dtype=np.int64)
    col = np.concatenate([np.sort(np.concatenate([[0, 1], rng.choice(np.arange(2, d), 2, replace=False)]))
                          for _ in range(n)]).astype(np.int32)
    val = rng.uniform(0.1, 1.0, 4 * n).astype(np.float32)
    for g, e in zip(apss.transpose(rp, col, val, d, 0), oracle.transpose(rp, col, val, d, 0)):
        assert np.array_equal(g, e)
    z = apss.transpose(np.zeros(1, np.int64), np.zeros(0, np.int32), np.zeros(0, np.float32), 10, 0)
    assert np.array_equal(z[0], np.zeros(11, np.int64)) and len(z[1]) == 0


def _rows(rp, col, val, a, b):
    return ((rp[a:b + 1] - rp[a]).astype(np.int64), col[rp[a]:rp[b]].copy(), val[rp[a]:rp[b]].copy())


def test_serving_workload_full_size():
    # the serving configuration's 2e6-row corpus as a resident index: query
    # rows from the serving query pool through apss_index_query (t = 0.6 and
    # 0.8) and two requests packed by the server, against the FP64 oracle over
    # the whole corpus
    rp, col, val, d = apss.workload("serving")
    qrp, qcol, qval, _ = apss.workload("serving", n=1 << 14, seed_offset=1000)
    q1, q2 = _rows(qrp, qcol, qval, 100, 164), _rows(qrp, qcol, qval, 5000, 5040)
    corpus = (rp, col, val)
    exp = {t: oracle.query_fp64(q1, corpus, d, t) for t in (0.6, 0.8)}  # ~10 s each
    ix = apss.Index(rp, col, val, d)
    try:
        for t in (0.6, 0.8):
            got, info = ix.query(*q1, t)
            print(f"serving corpus n={len(rp) - 1}: query t={t}: {len(got)} pairs", flush=True)
            check_pairs(got, exp[t])
        srv = apss.Server(ix, slots=2)
        try:
            tk1, tk2 = srv.submit(*q1, 0.6), srv.submit(*q2, 0.8)
            g1, g2 = srv.wait(tk1)[0], srv.wait(tk2)[0]
        finally:
            srv.close()
    finally:
        ix.close()
    check_pairs(g1, exp[0.6])
    e2 = oracle.query_fp64(q2, corpus, d, 0.8)
    check_pairs(g2, e2)
    assert len(g1) > 0 and len(e2) > 0
