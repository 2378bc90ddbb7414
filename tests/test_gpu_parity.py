"""Parity of the HIP path (through the C-ABI) with the FP64 oracle.

Exact pair-set equality and scores within 1e-6 relative (the library's score
tolerance), on seeded synthetic matrices at sizes the oracle finishes in
seconds, plus the edge cases of the method.
"""
import numpy as np
import pytest

import apss
import oracle
from helpers import BF16_ATOL, check_pairs, check_pairs_abs, subset, tiny_matrix

pytestmark = pytest.mark.gpu


def run_and_check(rp, col, val, d, t, p=0, **opts):
    got, info = apss.run(rp, col, val, d, t, p, **opts)
    exp = oracle.apss_fp64(rp, col, val, d, t)
    err = check_pairs(got, exp)
    return got, exp, info, err


def test_tiny_closed_forms():
    # r0 == r1 (s = 1), r2 disjoint, r3 shares one feature with r0
    x = 0.6, 0.8
    rows = [{1: x[0], 7: x[1]}, {1: x[0], 7: x[1]}, {3: 1.0}, {7: 1.0}, {}]
    rp, col, val, d = tiny_matrix(rows, 10)
    got, info = apss.run(rp, col, val, d, 0.5, 0, sort_output=1)
    pairs = [(int(a), int(b)) for a, b in zip(got["i"], got["j"])]
    assert pairs == [(0, 1), (0, 3), (1, 3)]
    s01 = float(np.float32(0.6) * np.float32(0.6)) + float(np.float32(0.8) * np.float32(0.8))
    assert abs(got["score"][0] - s01) <= 1e-6 * s01
    assert got["score"][1] == np.float32(0.8) and got["score"][2] == np.float32(0.8)


def test_threshold_equal_to_score_is_included():
    rp, col, val, d = tiny_matrix([{2: 0.6, 5: 0.8}, {2: 0.8, 5: 0.6}, {5: 1.0}], 8)
    s = oracle.pair_fp64(rp, col, val, 0, 1)   # exact FP64 score of (0,1)
    got, _ = apss.run(rp, col, val, d, s, 0)
    assert (0, 1) in set(zip(got["i"].tolist(), got["j"].tolist()))
    nxt = np.nextafter(s, 2.0)
    got2, _ = apss.run(rp, col, val, d, nxt, 0)
    assert (0, 1) not in set(zip(got2["i"].tolist(), got2["j"].tolist()))


def test_empty_and_degenerate():
    rp, col, val, d = tiny_matrix([], 5)
    got, info = apss.run(rp, col, val, d, 0.1, 0)
    assert len(got) == 0
    rp, col, val, d = tiny_matrix([{}, {}, {1: 1.0}], 5)
    got, info = apss.run(rp, col, val, d, 0.1, 0)
    assert len(got) == 0
    rp, col, val, d = tiny_matrix([{1: 1.0}, {1: 1.0}], 5)
    got, info = apss.run(rp, col, val, d, 0.1, 5)  # all-head
    assert len(got) == 1


@pytest.mark.parametrize("t", [0.3, 0.5, 0.7])
@pytest.mark.parametrize("p", [0, 64, 4096])
def test_small_config_subset(t, p):
    rp, col, val, d = apss.workload("small")
    rp, col, val = subset(rp, col, val, 3000)
    got, exp, info, err = run_and_check(rp, col, val, d, t, p)
    assert info["p_used"] == p
    assert len(exp) > 0


@pytest.mark.parametrize("cap", [1, 64, 512])
def test_cursor_tier_forced(cap):
    # a small bitmap tier pushes most tail features into the cursor tier
    rp, col, val, d = apss.workload("small")
    rp, col, val = subset(rp, col, val, 4000)
    run_and_check(rp, col, val, d, 0.4, 0, dense_features=cap)


def test_long_cursor_lists_merge_sorted():
    # three features in every one of 30000 rows: with a one-feature bitmap tier
    # and no dense head their lists (30000 entries) go to the cursor tier, are
    # sorted in 8192-entry pieces and merged in two passes
    rng = np.random.default_rng(11)
    n, d = 30000, 2000
    rows = []
    for _ in range(n):
        extra = rng.choice(np.arange(3, d), size=3, replace=False)
        cols = np.concatenate([[0, 1, 2], np.sort(extra)])
        v = rng.uniform(0.05, 1.0, size=len(cols))
        rows.append(dict(zip(cols.tolist(), (v / np.linalg.norm(v)).astype(np.float32).tolist())))
    rp, col, val, d = tiny_matrix(rows, d)
    got, exp, info, err = run_and_check(rp, col, val, d, 0.9, 0, dense_features=1, mfma_head=-1)
    assert info["cursor_products"] > 0 and len(exp) > 0


def test_full_small_config_long_lists():
    # the full small configuration: the most frequent lists exceed 8192 entries
    rp, col, val, d = apss.workload("small")
    run_and_check(rp, col, val, d, 0.7, 0, dense_features=1, mfma_head=-1)


def test_multi_tile_lognormal():
    rp, col, val, d = apss.workload("large", n=6000)
    got, exp, info, err = run_and_check(rp, col, val, d, 0.5, 0)
    assert info["cells"] == 6000 * 5999 // 2


def test_split_invariance():
    rp, col, val, d = apss.workload("small")
    rp, col, val = subset(rp, col, val, 2500)
    sets = []
    for p in (0, 1, 300, 4096, d):
        got, _ = apss.run(rp, col, val, d, 0.45, p)
        g = apss.sort_pairs(got)
        sets.append((g["i"].tobytes(), g["j"].tobytes()))
    assert all(s == sets[0] for s in sets)


def test_rescore_all_scores_exact():
    rp, col, val, d = apss.workload("small")
    rp, col, val = subset(rp, col, val, 2000)
    got, info = apss.run(rp, col, val, d, 0.5, 0, rescore_all=1)
    exp = oracle.apss_fp64(rp, col, val, d, 0.5)
    g = apss.sort_pairs(got)
    assert np.array_equal(g["i"], exp["i"]) and np.array_equal(g["j"], exp["j"])
    assert np.array_equal(g["score"], exp["score"].astype(np.float32))


def test_row_permutation():
    rp, col, val, d = apss.workload("small")
    rp, col, val = subset(rp, col, val, 2000)
    n = len(rp) - 1
    perm = np.random.default_rng(7).permutation(n)
    lens = np.diff(rp)[perm]
    rp2 = np.zeros(n + 1, np.int64)
    rp2[1:] = np.cumsum(lens)
    col2 = np.concatenate([col[rp[i]:rp[i + 1]] for i in perm])
    val2 = np.concatenate([val[rp[i]:rp[i + 1]] for i in perm])
    a, _ = apss.run(rp, col, val, d, 0.5, 0)
    b, _ = apss.run(rp2, col2, val2, d, 0.5, 0)
    inv = np.argsort(perm)  # new id -> old id is perm; old id of new row k is perm[k]
    bi, bj = perm[b["i"]], perm[b["j"]]
    lo, hi = np.minimum(bi, bj), np.maximum(bi, bj)
    sb = set(zip(lo.tolist(), hi.tolist()))
    sa = set(zip(a["i"].tolist(), a["j"].tolist()))
    assert sa == sb and inv is not None


def test_invalid_inputs_rejected():
    rp, col, val, d = tiny_matrix([{1: 1.0, 3: 0.5}, {2: 1.0}], 5)
    bad = col.copy()
    bad[0], bad[1] = 3, 1  # descending inside a row
    with pytest.raises(apss.ApssError) as e:
        apss.run(rp, bad, val, d, 0.5, 0)
    assert e.value.code == apss.APSS_E_INVALID
    badv = val.copy()
    badv[0] = np.nan
    with pytest.raises(apss.ApssError) as e:
        apss.run(rp, col, badv, d, 0.5, 0)
    assert e.value.code == apss.APSS_E_INVALID
    badc = col.copy()
    badc[2] = 9
    with pytest.raises(apss.ApssError) as e:
        apss.run(rp, badc, val, d, 0.5, 0)
    assert e.value.code == apss.APSS_E_INVALID


def test_dense_output_overflows_buffers():
    # a very low threshold makes most cells hits: the LDS staging buffers
    # overflow into HBM directly (the HBM regrowth rerun is forced in
    # test_gpu_fullsize.py::test_query_result_outgrows_first_buffer)
    rp, col, val, d = apss.workload("small")
    rp, col, val = subset(rp, col, val, 1200)
    got, exp, info, err = run_and_check(rp, col, val, d, 0.01, 0)
    assert len(exp) > 100_000


# ---- dense MFMA head ---------------------------------------------------------

@pytest.mark.parametrize("mfma_head", [0, -1])
@pytest.mark.parametrize("p", [0, 64])
def test_dense_head_on_and_off(mfma_head, p):
    rp, col, val, d = apss.workload("large", n=5000)
    got, exp, info, err = run_and_check(rp, col, val, d, 0.4, p, mfma_head=mfma_head)
    assert info["mfma_features"] == (64 if mfma_head == 0 else 0)


def test_dense_head_signed_and_unnormalised_values():
    # random signs exercise the balanced limb digits; a row scale of up to 3.7
    # moves the fixed-point exponents away from the unit-norm case
    rp, col, val, d = apss.workload("mixed", n=4000)
    rng = np.random.default_rng(7)
    sign = np.where(rng.random(len(val)) < 0.3, -1.0, 1.0).astype(np.float32)
    scale = np.repeat(rng.uniform(0.2, 3.7, len(rp) - 1), np.diff(rp)).astype(np.float32)
    v2 = (val * sign * scale).astype(np.float32)
    for t in (0.5, 2.0):
        got, info = apss.run(rp, col, v2, d, t, 0)
        check_pairs(got, oracle.apss_fp64(rp, col, v2, d, t))
        assert info["mfma_features"] == 64


@pytest.mark.parametrize("t", [0.5, 0.8])
def test_bf16_head_pair_set_exact(t):
    # the BF16 head keeps the pair set exact (band rescored in FP64); scores
    # of unit rows carry at most (2^-7 + 2^-16 + 2 gamma_128) + FP32 tail error
    rp, col, val, d = apss.workload("mixed", n=6000)
    got, info = apss.run(rp, col, val, d, t, 0, precision=apss.APSS_PREC_BF16_HEAD)
    exp = oracle.apss_fp64(rp, col, val, d, t)
    err = check_pairs_abs(got, exp, BF16_ATOL)
    assert info["mfma_features"] == 128
    assert info["band"] > 0  # the BF16 bound sends near-threshold cells to the FP64 rescore
    assert err > 0.0


def test_bf16_head_with_split_and_rescore_all():
    rp, col, val, d = apss.workload("large", n=3000)
    exp = oracle.apss_fp64(rp, col, val, d, 0.45)
    got, _ = apss.run(rp, col, val, d, 0.45, 32, precision=apss.APSS_PREC_BF16_HEAD)
    check_pairs_abs(got, exp, BF16_ATOL)
    got, _ = apss.run(rp, col, val, d, 0.45, 0, precision=apss.APSS_PREC_BF16_HEAD, rescore_all=1)
    check_pairs(got, exp)  # every score from the FP64 rescore


def test_bf16_head_signed_unit_rows():
    # negative values: the BF16 head's error is bounded by the head norms (the
    # bound from the cell's own score needs non-negative products)
    rp, col, val, d = apss.workload("mixed", n=5000)
    rng = np.random.default_rng(11)
    v2 = np.where(rng.random(len(val)) < 0.3, -val, val).astype(np.float32)
    for t in (0.3, 0.6):
        got, info = apss.run(rp, col, v2, d, t, 0, precision=apss.APSS_PREC_BF16_HEAD)
        check_pairs_abs(got, oracle.apss_fp64(rp, col, v2, d, t), BF16_ATOL)
        assert info["mfma_features"] == 128


def test_bf16_head_bound_switches_on_a_negative_value():
    # the same matrix with and without one more row holding a single negative
    # value (it pairs with nothing): both pair sets are exact, and without a
    # negative value the bound from the cell's own score sends fewer cells to
    # the FP64 rescore than the head-norm bound
    rp, col, val, d = apss.workload("mixed", n=6000)
    exp = oracle.apss_fp64(rp, col, val, d, 0.5)
    got, info_pos = apss.run(rp, col, val, d, 0.5, 0, precision=apss.APSS_PREC_BF16_HEAD)
    check_pairs_abs(got, exp, BF16_ATOL)
    rp2 = np.append(rp, rp[-1] + 1).astype(np.int64)
    col2 = np.append(col, 0).astype(np.int32)
    val2 = np.append(val, -1.0).astype(np.float32)
    got2, info_neg = apss.run(rp2, col2, val2, d, 0.5, 0, precision=apss.APSS_PREC_BF16_HEAD)
    check_pairs_abs(got2, exp, BF16_ATOL)
    print(f"rescored: {info_pos['band']} without, {info_neg['band']} with a negative value", flush=True)
    assert 0 < info_pos["band"] < info_neg["band"]


def test_text_file_to_pair_file(tmp_path):
    # Rows written in the reference's text format, loaded through
    # apss_load_text, joined on the GPU, written with apss_write_pairs and
    # compared as a set with the oracle's pair file.
    rp, col, val, d = apss.workload("small", n=3000)
    src = str(tmp_path / "rows.txt")
    apss.save_text(src, rp, col, val, d)
    r2, c2, v2, d2 = apss.load_text(src, n_cols=d)
    got, _ = apss.run(r2, c2, v2, d2, 0.5, 0)
    exp = oracle.apss_fp64(rp, col, val, d, 0.5)
    a, b = str(tmp_path / "gpu.txt"), str(tmp_path / "oracle.txt")
    apss.write_pairs(a, got)
    e = np.zeros(len(exp), apss.PAIR_DTYPE)
    e["i"], e["j"], e["score"] = exp["i"], exp["j"], exp["score"]
    apss.write_pairs(b, e)
    assert len(got) > 0
    assert apss.compare_pair_files(a, b) == (0, 0)


def test_cursor_walk_alignment():
    # The cursor walks read list entries as aligned 16-byte words from the
    # multiple of four at or below the cursor and mask the rest.  Lists of
    # lengths around multiples of 4 and 16 start at every offset of the global
    # list array, and with 70000 rows (1094 tiles, more than one 1024-tile
    # bucket chunk) walks also stop inside a step at a chunk end.  Each row has
    # six features out of 4000; the fixed-length lists carry dominant values,
    # so the rows sharing one form the result pairs.
    rng = np.random.default_rng(5)
    n, d = 70000, 4100
    lengths = [1, 2, 3, 4, 5, 7, 8, 9, 15, 16, 17, 31, 32, 33, 47, 48, 49, 63, 64, 65, 255, 256, 257,
               1023, 1024, 1025]
    rows = [dict() for _ in range(n)]
    for r in range(n):
        for f in rng.choice(np.arange(100, d), size=6, replace=False):
            rows[r][int(f)] = float(rng.uniform(0.05, 0.3))
    for f, L in enumerate(lengths, start=10):
        for r in rng.choice(n, size=L, replace=False):
            rows[r][f] = float(rng.uniform(2.0, 4.0))
    for r in range(n):
        v = np.array(list(rows[r].values()))
        s = float(np.linalg.norm(v))
        rows[r] = {k: float(np.float32(x / s)) for k, x in rows[r].items()}
    rp, col, val, d = tiny_matrix(rows, d)
    got, exp, info, err = run_and_check(rp, col, val, d, 0.35, 0, dense_features=1, mfma_head=-1)
    assert info["cursor_products"] > 0 and len(exp) > 100000


def test_more_than_65535_tiles():
    # A corpus of 4.5e6 rows of one feature each is cut into 70313 tiles of 64
    # rows; a 100-row query meets every corpus row sharing its feature with
    # score 1 (the query is one band, so the scan stays small)
    rng = np.random.default_rng(21)
    n, d = 4_500_000, 2_000_000
    rp = np.arange(n + 1, dtype=np.int64)
    col = rng.integers(0, d, size=n).astype(np.int32)
    val = np.ones(n, np.float32)
    ix = apss.Index(rp, col, val, d, p=0)
    try:
        qcol = col[rng.choice(n, 100, replace=False)]
        qrp = np.arange(101, dtype=np.int64)
        got, info = ix.query(qrp, qcol, np.ones(100, np.float32), 0.99)
    finally:
        ix.close()
    assert info["tile_pairs"] >= 70313
    expected = {(q, int(r)) for q in range(100) for r in np.flatnonzero(col == qcol[q])}
    assert {(int(a), int(b)) for a, b in zip(got["i"], got["j"])} == expected
    assert np.all(got["score"] == np.float32(1.0))


@pytest.mark.parametrize("sort_output", [0, 1])
def test_device_inputs_and_device_output(sort_output):
    # The mode the benchmark runs in: the CSR already in device memory and the
    # pairs left there (copied back here only to check them); with
    # sort_output the device result is sorted by (i, j).
    rp, col, val, d = apss.workload("small", n=4000)
    hip = apss.Hip()
    X = apss.DeviceCSR(hip, rp, col, val, d)
    try:
        r = apss.run_csr(X.csr(), 0.5, 0, output_on_device=1, sort_output=sort_output)
        try:
            assert r.on_device == 1
            got = np.zeros(r.count, apss.PAIR_DTYPE)
            if r.count:
                hip.d2h(got, r.pairs)
        finally:
            apss.lib().apss_result_free(apss.ctypes.byref(r))
    finally:
        X.release()
    exp = oracle.apss_fp64(rp, col, val, d, 0.5)
    check_pairs(got, exp)
    if sort_output:
        key = got["i"].astype(np.int64) * (1 << 32) + got["j"]
        assert np.all(np.diff(key) > 0)


@pytest.mark.parametrize("mfma_head", [0, -1])
def test_mixed_workload_subset_exact(mfma_head):
    # the heavy Zipf(1.2) head of the mixed workload through the exact int8
    # MFMA head and through the plain sparse path: exact pair set, scores
    # within 1e-6 relative
    rp, col, val, d = apss.workload("mixed", n=8000)
    got, exp, info, err = run_and_check(rp, col, val, d, 0.5, 0, mfma_head=mfma_head)
    assert info["mfma_features"] == (64 if mfma_head == 0 else 0)
    assert len(exp) > 10000
