"""Tile-pair skipping for bands whose rows hold rare features only (the
sparse form of the band kernel): tiles that share no feature with the band are
skipped, and tiles whose block bound (the L1 x max bound of the two blocks, or
the sum of their products) stays below t are not settled.  Every result is
checked against the FP64 oracle."""
import numpy as np
import pytest

import apss
import oracle
from helpers import check_pairs

pytestmark = pytest.mark.gpu


def rare_rows(n, d, per_row, seed, lo=0.05):
    """Rows of `per_row` features drawn uniformly from d (no frequent
    feature), uniform values, L2-normalised."""
    rng = np.random.default_rng(seed)
    col = np.sort(rng.integers(0, d, size=(n, per_row)), axis=1)
    # distinct ids inside a row
    col[:, 1:] = np.maximum(col[:, 1:], col[:, :-1] + 1)
    col = np.minimum(col, d - 1)
    keep = np.concatenate([np.ones((n, 1), bool), col[:, 1:] > col[:, :-1]], axis=1)
    lens = keep.sum(1)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    col = col[keep].astype(np.int32)
    val = rng.uniform(lo, 1.0, len(col))
    for r in range(n):
        v = val[rp[r]:rp[r + 1]]
        val[rp[r]:rp[r + 1]] = v / np.linalg.norm(v)
    return rp, col, val.astype(np.float32), d


def test_corpus_without_frequent_features_skips_tiles():
    rp, col, val, d = rare_rows(120_000, 400_000, 3, 1)
    got, info = apss.run(rp, col, val, d, 0.5, 0)
    print(f"{info['tile_pairs']} tile pairs, {info['tile_pairs_skipped']} skipped, "
          f"{info['tile_pairs_pruned']} pruned, scan {info['ms_scan']:.2f} ms", flush=True)
    check_pairs(got, oracle.apss_fp64(rp, col, val, d, 0.5))
    assert info["bitmap_features"] == 0 and info["mfma_features"] == 0
    assert info["tile_pairs_skipped"] > 0.5 * info["tile_pairs"]
    assert info["tile_pairs_skipped"] + info["tile_pairs_pruned"] > 0.9 * info["tile_pairs"]


def test_sparse_and_dense_bands_in_one_run():
    # the first rows from the Zipf workload (bands of the full kernel), the
    # rest rare-only (bands of the sparse form); with p = 0 and p = 1000
    zrp, zcol, zval, d = apss.workload("large", n=6000)
    rrp, rcol, rval, _ = rare_rows(20_000, d, 4, 2)
    rp = np.concatenate([zrp, zrp[-1] + rrp[1:]]).astype(np.int64)
    col = np.concatenate([zcol, rcol])
    val = np.concatenate([zval, rval])
    exp = oracle.apss_fp64(rp, col, val, d, 0.4)
    for p in (0, 1000):
        got, info = apss.run(rp, col, val, d, 0.4, p)
        check_pairs(got, exp)
        assert info["tile_pairs_skipped"] > 0


def test_block_bound_prunes_tiles():
    # every row has one of 1000 group features (about 30 rows each, a rare
    # list) with a small value and one feature of its own: tiles share group
    # features, but their products sum far below t, so no tile is settled
    n, d = 30_000, 100_000
    rng = np.random.default_rng(3)
    rp = np.arange(0, 2 * n + 1, 2, dtype=np.int64)
    col = np.empty(2 * n, np.int32)
    col[0::2] = rng.integers(0, 1000, n)
    col[1::2] = 1000 + rng.permutation(d - 1000)[:n]
    val = np.empty(2 * n, np.float32)
    val[0::2] = 0.01
    val[1::2] = np.sqrt(1 - 0.01 ** 2)
    got, info = apss.run(rp, col, val, d, 0.3, 0)
    exp = oracle.apss_fp64(rp, col, val, d, 0.3)
    check_pairs(got, exp)
    assert len(exp) == 0
    assert info["tile_pairs_pruned"] > 0 and info["bitmap_features"] == 0


def test_sparse_batched_and_query():
    inst = [rare_rows(n, 50_000, 3, 10 + k) for k, n in enumerate((9000, 4000, 12000))]
    res = apss.run_batched(inst, [0.5] * 3, p=0)
    for x, (pairs, info) in zip(inst, res):
        check_pairs(pairs, oracle.apss_fp64(*x, 0.5))
    assert res[0][1]["tile_pairs_skipped"] > 0
    crp, ccol, cval, d = inst[2]
    qrp, qcol, qval, _ = inst[1]
    ix = apss.Index(crp, ccol, cval, d)
    try:
        got, info = ix.query(qrp, qcol, qval, 0.5)
    finally:
        ix.close()
    check_pairs(got, oracle.query_fp64((qrp, qcol, qval), (crp, ccol, cval), d, 0.5))
