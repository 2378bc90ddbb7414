"""The reference's two candidate bounds at the granularity of this build's
band x tile pairs, on the benchmark workloads (CPU, no GPU):

  block-max   sum_f max_{i in I} |x_if| * max_{j in J} |y_jf|
              (filter_candidate, similarity_search.cu:33-103, with the
              per-partition column maxima of main.cu:56-72);
  L1 x max    min(max_I |x|_1 * max_J max|y|, max_I max|x| * max_J |y|_1)
              (filter_pair2, similarity_search.cu:533-558, per pair there).

Both bound every score of the two blocks (checked against the exact block
maxima on a few pairs); on the Zipf workloads they stay above 1 for every
sampled pair that shares a feature, so neither would skip a tile pair at any
threshold of the configurations (t <= 0.9).  The band kernel therefore
applies them only to bands of rare features (the sparse form), where they do
skip tiles.  Run with -s to see the quantiles."""
import numpy as np
import pytest

import apss


def blocks(rp, col, val, rows):
    a, b = rows
    c = col[rp[a]:rp[b]]
    v = np.abs(val[rp[a]:rp[b]])
    order = np.argsort(c, kind="stable")
    c, v = c[order], v[order]
    uniq, start = np.unique(c, return_index=True)
    fmax = np.maximum.reduceat(v, start) if len(v) else v
    l1 = np.add.reduceat(np.abs(val[rp[a]:rp[b]]), rp[a:b] - rp[a]) if b > a else np.zeros(0)
    mx = np.maximum.reduceat(np.abs(val[rp[a]:rp[b]]), rp[a:b] - rp[a]) if b > a else np.zeros(0)
    return uniq, fmax, l1.max(), mx.max()


def dense(rp, col, val, d, rows):
    a, b = rows
    m = np.zeros((b - a, d))
    for r in range(a, b):
        m[r - a, col[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
    return m


def bounds(rp, col, val, I, J):
    fi, mi, l1i, mxi = blocks(rp, col, val, I)
    fj, mj, l1j, mxj = blocks(rp, col, val, J)
    common, ii, jj = np.intersect1d(fi, fj, assume_unique=True, return_indices=True)
    bm = float(np.dot(mi[ii], mj[jj]))
    lm = float(min(l1i * mxj, mxi * l1j))
    return len(common), bm, lm


@pytest.mark.parametrize("name,n", [("small", 20_000), ("large", 200_000), ("mixed", 100_000)])
def test_block_bounds_never_prune_zipf_workloads(name, n):
    rp, col, val, d = apss.workload(name, n=n)
    rng = np.random.default_rng(8)
    nb, nt = n // 128, n // 64
    bm_all, lm_all = [], []
    for _ in range(400):
        b = int(rng.integers(0, nb))
        t = int(rng.integers(2 * b, nt))  # tiles at or after the band's rows (self-join)
        I, J = (128 * b, min(n, 128 * b + 128)), (64 * t, min(n, 64 * t + 64))
        k, bm, lm = bounds(rp, col, val, I, J)
        if k:
            bm_all.append(bm)
            lm_all.append(lm)
    bm_all, lm_all = np.array(bm_all), np.array(lm_all)
    print(f"\n{name}: {len(bm_all)} sampled tile pairs sharing a feature; block-max q01/q50 "
          f"{np.quantile(bm_all, 0.01):.2f}/{np.quantile(bm_all, 0.5):.2f} (min {bm_all.min():.2f}); "
          f"L1 x max min {lm_all.min():.2f}")
    assert len(bm_all) > 300
    assert bm_all.min() > 1.0 and lm_all.min() > 1.0


def test_block_bounds_hold():
    # the bounds are bounds: on a few pairs of blocks they are at least the
    # largest exact score of the block pair
    rp, col, val, d = apss.workload("small", n=4000)
    for b, t in ((0, 5), (3, 20), (7, 30), (10, 31)):
        I, J = (128 * b, 128 * b + 128), (64 * t, 64 * t + 64)
        _, bm, lm = bounds(rp, col, val, I, J)
        smax = np.abs(dense(rp, col, val, d, I) @ dense(rp, col, val, d, J).T).max()
        assert 0 < smax <= bm * (1 + 1e-6) and smax <= lm * (1 + 1e-6)
