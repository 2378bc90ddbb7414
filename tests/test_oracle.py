"""The oracle (oracle/liboracle.so) pinned against independent routines.

No GPU.  The reference ships no tests, golden vectors or fixtures, so the
oracle is pinned by: dense float64 Gram matrices (numpy matmul, committed as
fixtures and recomputed on random inputs), scipy's CSR->CSC, closed forms and
brute force on tiny inputs.  Each check is chosen so that a plausible mistake
in the oracle (a dropped term, a transposed operand, an off-by-one in the
triangular restriction, a wrong comparison at t) fails it.
"""
import glob
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from helpers import tiny_matrix

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def dense_gram(rp, col, val, d):
    A = sp.csr_matrix((np.asarray(val, np.float64), col, rp), shape=(len(rp) - 1, d)).toarray()
    return A @ A.T


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "pairs_*.npz"))))
def test_apss_fp64_matches_golden(path):
    z = np.load(path)
    got = oracle.apss_fp64(z["rp"], z["col"], z["val"], int(z["d"]), float(z["t"]))
    assert np.array_equal(got["i"], z["i"]) and np.array_equal(got["j"], z["j"])
    assert np.allclose(got["score"], z["s"], rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_apss_fp64_matches_dense_random(seed):
    rng = np.random.default_rng(seed)
    n, d = 120, 40
    rows = []
    for _ in range(n):
        k = int(rng.integers(0, 9))
        c = rng.choice(d, size=k, replace=False)
        v = rng.normal(size=k)
        rows.append(dict(zip(c.tolist(), (v / max(np.linalg.norm(v), 1e-12)).tolist())))
    rp, col, val, d = tiny_matrix(rows, d)
    G = dense_gram(rp, col, val, d)
    iu, ju = np.triu_indices(n, 1)
    s = G[iu, ju]
    t = 0.25
    keep = np.abs(s - t) > 1e-9
    exp = set(zip(iu[(s >= t) & keep].tolist(), ju[(s >= t) & keep].tolist()))
    got = oracle.apss_fp64(rp, col, val, d, t)
    gs = set(zip(got["i"].tolist(), got["j"].tolist()))
    near = set(zip(iu[~keep].tolist(), ju[~keep].tolist()))
    assert gs - near == exp
    assert np.all(got["i"] < got["j"])


def test_query_fp64_matches_dense():
    rng = np.random.default_rng(5)
    d = 30

    def rand_rows(n):
        out = []
        for _ in range(n):
            k = int(rng.integers(1, 7))
            c = rng.choice(d, size=k, replace=False)
            v = rng.uniform(0.1, 1.0, size=k)
            out.append(dict(zip(c.tolist(), (v / np.linalg.norm(v)).tolist())))
        return out

    q = tiny_matrix(rand_rows(25), d)
    c = tiny_matrix(rand_rows(70), d)
    Q = sp.csr_matrix((q[2].astype(np.float64), q[1], q[0]), shape=(25, d)).toarray()
    C = sp.csr_matrix((c[2].astype(np.float64), c[1], c[0]), shape=(70, d)).toarray()
    S = Q @ C.T
    t = 0.3
    exp = {(a, b) for a, b in zip(*np.nonzero(S >= t)) if abs(S[a, b] - t) > 1e-9}
    got = oracle.query_fp64(q[:3], c[:3], d, t)
    assert {(int(a), int(b)) for a, b in zip(got["i"], got["j"])} == exp


def test_closed_forms():
    # identical rows -> sum of squares; disjoint -> nothing; one shared feature -> x*y
    rows = [{0: 0.6, 3: 0.8}, {0: 0.6, 3: 0.8}, {5: 1.0}, {3: 0.5, 6: np.sqrt(0.75)}]
    rp, col, val, d = tiny_matrix(rows, 8)
    got = oracle.apss_fp64(rp, col, val, d, 0.1)
    pairs = {(int(a), int(b)): s for a, b, s in zip(got["i"], got["j"], got["score"])}
    x, y = float(np.float32(0.6)), float(np.float32(0.8))
    assert pairs[(0, 1)] == x * x + y * y
    assert pairs[(0, 3)] == y * float(np.float32(0.5))
    assert pairs[(1, 3)] == pairs[(0, 3)]
    assert (0, 2) not in pairs and (1, 2) not in pairs and (2, 3) not in pairs


def test_threshold_is_inclusive():
    rp, col, val, d = tiny_matrix([{1: 0.6, 2: 0.8}, {1: 0.8, 2: 0.6}], 4)
    s = oracle.pair_fp64(rp, col, val, 0, 1)
    assert len(oracle.apss_fp64(rp, col, val, d, s)) == 1
    assert len(oracle.apss_fp64(rp, col, val, d, np.nextafter(s, 2.0))) == 0


def test_pair_fp64_is_ascending_order_sum():
    # three terms whose float64 sum depends on the order: the rescore rule is
    # ascending feature id, starting from 0.0
    a = {0: 1.0, 1: 1.0, 2: 1.0}
    b = {0: np.float32(1e8), 1: np.float32(1.0), 2: np.float32(-1e8)}
    rp, col, val, d = tiny_matrix([a, b], 3)
    exp = 0.0
    for k in range(3):
        exp += float(np.float32(a[k])) * float(np.float32(b[k]))
    assert oracle.pair_fp64(rp, col, val, 0, 1) == exp


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "csc_*.npz"))))
def test_transpose_matches_scipy(path):
    z = np.load(path)
    cptr, crow, cval = oracle.transpose(z["rp"], z["col"], z["val"], int(z["d"]), int(z["p"]))
    assert np.array_equal(cptr, z["cptr"])
    assert np.array_equal(crow, z["crow"])
    assert np.array_equal(cval, z["cval"])


def test_hybrid_matches_exact_up_to_fp32_rounding():
    z = np.load(os.path.join(GOLDEN, "pairs_small400_t0.3.npz"))
    rp, col, val, d = z["rp"], z["col"], z["val"], int(z["d"])
    exact = oracle.apss_fp64(rp, col, val, d, 0.3)
    for p in (0, 64, 4096, d):
        hyb, st = oracle.hybrid_fp32(rp, col, val, d, 0.3, p)
        a = set(zip(exact["i"].tolist(), exact["j"].tolist()))
        b = set(zip(hyb["i"].tolist(), hyb["j"].tolist()))
        for i, j in a ^ b:  # any disagreement must sit within FP32 rounding of t
            assert abs(oracle.pair_fp64(rp, col, val, i, j) - 0.3) < 1e-5
        assert st["survivors"] <= st["candidates"]
    # p = d: no tail, so the filter bound is |x||y| and every pair is a candidate
    _, st = oracle.hybrid_fp32(rp, col, val, d, 0.3, d)
    n = len(rp) - 1
    assert st["candidates"] == n * (n - 1) // 2


def test_hybrid_row_subset_is_a_restriction():
    z = np.load(os.path.join(GOLDEN, "pairs_small400_t0.5.npz"))
    rp, col, val, d = z["rp"], z["col"], z["val"], int(z["d"])
    full, _ = oracle.hybrid_fp32(rp, col, val, d, 0.5, 4096)
    rows = np.array([3, 50, 200], np.int64)
    part, _ = oracle.hybrid_fp32(rp, col, val, d, 0.5, 4096, rows)
    fs = {(int(i), int(j)) for i, j in zip(full["i"], full["j"]) if i in set(rows.tolist())}
    assert {(int(i), int(j)) for i, j in zip(part["i"], part["j"])} == fs


@pytest.mark.parametrize("t", [0.3, 0.5, 0.8])
def test_inverted_search_counts_match_gram(t):
    z = np.load(os.path.join(GOLDEN, "pairs_small400_t0.3.npz"))
    rp, col, val, d = z["rp"], z["col"], z["val"], int(z["d"])
    G = dense_gram(rp, col, val, d)
    iu = np.triu_indices(len(rp) - 1, 1)
    s = G[iu]
    exact = int(np.count_nonzero(s >= t))
    near = int(np.count_nonzero(np.abs(s - t) < 1e-5))
    cnt, sec, done = oracle.inverted_search_fp32(rp, col, val, d, t)
    assert abs(cnt - exact) <= near and sec >= 0.0 and done == len(rp) - 1
    # a row subset counts exactly its rows' pairs (j > i)
    rows = np.array([0, 7, 123, 398], np.int64)
    sub, _, _ = oracle.inverted_search_fp32(rp, col, val, d, t, rows)
    mask = np.isin(iu[0], rows)
    assert abs(sub - int(np.count_nonzero(s[mask] >= t))) <= int(np.count_nonzero(np.abs(s[mask] - t) < 1e-5))


def test_inverted_search_closed_forms():
    # identical rows score 1; disjoint rows 0; one shared feature a*b
    rp, col, val, d = tiny_matrix([{1: 0.6, 3: 0.8}, {1: 0.6, 3: 0.8}, {5: 1.0}, {3: 1.0}], 6)
    assert oracle.inverted_search_fp32(rp, col, val, d, 0.99)[0] == 1
    assert oracle.inverted_search_fp32(rp, col, val, d, 0.8)[0] == 3      # (0,1), (0,3), (1,3)
    assert oracle.inverted_search_fp32(rp, col, val, d, 0.81)[0] == 1
