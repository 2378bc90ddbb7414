"""Writes the golden fixtures of the test suite (run once; outputs are committed).

The reference ships no tests, fixtures or data (its only dataset file is not in
the repository), so these vectors pin the oracle and the generator from
independent routines:

  * expected pair sets and scores: a dense float64 Gram matrix computed with
    numpy matmul over the densified rows (a textbook routine, summation order
    unrelated to the oracle's ascending-feature loop);
  * CSC layout of the inverted index: scipy.sparse csr_matrix.tocsc();
  * generator output: the first rows of the named workloads, by value.

Each pair fixture uses thresholds at least 1e-9 away from every score, so the
set does not depend on the summation order.

Usage:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "apss-cuda_amd"))

import apss  # noqa: E402


def dense_pairs(rp, col, val, d, t):
    n = len(rp) - 1
    A = sp.csr_matrix((val.astype(np.float64), col, rp), shape=(n, d)).toarray()
    G = A @ A.T
    iu, ju = np.triu_indices(n, 1)
    s = G[iu, ju]
    gap = np.min(np.abs(s - t)) if len(s) else 1.0
    assert gap > 1e-9, f"threshold {t} too close to a score ({gap})"
    m = s >= t
    return iu[m].astype(np.uint32), ju[m].astype(np.uint32), s[m]


def take(rp, col, val, n):
    return rp[: n + 1].copy(), col[: rp[n]].copy(), val[: rp[n]].copy()


def signed_matrix(n, d, seed):
    rng = np.random.default_rng(seed)
    rp = [0]
    cols, vals = [], []
    for _ in range(n):
        k = int(rng.integers(0, 12))
        c = np.sort(rng.choice(d, size=k, replace=False))
        v = rng.uniform(-1, 1, size=k)
        nv = np.linalg.norm(v)
        v = v / nv if nv > 0 else v
        cols.extend(c.tolist())
        vals.extend(v.astype(np.float32).tolist())
        rp.append(len(cols))
    return np.array(rp, np.int64), np.array(cols, np.int32), np.array(vals, np.float32)


def main():
    out = {}
    rp, col, val, d = apss.workload("small")
    r = take(rp, col, val, 400)
    for t in (0.3, 0.5):
        i, j, s = dense_pairs(*r, d, t)
        out[f"small400_t{t}"] = dict(rp=r[0], col=r[1], val=r[2], d=d, t=t, i=i, j=j, s=s)
    rp, col, val, d = apss.workload("large", n=300)
    i, j, s = dense_pairs(rp, col, val, d, 0.4)
    out["large300_t0.4"] = dict(rp=rp, col=col, val=val, d=d, t=0.4, i=i, j=j, s=s)
    r = signed_matrix(150, 60, 11)
    i, j, s = dense_pairs(*r, 60, 0.3)
    out["signed150_t0.3"] = dict(rp=r[0], col=r[1], val=r[2], d=60, t=0.3, i=i, j=j, s=s)
    for name, dct in out.items():
        np.savez_compressed(os.path.join(HERE, f"pairs_{name}.npz"), **dct)
        print(name, "pairs:", len(dct["i"]))

    # inverted index (CSC of ids >= p) of the signed matrix and of small rows
    r = signed_matrix(150, 60, 11)
    rs = take(*apss.workload("small")[:3], 200)
    for name, (m_rp, m_col, m_val, m_d) in {"signed150": (*r, 60), "small200": (*rs, 50_000)}.items():
        for p in (0, 7):
            keep = m_col >= p
            rows = np.repeat(np.arange(len(m_rp) - 1), np.diff(m_rp))
            A = sp.csr_matrix((m_val[keep], (rows[keep], m_col[keep])), shape=(len(m_rp) - 1, m_d))
            C = A.tocsc()
            C.sort_indices()
            np.savez_compressed(os.path.join(HERE, f"csc_{name}_p{p}.npz"), rp=m_rp, col=m_col, val=m_val,
                                d=m_d, p=p, cptr=C.indptr.astype(np.int64), crow=C.indices.astype(np.int32),
                                cval=C.data.astype(np.float32))

    # generator output by value
    for wl in ("small", "large", "mixed", "serving"):
        rp, col, val, d = apss.workload(wl, n=40)
        np.savez_compressed(os.path.join(HERE, f"gen_{wl}_40rows.npz"), rp=rp, col=col, val=val, d=d)


if __name__ == "__main__":
    main()
