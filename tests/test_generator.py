"""The synthetic workload generator (libapss_host.so).  No GPU."""
import glob
import os

import numpy as np
import pytest

import apss

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "gen_*_40rows.npz"))))
def test_generator_is_pinned(path):
    name = os.path.basename(path).split("_")[1]
    z = np.load(path)
    rp, col, val, d = apss.workload(name, n=40)
    assert d == int(z["d"])
    assert np.array_equal(rp, z["rp"]) and np.array_equal(col, z["col"]) and np.array_equal(val, z["val"])


def test_rows_independent_of_batch():
    # rows are a function of (seed, global row id) only
    w = apss.WORKLOADS["small"]
    full = apss.generate(300, w["d"], w["mean"], w["kmin"], w["kmax"], w["zipf"], w["values"], seed=7)
    part = apss.generate(100, w["d"], w["mean"], w["kmin"], w["kmax"], w["zipf"], w["values"], seed=7,
                         row_offset=150)
    a0, a1 = full[0][150], full[0][250]
    assert np.array_equal(full[1][a0:a1], part[1])
    assert np.array_equal(full[2][a0:a1], part[2])


@pytest.mark.parametrize("name", ["small", "large", "mixed"])
def test_row_invariants(name):
    rp, col, val, d = apss.workload(name, n=3000)
    w = apss.WORKLOADS[name]
    lens = np.diff(rp)
    assert lens.min() >= w["kmin"] and lens.max() <= w["kmax"]
    assert abs(lens.mean() - w["mean"]) / w["mean"] < 0.2
    assert col.min() >= 0 and col.max() < d
    for i in range(0, 3000, 97):  # ascending, distinct, unit norm
        c = col[rp[i]:rp[i + 1]]
        assert np.all(np.diff(c) > 0)
        nv = np.linalg.norm(val[rp[i]:rp[i + 1]].astype(np.float64))
        assert abs(nv - 1.0) < 1e-6
    if w["values"] == "uniform":
        assert val.min() > 0 and val.max() <= 1.0


def test_zipf_rank_frequency():
    rp, col, val, d = apss.workload("large", n=20000)
    cnt = np.bincount(col, minlength=64)[:64]
    # id 0 the most frequent; counts decay with rank
    assert cnt[0] == cnt.max()
    assert cnt[:8].sum() > cnt[8:16].sum() > cnt[32:40].sum()


def test_alpha_solves_mean():
    H = apss.host_lib()
    for kmin, kmax, mean in [(5, 500, 40), (5, 5000, 100), (5, 5000, 60)]:
        a = H.apss_gen_alpha(kmin, kmax, mean)
        k = np.arange(kmin, kmax + 1, dtype=np.float64)
        w = k ** -a
        assert abs((w * k).sum() / w.sum() - mean) < 1e-6
