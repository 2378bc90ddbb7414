"""Batched runs, the corpus index (queries and inserts), the continuous-batching
server and the reference-signature shim, each against the FP64 oracle."""
import os
import subprocess
import threading

import numpy as np
import pytest

import apss
import oracle
from helpers import check_pairs, subset, tiny_matrix

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def small_rows(n, seed_offset=0, name="small"):
    rp, col, val, d = apss.workload(name, n=n, seed_offset=seed_offset)
    return rp, col, val, d


def test_batched_each_instance_matches_alone():
    sizes = [700, 0, 1, 2300, 64, 1500]
    ts = [0.5, 0.3, 0.9, 0.45, 0.2, 0.6]
    inst = [small_rows(n, seed_offset=10 + k) for k, n in enumerate(sizes)]
    res = apss.run_batched(inst, ts, p=0)
    for (rp, col, val, d), t, (pairs, info) in zip(inst, ts, res):
        exp = oracle.apss_fp64(rp, col, val, d, t)
        check_pairs(pairs, exp)


def test_batched_with_split_and_small_bitmap_tier():
    inst = [small_rows(n, seed_offset=30 + k, name="large") for k, n in enumerate([900, 1300, 400])]
    ts = [0.4, 0.5, 0.35]
    res = apss.run_batched(inst, ts, p=50, dense_features=64)
    for (rp, col, val, d), t, (pairs, _) in zip(inst, ts, res):
        check_pairs(pairs, oracle.apss_fp64(rp, col, val, d, t))


def test_batched_thresholds_are_per_instance():
    # instance A holds a pair whose exact FP64 score s sits on A's threshold;
    # instance B, in the same launch, has a much lower threshold.  Pairs near
    # a threshold are settled by the FP64 rescore, which must use A's own.
    rpA, colA, valA, dA = tiny_matrix([{2: 0.6, 5: 0.8}, {2: 0.8, 5: 0.6}, {5: 1.0}], 8)
    s = oracle.pair_fp64(rpA, colA, valA, 0, 1)
    B = small_rows(300, seed_offset=40)
    A = (rpA, colA, valA, B[3])
    for tA, expect in ((float(np.nextafter(s, 2.0)), False), (s, True)):
        res = apss.run_batched([A, B], [tA, 0.2])
        got = set(zip(res[0][0]["i"].tolist(), res[0][0]["j"].tolist()))
        assert ((0, 1) in got) == expect
        check_pairs(res[1][0], oracle.apss_fp64(*B, 0.2))


def test_batched_graph_repeated_with_eager_work_between():
    # every batched call runs its device job as one graph launch; the graph is
    # built once and later calls only update its arguments.  Repeat it with
    # eager runs and synchronisations in between, and with other batches whose
    # buffers differ, checking every result
    inst = [small_rows(n, seed_offset=50 + k) for k, n in enumerate([900, 400, 1300])]
    ts = [0.5, 0.6, 0.45]
    exp = [oracle.apss_fp64(*x, t) for x, t in zip(inst, ts)]
    other = [small_rows(n, seed_offset=90 + k) for k, n in enumerate([2000, 700])]
    exp_other = [oracle.apss_fp64(*x, 0.5) for x in other]
    reused = []
    for rep in range(4):
        res = apss.run_batched(inst, ts, p=0)
        for (pairs, _), e in zip(res, exp):
            check_pairs(pairs, e)
        reused.append(res[0][1]["graph_reused"])
        got, _ = apss.run(*inst[rep % 3], ts[rep % 3], 0)  # eager work between the graph launches
        check_pairs(got, exp[rep % 3])
        res = apss.run_batched(other, [0.5, 0.5], p=0)  # same graph, other arguments
        for (pairs, _), e in zip(res, exp_other):
            check_pairs(pairs, e)
        reused.append(res[0][1]["graph_reused"])
    assert all(reused[1:]), reused


def test_autotuned_runs_match_oracle():
    rp, col, val, d = small_rows(3000)
    got, info = apss.run(rp, col, val, d, 0.5, -1)
    check_pairs(got, oracle.apss_fp64(rp, col, val, d, 0.5))
    assert info["p_used"] in (0, 1024, 4096) and info["ms_tune"] > 0
    inst = [small_rows(n, seed_offset=60 + k, name="large") for k, n in enumerate([1200, 700])]
    res = apss.run_batched(inst, [0.4, 0.5], p=-1)
    for x, t, (pairs, info) in zip(inst, [0.4, 0.5], res):
        check_pairs(pairs, oracle.apss_fp64(*x, t))
        assert info["ms_tune"] > 0
        assert info["p_used"] in (0, 1024, 4096) and info["items_per_cu"] in (4, 8, 16)
        assert (info["band_rows"], info["tile_rows"]) in ((128, 64), (64, 32))


@pytest.mark.parametrize("p,br", [(1024, 64), (4096, 128), (0, 64), (0, 16)])
def test_tuned_shapes_match_oracle(p, br):
    # the configurations the autotuner can choose, set by hand: the split
    # p > 0 and lower band / tile height caps
    rp, col, val, d = small_rows(2500, seed_offset=7, name="large")
    got, info = apss.run(rp, col, val, d, 0.5, p, band_rows=br)
    check_pairs(got, oracle.apss_fp64(rp, col, val, d, 0.5))
    assert (info["band_rows"], info["tile_rows"]) == (br, br // 2)


def test_autotuner_choice_kept_per_job_shape():
    # a second tuned run of the same job shape takes the kept choice without
    # timing candidates: same configuration, same pairs; another threshold is
    # another shape and is tuned again
    rp, col, val, d = small_rows(40_001, seed_offset=131, name="large")
    a, ia = apss.run(rp, col, val, d, 0.6125, -1)
    b, ib = apss.run(rp, col, val, d, 0.6125, -1)
    print(f"tune {ia['ms_tune']:.2f} ms, then {ib['ms_tune']:.3f} ms", flush=True)
    assert ia["ms_tune"] > 2.0 and ib["ms_tune"] < 1.0
    for f in ("p_used", "band_rows", "tile_rows", "items_per_cu", "mfma_features", "bitmap_features"):
        assert ia[f] == ib[f], f
    ga, gb = apss.sort_pairs(a), apss.sort_pairs(b)
    assert np.array_equal(ga["i"], gb["i"]) and np.array_equal(ga["j"], gb["j"])
    check_pairs(b, oracle.apss_fp64(rp, col, val, d, 0.6125))
    _, ic = apss.run(rp, col, val, d, 0.6375, -1)
    assert ic["ms_tune"] > 2.0


def test_autotuned_batch_prefix_inside_an_instance():
    # the autotuner times candidates on a row prefix; with two 40,000-row
    # instances the prefix (16,384 rows) ends inside the first instance, whose
    # bands must then stop at the prefix's last tile
    inst = [small_rows(40_000, seed_offset=120 + k) for k in range(2)]
    res = apss.run_batched(inst, [0.7, 0.7], p=-1)
    for x, (pairs, info) in zip(inst, res):
        check_pairs(pairs, oracle.apss_fp64(*x, 0.7))
        assert info["ms_tune"] > 0


def test_batched_long_lists_across_instances():
    # instances of the batched workload (scaled down): concatenated, nine of
    # their common features form lists longer than one sort piece; a
    # one-feature bitmap tier leaves eight of those to the cursor tier
    inst = apss.batched_instances(count=6, n_lo=3000, n_hi=9000)
    ts = [0.8] * len(inst)
    res = apss.run_batched(inst, ts, p=0, dense_features=1, mfma_head=-1)
    for (rp, col, val, d), t, (pairs, _) in zip(inst, ts, res):
        check_pairs(pairs, oracle.apss_fp64(rp, col, val, d, t))


@pytest.mark.parametrize("p", [0, 64])
def test_index_query_matches_oracle(p):
    crp, ccol, cval, d = small_rows(3000, name="serving")
    qrp, qcol, qval, _ = small_rows(400, seed_offset=100, name="serving")
    ix = apss.Index(crp, ccol, cval, d, p=p)
    try:
        pairs, info = ix.query(qrp, qcol, qval, 0.5)
    finally:
        ix.close()
    exp = oracle.query_fp64((qrp, qcol, qval), (crp, ccol, cval), d, 0.5)
    check_pairs(pairs, exp)


def test_index_inserts_and_merge():
    crp, ccol, cval, d = small_rows(2000, name="serving")
    ix = apss.Index(crp, ccol, cval, d)
    allr = [(crp, ccol, cval)]
    try:
        for k in range(6):  # more inserts than segments kept: forces a merge
            rp, col, val, _ = small_rows(150 + 37 * k, seed_offset=200 + k, name="serving")
            ix.insert(rp, col, val)
            allr.append((rp, col, val))
        n_all = sum(len(r[0]) - 1 for r in allr)
        assert ix.rows() == n_all
        qrp, qcol, qval, _ = small_rows(300, seed_offset=300, name="serving")
        pairs, _ = ix.query(qrp, qcol, qval, 0.45)
    finally:
        ix.close()
    lens = np.concatenate([np.diff(r[0]) for r in allr])
    crp2 = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    ccol2 = np.concatenate([r[1] for r in allr])
    cval2 = np.concatenate([r[2] for r in allr])
    exp = oracle.query_fp64((qrp, qcol, qval), (crp2, ccol2, cval2), d, 0.45)
    check_pairs(pairs, exp)


def test_server_result_same_alone_and_in_the_mix():
    crp, ccol, cval, d = small_rows(2500, name="serving")
    ix = apss.Index(crp, ccol, cval, d)
    sizes = [1, 37, 600, 0, 5, 250, 1200, 64, 3]  # one request without rows
    ts = [0.6, 0.8, 0.6, 0.7, 0.5, 0.8, 0.6, 0.7, 0.9]
    qs = [small_rows(n, seed_offset=400 + k, name="serving")[:3] for k, n in enumerate(sizes)]
    try:
        alone = [ix.query(*q, t)[0] for q, t in zip(qs, ts)]
        srv = apss.Server(ix, slots=2)
        try:
            tickets = [None] * len(qs)

            def submit(k):
                tickets[k] = srv.submit(*qs[k], ts[k])

            th = [threading.Thread(target=submit, args=(k,)) for k in range(len(qs))]
            for x in th:
                x.start()
            for x in th:
                x.join()
            mixed = [srv.wait(tk) for tk in tickets]
        finally:
            srv.close()
    finally:
        ix.close()
    for q, t, a, (m, lat) in zip(qs, ts, alone, mixed):
        exp = oracle.query_fp64(q, (crp, ccol, cval), d, t)
        check_pairs(a, exp)
        check_pairs(m, exp)
        assert lat >= 0


def test_server_segments_under_request_budget():
    # A corpus of a base segment and four delta segments, served with a device
    # budget that fits about one 4096-row request: each launch holds one
    # segment's query structures and scan scratch at a time, so every request
    # completes, and each result equals the oracle's over the whole corpus.
    crp, ccol, cval, d = small_rows(40_000, name="serving")
    ix = apss.Index(crp, ccol, cval, d)
    allr = [(crp, ccol, cval)]
    try:
        for k in range(4):
            rows = small_rows(1500, seed_offset=700 + k, name="serving")[:3]
            ix.insert(*rows)
            allr.append(rows)
        qs = [small_rows(4096, seed_offset=710 + k, name="serving")[:3] for k in range(3)]
        ts = [0.6, 0.8, 0.6]
        srv = apss.Server(ix, slots=2, mem_budget=5 << 29)  # 2.5 GiB
        try:
            tks = [srv.submit(*q, t) for q, t in zip(qs, ts)]
            print(f"submitted {len(tks)} requests of 4096 rows", flush=True)
            got = [srv.wait(tk)[0] for tk in tks]
        finally:
            srv.close()
    finally:
        ix.close()
    lens = np.concatenate([np.diff(r[0]) for r in allr])
    crp2 = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    ccol2 = np.concatenate([r[1] for r in allr])
    cval2 = np.concatenate([r[2] for r in allr])
    for q, t, g in zip(qs, ts, got):
        exp = oracle.query_fp64(q, (crp2, ccol2, cval2), d, t)
        assert len(exp) > 0
        check_pairs(g, exp)


def test_server_cancel_and_budget():
    crp, ccol, cval, d = small_rows(800, name="serving")
    ix = apss.Index(crp, ccol, cval, d)
    try:
        srv = apss.Server(ix, slots=1, mem_budget=1 << 20)  # 1 MiB: every request is too large
        try:
            q = small_rows(10, seed_offset=500, name="serving")[:3]
            with pytest.raises(apss.ApssError) as e:
                srv.submit(*q, 0.5)
            assert e.value.code == apss.APSS_E_NOMEM
        finally:
            srv.close()
        srv = apss.Server(ix, slots=1, max_queue=64)
        try:
            tks = [srv.submit(*small_rows(300, seed_offset=600 + k, name="serving")[:3], 0.5) for k in range(6)]
            cancelled = [tk for tk in tks[3:] if srv.cancel(tk) == apss.APSS_OK]
            for tk in tks:
                if tk in cancelled:
                    with pytest.raises(apss.ApssError) as e:
                        srv.wait(tk)
                    assert e.value.code == apss.APSS_E_CANCELLED
                else:
                    srv.wait(tk)
        finally:
            srv.close()
    finally:
        ix.close()


def test_reference_signature_shim_counts_pairs():
    rp, col, val, d = small_rows(2400)
    # eight partitions, the last one smaller (as in main.cu:180-187)
    cuts = [0, 300, 600, 900, 1200, 1500, 1800, 2100, 2400]
    parts = []
    for a, b in zip(cuts[:-1], cuts[1:]):
        prp = (rp[a:b + 1] - rp[a]).astype(np.int64)
        parts.append((prp, col[rp[a]:rp[b]].copy(), val[rp[a]:rp[b]].copy()))
    count = apss.similarity_join_device(parts)
    exp = oracle.apss_fp64(rp, col, val, max(d, 198456), float(np.float32(0.3)))
    assert count == len(exp)


def test_text_file_through_reference_shim(tmp_path):
    # A main.cu-style caller: rows read from the reference's text format, cut
    # into eight partitions (the last one smaller), handed to the 70-argument
    # shim as SELL-32 slices; the count equals the oracle's at t = 0.3f.
    rp, col, val, d = small_rows(2000)
    path = str(tmp_path / "rows.txt")
    apss.save_text(path, rp, col, val, d)
    rp, col, val, d = apss.load_text(path)
    cuts = [0, 256, 512, 768, 1024, 1280, 1536, 1792, 2000]
    parts = [((rp[a:b + 1] - rp[a]).astype(np.int64), col[rp[a]:rp[b]].copy(), val[rp[a]:rp[b]].copy())
             for a, b in zip(cuts[:-1], cuts[1:])]
    count = apss.similarity_join_device(parts)
    exp = oracle.apss_fp64(rp, col, val, max(d, 198456), float(np.float32(0.3)))
    assert count == len(exp) > 0


def test_cxx_caller_binary_counts_pairs(tmp_path):
    # The reference-style C++ caller (tests/caller/ref_caller.cpp, built by the
    # package Makefile and linked through the C++-linkage symbol) reads the
    # reference text format, builds eight SELL-32 partitions and calls
    # similarity_join_device; its count equals the oracle's at t = 0.3f.
    exe = os.path.join(ROOT, "apss-cuda_amd", "_build", "ref_caller")
    assert os.path.exists(exe), "ref_caller not built (run the package Makefile)"
    rp, col, val, d = small_rows(1800)
    path = str(tmp_path / "rows.txt")
    apss.save_text(path, rp, col, val, d)
    print(f"ref_caller on {path}", flush=True)
    out = subprocess.run([exe, path], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    count = int(out.stdout.strip().splitlines()[-1].split()[1])
    exp = oracle.apss_fp64(rp, col, val, max(d, 198456), float(np.float32(0.3)))
    assert count == len(exp) > 0


def test_concurrent_runs_share_the_pools():
    # apss_run from several host threads at once: each call has its own
    # stream, and the device-block and pinned-host pools are shared.  Every
    # result must equal the single-threaded one, over repeated calls (blocks
    # freed by one call are taken by the others).
    mats = [apss.workload("small", n=n, seed_offset=k) for k, n in enumerate((3000, 5000, 7000))]
    ref = []
    for rp, col, val, d in mats:
        pairs, _ = apss.run(rp, col, val, d, 0.5, 0, sort_output=1)
        ref.append(pairs)
    errors = []

    def worker(k):
        try:
            rp, col, val, d = mats[k]
            for _ in range(4):
                pairs, _ = apss.run(rp, col, val, d, 0.5, 0, sort_output=1)
                if not np.array_equal(pairs, ref[k]):
                    errors.append(f"matrix {k}: result differs")
        except Exception as e:  # reported below
            errors.append(f"matrix {k}: {e!r}")

    threads = [threading.Thread(target=worker, args=(k % 3,)) for k in range(6)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors
    assert sum(len(r) for r in ref) > 1000
