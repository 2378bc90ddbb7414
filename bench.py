"""APSS benchmark on one MI355X.

Headline workload: the large synthetic configuration -- n = 1,000,000 rows,
d = 1,000,000 features, ~1e8 nonzeros (power-law row lengths, mean 100, in
[5, 5000]), Zipf(1.05) feature ids, lognormal(0,1) values, L2-normalised --
self-joined at t = 0.5 with split p = 0 (every feature on the inverted-list /
dense-head path), FP32 mode (pair set identical to the FP64 result, scores
within 1e-6 relative).

A step is one full APSS job on the device-resident CSR: validation + norms,
feature histogram and tier selection, tiled inverted index and dense head
records, the band-scan kernel over all tile pairs, the FP64 rescore of the
error band; the pair list stays in device memory.  value = n(n-1)/2 / wall
(pair-space throughput, comparable across t); result pairs/s beside it.

Legs printed on the same JSON line (each can be switched off):
  t0.9      one step of the large workload at t = 0.9;
  variants  the large workload at t = 0.5 without the dense MFMA head (the
            plain sparse path) and with the BF16 head (mixed precision);
  mixed     the mixed-precision configuration (n = d = 5e5, mean 80 nnz/row,
            Zipf(1.2), uniform values, t = 0.5): plain path, exact int8 head,
            BF16 head -- the pair sets must be identical;
  split     the split p swept over {0, 1024, 2048, 4096, 8192, 16384} at t = 0.5 and 0.9 on
            the first 50,000 rows of the large workload (p > 0 runs the
            reference's prefix-norm filter + forward-list head dot); the pair
            counts must not depend on p;
  batched   256 independent instances (n log-uniform in [500, 5e4], d = 1e5)
            in one captured band-scan graph with the tiering autotuned (p = -1),
            against the same instances run one by one;
  serving   2e6-row resident corpus (d = 1e6, mean 60 nnz/row, Zipf(1.1),
            lognormal), query batches of 1..4096 rows (log-uniform), t drawn
            from {0.6, 0.8}, Poisson arrivals at fractions of the measured
            capacity, 1% of the submitted rows inserted into the index while
            queries run; p50 / p99 request latency (submit -> result);
  h2h       the headline job from host CSR to host pairs (PCIe included),
            wall time of two calls;
  roofline  band-scan kernel: algorithmic bytes (alg_terms below: tile-side
            list entries per tile pair, band-side entries once per work item,
            cursor walks and products, dense head records, segment-start rows,
            pairs out) over its hipEvent-timed duration on the stream it runs
            on, against 8 TB/s; beside it the rocprofv3 HBM traffic of the same
            launch (profiles/) and its ratio to the model;
  cpu       the reference algorithm restated for the CPU (oracle/), timed on
            the host cores on a stratified sample of query rows, each joined
            with every later row (pair-space throughput).
  cpu_serial  the reference's serial inverted-index search (main.cu:104-141,
            oracle/), one host thread, on a seeded random sample of query rows
            (a second CPU data point).
"""
import argparse
import json
import math
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "apss-cuda_amd"))

import numpy as np  # noqa: E402

import apss  # noqa: E402

HBM_PEAK_GBS = 8000.0
ALL_LEGS = ("t09", "variants", "h2h", "split", "mixed", "batched", "serving", "cpu")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profiles():
    """The committed rocprofv3 PMC summary of the band scan (FETCH_SIZE and
    WRITE_SIZE per dispatch, raw and scaled by a copy kernel of known size), or
    None."""
    path = os.path.join(ROOT, "profiles", "band_scan_large_traffic.json")
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def alg_terms(info):
    """Algorithmic bytes of one band-scan launch, term by term, from its counters.

    Every operand fetched from HBM counts once per logical use; data the kernel
    stages in registers or LDS for a whole work item counts once per item.
      tile_lists    8 B (row id + value) per tile-side list entry: the J tile's
                    segment of every feature the band holds, per tile pair
      band_lists    8 B per band-side bitmap-tier entry, once per work item
      band_records  16 B per band cursor record, once per work item
      walks         4 B per list position read by the counting walks, 8 B
                    (position + value) per entry read by the scatter walks
      cursor        16 B per cursor-tier product (bucket write + read)
      dense         256 B per dense-head record fetched (tile rows per tile
                    pair, band rows once per item)
      seg_rows      one J tile's segment-start row per tile pair
      pairs_out     12 B per pair leaving the tile (output and error band)
    """
    return {
        "tile_lists": 8.0 * info["list_visits"],
        "band_lists": 8.0 * info["band_entries"],
        "band_records": 16.0 * info["band_records"],
        "walks": 4.0 * info["walk_entries"] + 8.0 * info["walk_scatter_entries"],
        "cursor": 16.0 * info["cursor_products"],
        "dense": 256.0 * info["dense_rows"],
        "seg_rows": float(info["seg_row_bytes"]) * info["tile_pairs"],
        "pairs_out": 12.0 * (info["count"] + info["band"]),
    }


def alg_bytes(info):
    """Algorithmic bytes of one band-scan launch (sum of alg_terms)."""
    return sum(alg_terms(info).values())


def cpu_baseline(rp, col, val, d, t, budget_s):
    """Reference hybrid (oracle/, FP32, p = 4096 as in the reference) on a
    stratified row sample; returns the pair-space throughput dict."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    n = len(rp) - 1
    threads = oracle.num_threads()
    # calibrate on a small sample that keeps every thread busy (the restatement
    # hands out query rows in blocks of 16), then size the sample to the budget
    probe = np.linspace(0, n - 2, num=min(n - 1, 32 * threads), dtype=np.int64)
    _, st = oracle.hybrid_fp32(rp, col, val, d, t, 4096, probe)
    wall_per_row = max(st["search_seconds"], 1e-6) / len(probe)
    m = int(min(n - 1, max(16 * threads, budget_s / wall_per_row)))
    rows = np.linspace(0, n - 2, num=m, dtype=np.int64)
    _, st = oracle.hybrid_fp32(rp, col, val, d, t, 4096, rows)
    pairs = float(np.sum((n - 1 - rows).astype(np.float64)))
    return {
        "value": pairs / st["search_seconds"],
        "unit": "pairs/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{m} query rows stratified over the {n} rows, each joined with every later row "
                   f"({pairs:.3e} pair cells, {st['search_seconds']:.1f} s); reference hybrid, p=4096, FP32"),
    }


def cpu_serial(rp, col, val, d, t, budget_s):
    """Serial inverted-index search (oracle/, one thread, FP32; the reference's
    InvertedIndex_search) on a stratified row sample: a second CPU data point."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    n = len(rp) - 1
    # a stratified grid in a fixed random order: any prefix is a sample of the
    # whole range; the search stops after budget_s seconds
    grid = np.linspace(0, n - 2, num=min(n - 1, 65536), dtype=np.int64)
    grid = grid[np.random.default_rng(7).permutation(len(grid))]
    cnt, sec, m = oracle.inverted_search_fp32(rp, col, val, d, t, grid, budget_s)
    rows = grid[:m]
    pairs = float(np.sum((n - 1 - rows).astype(np.float64)))
    return {
        "value": pairs / sec,
        "unit": "pairs/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"{m} query rows (a seeded random prefix of a stratified grid over the {n} rows), each "
                   f"scored against every row through "
                   f"the full inverted index ({pairs:.3e} pair cells of j > i, {cnt} pairs, {sec:.1f} s); "
                   f"serial inverted-index search, FP32"),
    }


def host_to_host(rp, col, val, d, t, p, runs=2, **opts):
    """The headline job from host CSR to host pairs (upload, search, download
    into host memory): wall time of each of `runs` calls through the C-ABI."""
    L = apss.lib()
    X = apss.make_csr(rp, col, val, d)
    o = apss.make_opts(**opts)
    walls, info = [], None
    for _ in range(runs):
        r = apss.Result()
        t0 = time.perf_counter()
        rc = L.apss_run(apss.ctypes.byref(X), float(t), int(p), apss.ctypes.byref(o), apss.ctypes.byref(r))
        walls.append((time.perf_counter() - t0) * 1e3)
        if rc != apss.APSS_OK:
            raise apss.ApssError(rc, L.apss_last_error().decode())
        info = info_of(r)
        L.apss_result_free(apss.ctypes.byref(r))
    n = len(rp) - 1
    return {"wall_ms": [round(w, 1) for w in walls], "pairs": info["count"],
            "pairs_per_s": n * (n - 1) / 2.0 / (min(walls) * 1e-3),
            "phases_ms": {k: round(info[k], 1) for k in ("ms_upload", "ms_index", "ms_scan", "ms_download")}}


def info_of(r):
    return {f: getattr(r, f) for f, _ in apss.Result._fields_ if f not in ("pairs", "internal_")}


def run_device(Xc, t, p, stream, **opts):
    r = apss.run_csr(Xc, t, p, output_on_device=1, stream=stream, **opts)
    info = info_of(r)
    apss.lib().apss_result_free(apss.ctypes.byref(r))
    return info


def leg_mixed(hip, stream):
    """Mixed-precision configuration: plain sparse path vs exact int8 head vs BF16 head."""
    t0 = time.time()
    rp, col, val, d = apss.workload("mixed")
    n = len(rp) - 1
    log(f"mixed: generated n={n} nnz={len(col)} in {time.time() - t0:.1f}s")
    X = apss.DeviceCSR(hip, rp, col, val, d)
    Xc = X.csr()
    out = {"workload": f"mixed: n={n} d={d} nnz={len(col)} Zipf(1.2) uniform, t=0.5, p=0"}
    forms = (("plain", dict(mfma_head=-1)), ("int8_head", dict(mfma_head=0)),
             ("bf16_head", dict(mfma_head=0, precision=apss.PREC_BF16_HEAD)))
    counts = {}
    for name, o in forms:
        run_device(Xc, 0.5, 0, stream, **o)  # warm-up (first-call costs)
        info = run_device(Xc, 0.5, 0, stream, **o)
        counts[name] = info["count"]
        out[name] = {"wall_ms": round(info["ms_total"], 1), "scan_ms": round(info["ms_scan"], 1),
                     "pairs": info["count"], "rescored": info["band"], "mfma_features": info["mfma_features"]}
        log(f"mixed {name}: {info['ms_total']:.1f} ms, {info['count']} pairs")
    out["pair_sets_equal_counts"] = len(set(counts.values())) == 1
    out["bf16_speedup_vs_plain"] = round(out["plain"]["wall_ms"] / out["bf16_head"]["wall_ms"], 3)
    out["int8_speedup_vs_plain"] = round(out["plain"]["wall_ms"] / out["int8_head"]["wall_ms"], 3)
    X.release()
    return out


def leg_split(n=50_000):
    """Split sweep on a slice of the large workload: time and pair count per p.
    With p > 0 the ids below p run the reference's prefix-norm filter and
    forward-list head dot for every surviving pair, whose cost grows with p,
    so the sweep uses the first 50,000 rows; the pair set must not change."""
    rp, col, val, d = apss.workload("large", n=n)
    out = {"rows": n, "sweep": []}
    for t in (0.5, 0.9):
        counts = set()
        for p in (0, 1024, 2048, 4096, 8192, 16384):
            _, info = apss.run(rp, col, val, d, t, p, output_on_device=1)
            counts.add(info["count"])
            out["sweep"].append({"t": t, "p": p, "scan_ms": round(info["ms_scan"], 1),
                                 "wall_ms": round(info["ms_total"], 1), "pairs": info["count"],
                                 "head_pairs": info["head_pairs"]})
            log(f"split t={t} p={p}: {info['ms_scan']:.1f} ms, {info['count']} pairs")
        out[f"pairs_independent_of_p_t{t}"] = len(counts) == 1
    return out


def leg_batched():
    """256 instances in one launch sequence vs the same instances one by one."""
    t0 = time.time()
    inst = apss.batched_instances()
    rows = sum(len(x[0]) - 1 for x in inst)
    nnz = sum(len(x[1]) for x in inst)
    cells = sum((len(x[0]) - 1) * (len(x[0]) - 2) / 2.0 for x in inst)
    log(f"batched: generated {len(inst)} instances, {rows} rows, {nnz} nnz in {time.time() - t0:.1f}s")
    ts = [0.8] * len(inst)
    apss.run_batched(inst[:4], ts[:4])  # warm-up
    t1 = time.perf_counter()
    res = apss.run_batched(inst, ts, p=-1)
    batched_s = time.perf_counter() - t1
    tuned = res[0][1] if res else {}
    pairs_b = sum(info["count"] for _, info in res)
    # the same batch again: the autotuner's choice for this job shape is kept
    # by the library, so the second call skips the timing of candidates
    t1 = time.perf_counter()
    res2 = apss.run_batched(inst, ts, p=-1)
    warm_s = time.perf_counter() - t1
    warm = res2[0][1] if res2 else {}
    pairs_w = sum(info["count"] for _, info in res2)
    del res2
    t1 = time.perf_counter()
    pairs_s = 0
    for x in inst:
        _, info = apss.run(*x, 0.8)
        pairs_s += info["count"]
    single_s = time.perf_counter() - t1
    return {"instances": len(inst), "rows": rows, "nnz": nnz, "pair_cells": cells, "t": 0.8,
            "batched_wall_ms": round(batched_s * 1e3, 1), "one_by_one_wall_ms": round(single_s * 1e3, 1),
            "speedup": round(single_s / batched_s, 3), "pairs": pairs_b, "pairs_equal": pairs_b == pairs_s,
            "pair_cells_per_s": cells / batched_s,
            "warm_wall_ms": round(warm_s * 1e3, 1), "warm_speedup": round(single_s / warm_s, 3),
            "warm_pairs_equal": pairs_w == pairs_b, "warm_tune_ms": round(warm.get("ms_tune", 0.0), 2),
            "graph_ms": round(tuned.get("ms_scan", 0.0), 1),
            "autotune": {"ms": round(tuned.get("ms_tune", 0.0), 1), "p": tuned.get("p_used"),
                         "mfma_features": tuned.get("mfma_features"),
                         "bitmap_features": tuned.get("bitmap_features")}}


def leg_serving(args):
    """Continuous-batching server over a resident corpus under Poisson load."""
    t0 = time.time()
    rp, col, val, d = apss.workload("serving", n=args.serving_rows)
    n = len(rp) - 1
    log(f"serving: generated corpus n={n} nnz={len(col)} in {time.time() - t0:.1f}s")
    t0 = time.time()
    ix = apss.Index(rp, col, val, d, p=0)
    build_s = time.time() - t0
    log(f"serving: index built in {build_s:.1f}s")
    del rp, col, val
    pool_n = 1 << 16
    qrp, qcol, qval, _ = apss.workload("serving", n=pool_n, seed_offset=1000)
    irp, icol, ival, _ = apss.workload("serving", n=pool_n, seed_offset=2000)
    rng = np.random.default_rng(0x5E4)

    def rows_slice(arp, acol, aval, a, b):
        return ((arp[a:b + 1] - arp[a]).astype(np.int64), acol[arp[a]:arp[b]].copy(), aval[arp[a]:arp[b]].copy())

    def make_batch():
        size = int(round(math.exp(rng.uniform(0.0, math.log(4096.0)))))
        a = int(rng.integers(0, pool_n - size))
        return rows_slice(qrp, qcol, qval, a, a + size), float(rng.choice([0.6, 0.8])), size

    srv = apss.Server(ix, slots=args.slots, mem_budget=args.serving_budget)
    out = {"corpus_rows": n, "features": d, "index_build_s": round(build_s, 2), "slots": args.slots,
           "mem_budget": args.serving_budget}
    try:
        # capacity: closed loop, one request at a time
        lat, rows_done, t1 = [], 0, time.perf_counter()
        for _ in range(args.serving_probe):
            q, t, size = make_batch()
            tk = srv.submit(*q, t)
            srv.wait_info(tk, keep_pairs=False)
            rows_done += size
        closed_s = time.perf_counter() - t1
        cap = args.serving_probe / closed_s
        out["capacity_batches_per_s"] = round(cap, 2)
        out["closed_loop_rows_per_s"] = round(rows_done / closed_s, 1)
        log(f"serving: capacity {cap:.1f} batches/s")
        ins_next = [0]
        sweeps = []
        for frac in args.serving_rates:
            rate = frac * cap
            nreq = int(min(args.serving_max, max(args.serving_min, rate * args.serving_seconds)))
            batches = [make_batch() for _ in range(nreq)]
            gaps = rng.exponential(1.0 / rate, size=nreq)
            tickets, lats, errs, parts = [], [], [], []
            done = threading.Event()
            cv = threading.Condition()

            def waiter():
                k = 0
                while True:
                    with cv:
                        while k >= len(tickets) and not done.is_set():
                            cv.wait(0.05)
                        if k >= len(tickets) and done.is_set():
                            return
                        tk = tickets[k]
                    k += 1
                    try:
                        _, info, ms = srv.wait_info(tk, keep_pairs=False)
                        lats.append(ms)
                        parts.append((info["ms_queue"], ms - info["ms_queue"] - info["ms_handover"],
                                      info["ms_handover"]))
                    except apss.ApssError as e:
                        errs.append({"code": e.code, "detail": str(e)})

            th = threading.Thread(target=waiter)
            th.start()
            ins_rows, sub_rows, inserts = 0, 0, 0
            t1 = time.perf_counter()
            due = t1
            for (q, t, size), g in zip(batches, gaps):
                due += g
                now = time.perf_counter()
                if due > now:
                    time.sleep(due - now)
                tk = srv.submit(*q, t)
                with cv:
                    tickets.append(tk)
                    cv.notify()
                sub_rows += size
                # 1% of the submitted rows are inserted into the corpus (256-row chunks)
                while (sub_rows // 100) - ins_rows >= 256:
                    a = ins_next[0] % (pool_n - 256)
                    ix.insert(*rows_slice(irp, icol, ival, a, a + 256))
                    ins_next[0] += 256
                    ins_rows += 256
                    inserts += 1
            done.set()
            th.join()
            wall = time.perf_counter() - t1
            la = np.array(lats) if lats else np.zeros(1)
            pa = np.array(parts) if parts else np.zeros((1, 3))

            def pct(a, q):
                return round(float(np.percentile(a, q)), 2)

            sw = {"load": frac, "offered_batches_per_s": round(rate, 2), "requests": nreq, "completed": len(lats),
                  "errors": len(errs), "p50_ms": pct(la, 50), "p99_ms": pct(la, 99), "mean_ms": round(float(la.mean()), 2),
                  # latency split: queue wait (submit -> launch start), launch (query upload,
                  # band build, scans and result copy of every corpus segment), hand-over
                  # (split of a packed batch's pairs by request)
                  "breakdown_ms": {name: {"p50": pct(pa[:, c], 50), "p99": pct(pa[:, c], 99)}
                                   for c, name in enumerate(("queue", "launch", "handover"))},
                  "throughput_batches_per_s": round(len(lats) / wall, 2),
                  "query_rows_per_s": round(sub_rows / wall, 1), "inserted_rows": ins_rows, "inserts": inserts}
            if errs:
                codes = {}
                for e in errs:
                    codes[str(e["code"])] = codes.get(str(e["code"]), 0) + 1
                sw["error_codes"] = codes
                sw["error_details"] = sorted({e["detail"] for e in errs})[:4]
            sweeps.append(sw)
            log(f"serving load {frac}: p50 {sw['p50_ms']} ms p99 {sw['p99_ms']} ms, "
                f"{sw['throughput_batches_per_s']} batches/s, {len(errs)} errors {sw.get('error_details', '')}")
        out["sweeps"] = sweeps
        out["index_rows_after"] = ix.rows()
    finally:
        srv.close()
        ix.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1_000_000, help="rows (default: the full large workload)")
    ap.add_argument("--t", type=float, default=0.5)
    ap.add_argument("--p", type=int, default=0)
    ap.add_argument("--mfma-head", type=int, default=0, help="0: dense MFMA head on, -1: off")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--skip", default="", help="comma list of legs to skip: " + ",".join(ALL_LEGS))
    ap.add_argument("--only-main", action="store_true", help="skip every leg beside the headline steps")
    ap.add_argument("--copy-check-mib", type=int, default=0,
                    help="before the steps, copy a buffer of this many MiB device to device three "
                         "times (a kernel of known traffic that scales HBM counters)")
    ap.add_argument("--slots", type=int, default=2)
    ap.add_argument("--serving-rows", type=int, default=2_000_000)
    ap.add_argument("--serving-budget", type=int, default=64 << 30)
    ap.add_argument("--serving-probe", type=int, default=24)
    ap.add_argument("--serving-rates", type=float, nargs="*", default=[0.25, 0.5, 0.75, 0.9])
    ap.add_argument("--serving-seconds", type=float, default=6.0)
    ap.add_argument("--serving-min", type=int, default=40)
    ap.add_argument("--serving-max", type=int, default=400)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              f"(the tile-pair space would shard across GPUs by work items)", file=sys.stderr)
        sys.exit(2)
    skip = set(ALL_LEGS) if args.only_main else {s for s in args.skip.split(",") if s}

    t0 = time.time()
    rp, col, val, d = apss.workload("large", n=args.n)
    n, nnz = len(rp) - 1, len(col)
    log(f"generated large workload n={n} nnz={nnz} in {time.time() - t0:.1f}s")

    hip = apss.Hip()
    if args.copy_check_mib > 0:
        nb = args.copy_check_mib << 20
        a, b = hip.malloc(nb), hip.malloc(nb)
        for _ in range(3):
            hip.d2d(b, a, nb)
        hip.sync()
        hip.free(a)
        hip.free(b)
        log(f"copy check: 3 x {args.copy_check_mib} MiB device-to-device")
    X = apss.DeviceCSR(hip, rp, col, val, d)
    stream = hip.stream()
    e0, e1 = hip.event(), hip.event()
    Xc = X.csr()
    opts = dict(mfma_head=args.mfma_head)

    for w in range(args.warmup):
        info = run_device(Xc, args.t, args.p, stream, **opts)
        log(f"warmup {w}: {info['ms_total']:.1f} ms, pairs {info['count']}")
    hip.sync()
    hip.record(e0, stream)
    infos = []
    for s in range(args.steps):
        infos.append(run_device(Xc, args.t, args.p, stream, **opts))
        log(f"step {s}: total {infos[-1]['ms_total']:.1f} ms scan {infos[-1]['ms_scan']:.1f} ms")
    hip.record(e1, stream)
    hip.sync()
    ms = hip.elapsed_ms(e0, e1) / args.steps
    last = infos[-1]
    cells = n * (n - 1) / 2.0

    # band-scan roofline: algorithmic bytes / hipEvent kernel time (same stream).
    # The timed steps run the band kernel without its work counters; one more
    # run of the same job with count_work = 1 (untimed) counts the visits and
    # products the byte model needs (the work is the same: the pair set and
    # the tiling do not change).
    counted = run_device(Xc, args.t, args.p, stream, count_work=1, **opts)
    assert counted["count"] == last["count"] and counted["tile_pairs"] == last["tile_pairs"]
    last = dict(last, list_visits=counted["list_visits"], products=counted["products"])
    terms = alg_terms(last)
    ab = sum(terms.values())
    scan_ms = float(np.mean([i["ms_scan"] for i in infos]))
    achieved = ab / (scan_ms * 1e-3) / 1e9
    prof = traffic_from_profiles()
    traffic = prof.get("hbm_bytes_per_launch") if prof else None
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "k_band_scan", "kernel_ms": round(scan_ms, 2), "alg_bytes_per_launch": ab,
                "alg_bytes_terms": {k: round(v) for k, v in terms.items()},
                "counters": {k: last[k] for k in ("list_visits", "band_entries", "band_records", "walk_entries",
                                                  "walk_scatter_entries", "cursor_products", "dense_rows",
                                                  "seg_row_bytes", "tile_pairs", "products")},
                "pairs_out": last["count"] + last["band"]}
    if prof:
        roofline["traffic_over_model"] = round(traffic / ab, 3)
        roofline["traffic_uncalibrated"] = prof.get("raw_bytes_per_launch")
        roofline["traffic_source"] = "profiles/" + prof.get("source", "band_scan_large_traffic.md")

    extra = {}
    if "t09" not in skip:
        info9 = run_device(Xc, 0.9, args.p, stream, **opts)
        extra["t0.9"] = {"wall_ms": round(info9["ms_total"], 1), "scan_ms": round(info9["ms_scan"], 1),
                         "pairs": info9["count"], "pairs_per_s": cells / (info9["ms_total"] * 1e-3)}
        log(f"t=0.9: {info9['ms_total']:.1f} ms, {info9['count']} pairs")
    if "variants" not in skip:
        v = {}
        for name, o in (("plain", dict(mfma_head=-1)),
                        ("bf16_head", dict(mfma_head=0, precision=apss.PREC_BF16_HEAD))):
            iv = run_device(Xc, args.t, args.p, stream, **o)
            v[name] = {"wall_ms": round(iv["ms_total"], 1), "scan_ms": round(iv["ms_scan"], 1),
                       "pairs": iv["count"], "pairs_equal": iv["count"] == last["count"],
                       "pairs_per_s": cells / (iv["ms_total"] * 1e-3)}
            log(f"variant {name}: {iv['ms_total']:.1f} ms, {iv['count']} pairs")
        v["headline_speedup_vs_plain"] = round(v["plain"]["wall_ms"] / ms, 3)
        extra["large_variants"] = v

    if "h2h" not in skip:
        h2h = host_to_host(rp, col, val, d, args.t, args.p, **opts)
        extra["host_to_host"] = h2h
        log(f"host to host: {h2h['wall_ms']} ms, {h2h['pairs']} pairs")
    cpu = None
    if "cpu" not in skip:
        log("cpu baseline ...")
        cpu = cpu_baseline(rp, col, val, d, args.t, args.cpu_seconds)
        log("cpu serial inverted-index search ...")
        extra["cpu_serial"] = cpu_serial(rp, col, val, d, args.t, args.cpu_seconds / 3)
    X.release()
    del rp, col, val

    for name, fn in (("split", leg_split), ("mixed", lambda: leg_mixed(hip, stream)), ("batched", leg_batched),
                     ("serving", lambda: leg_serving(args))):
        if name in skip:
            continue
        try:
            extra[name] = fn()
        except Exception as e:  # noqa: BLE001 -- a failed leg is reported, the line still prints
            log(f"{name} leg failed: {e!r}")
            extra[name] = {"error": repr(e)}

    out = {
        "metric": "APSS pair-space throughput n(n-1)/2 / wall, t=0.5",
        "value": cells / (ms * 1e-3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"large: n={n} d={d} nnz={nnz} Zipf(1.05) lognormal, t={args.t}, p={args.p}",
                   "rows": n, "features": d, "nnz": nnz, "t": args.t, "p": args.p,
                   "mfma_head": "int8-limb exact" if args.mfma_head == 0 else "off"},
        "result_pairs": last["count"],
        "result_pairs_per_s": last["count"] / (ms * 1e-3),
        "phases_ms": {k: round(last[k], 2) for k in ("ms_upload", "ms_index", "ms_scan", "ms_rescore")},
        "rescored_pairs": last["band"],
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
