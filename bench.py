#!/usr/bin/env python3
"""Headline benchmark: range + KNN queries/s on the N=10M, D=8 hierarchy (MI355X).

A step is one pass of the query path over one batch of synthetic queries
(range boxes with selectivity log-uniform in [1e-5, 1e-2] and KNN points with
K in {10, 100}), with the index and the queries resident in HBM, producing the
exact answers (range: sorted id sets; KNN: top-K ids and squared distances).

Reported:
  value           queries/s of the fastest path measured (whole job, 1 GPU)
  plain           the per-query GPU traversal (QueryHandler -> QueryEngine -> L0)
  roofline        the batched kernel with the largest time per step (the exact
                  box filter against VALU compare issue, or the hit sort
                  against HBM), from HIP events on the library's stream; the
                  other one and the plain path's dominant HBM kernel beside it
  cpu_baseline    the OpenMP CPU traversal (the oracle) on the host cores
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "t-blaeq_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

METRIC = "range+KNN queries/s at N=10M D=8, batched/learned vs plain; p99 latency"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# VALU lane-operations per second (non-packed FP64/FP32 compares): 256 CUs x 4
# SIMDs x 16 lanes x 2.4 GHz; half the FP64 FMA flop rate of 78.6 TFLOP/s
VALU_PEAK_TOPS = 39.3
PARAMS = (("knn_seed", 1), ("knn_wg_per_cu", 2), ("knn_mode", 3), ("range_blocks_per_cu", 4),
          ("step_blocks_per_cu", 5), ("radix_blocks_per_cu", 6), ("hit_stage", 7), ("range_row_queries", 8), ("knn_row_queries", 9), ("knn_leaf_refine", 10))


def final_traffic(cfg):
    """HBM bytes per launch of the exact box filter, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary for this workload (None if absent)."""
    path = os.path.join(ROOT, "profiles", "c2_final_filter_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("N") != cfg["N"] or t.get("D") != cfg["D"]:
        return None
    return t.get("bytes_per_launch")


def sort_traffic(cfg, passes):
    """HBM bytes per hit sort (counts + passes), from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary for this workload (None if absent)."""
    path = os.path.join(ROOT, "profiles", "c2_hit_sort_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("N") != cfg["N"] or t.get("D") != cfg["D"] or t.get("passes") != passes:
        return None
    return t.get("bytes_per_launch")


def _json_default(o, path=None):
    """NumPy values in the result record become plain numbers and lists."""
    if isinstance(o, np.ndarray):
        log("note: an array in the result record was written as a list, shape %s" % (o.shape,))
        return o.tolist()
    if isinstance(o, np.generic):
        return o.item()
    raise TypeError("not serialisable: %r" % type(o))


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def kernel_bytes(D, levels_surv, levels_out, coarsest, hits):
    """Algorithmic HBM bytes per kernel family for one query (FP64 values, 8-byte ids).

    prune/compact at level l:  p_l (8D + 16) + 8 s_l
    prolongation at level l:   s_l (8D + 16) + n_l (16D + 16)
    box filter (range refine): p (8D + 8) + 8 hits
    """
    b = {"compact": 0.0, "prolong": 0.0, "filter": 0.0}
    p = coarsest
    for s, n in zip(levels_surv, levels_out):
        b["compact"] += p * (8 * D + 16) + 8 * s
        b["prolong"] += s * (8 * D + 16) + n * (16 * D + 16)
        p = n
    b["filter"] = p * (8 * D + 8) + 8 * hits
    return b


def run_plain(h, boxes, knn_q, Ks, steps, warmup, timing):
    """Per-query GPU traversal + exact refine, one query at a time."""
    import tblaeq  # noqa: F401

    D = h.D
    batches = [("range", boxes, 0)] + [("knn", knn_q[Ks == k], int(k)) for k in np.unique(Ks)]
    for _ in range(warmup):
        for t, q, k in batches:
            h.query(t, q[: min(4, len(q))], K=k, refine=2)
    timing.reset()
    us = []
    totals = {"compact": 0.0, "prolong": 0.0, "filter": 0.0}
    t0 = time.perf_counter()
    results = []
    for _ in range(steps):
        for t, q, k in batches:
            results.append((t, h.query(t, q, K=k, refine=2)))
    wall = time.perf_counter() - t0
    nq = 0
    for t, r in results:
        for i in range(r.nq):
            info = r.info(i)
            us.append(info["us"])
            s, o = r.levels(i)
            b = kernel_bytes(D, s, o, h.mesh_sizes[-1], info["refined"] if t == "range" else 0)
            for key in totals:
                totals[key] += b[key]
            nq += 1
    return nq, wall, np.asarray(us), totals


def serving_mix(data, seed, mesh_sizes, vmin, vmax, n=2048, subsample=200_000):
    """The serving request mix over an index whose data (or a row sample of it) is `data`: n range
    boxes with selectivity log-uniform in 1e-6..1e-3, n KNN points with K log-uniform in 1..256."""
    from tblaeq import workload as W

    boxes, _ = W.range_boxes(data, n, seed + 3, 1e-6, 1e-3, subsample=subsample)
    pts = W.knn_points(data, n, seed + 4, mesh_sizes, vmin, vmax)
    rng = np.random.default_rng(seed + 5)
    Ks = np.clip(np.round(np.exp(rng.uniform(0.0, np.log(256.0), n))), 1, 256).astype(np.int64)
    return boxes, pts, Ks


def run_serving(h, boxes, pts, Ks, rates, seconds, log, tag="serving"):
    """Open-loop Poisson request stream through the continuous-batching server: 70% range, 30%
    KNN (K above the batched kernels' limit runs per query on the KNN worker), at each offered
    rate for `seconds`.  A consumer thread waits for the requests in submission order and
    releases each result as it completes, so results do not pile up on the host."""
    import queue
    import random
    import threading

    import tblaeq

    n = len(boxes)
    srv = tblaeq.Server(h, max_batch=4096)
    out = []
    try:
        for i in range(64):
            srv.wait(srv.submit("range", boxes[i % n]))
            srv.wait(srv.submit("knn", pts[i % n], int(Ks[i % n])))
        for rate in rates:
            r = random.Random(int(rate))
            srv.reset_stats()
            pending = queue.Queue()
            failures = []

            def consume():
                while True:
                    t = pending.get()
                    if t is None:
                        return
                    try:
                        srv.wait(t, 300.0)
                        srv.release(t)
                    except Exception as e:  # reported after the run
                        failures.append(e)

            th = threading.Thread(target=consume)
            th.start()
            sent = 0
            start = time.perf_counter()
            t_next = start
            i = 0
            try:
                while True:
                    now = time.perf_counter()
                    if now - start >= seconds:
                        break
                    if now < t_next:
                        if t_next - now > 2e-4:
                            time.sleep(t_next - now - 1e-4)
                        continue
                    if r.random() < 0.7:
                        pending.put(srv.submit("range", boxes[i % n]))
                    else:
                        pending.put(srv.submit("knn", pts[i % n], int(Ks[i % n])))
                    sent += 1
                    i += 1
                    t_next += r.expovariate(rate)
                offered_s = time.perf_counter() - start
            finally:
                pending.put(None)
                th.join()
            done_s = time.perf_counter() - start
            if failures:
                raise failures[0]
            st = srv.stats()
            rec = {"offered_rps": round(sent / offered_s, 1), "completed_rps": round(sent / done_s, 1),
                   "requests": sent, "p50_us": round(st["p50_us"], 1), "p99_us": round(st["p99_us"], 1),
                   "mean_batch": round(st["mean_batch"], 1), "batches": int(st["batches"]),
                   "deferred_by_hbm": int(st["deferred"])}
            for kind in ("range", "knn"):
                ks = srv.kind_stats(kind)
                rec[kind] = {"completed": int(ks["completed"]), "batches": int(ks["batches"]),
                             "reruns": int(ks["reruns"]), "per_query": int(ks["traversals"]),
                             "busy_s": round(ks["busy_us"] / 1e6, 3)}
            log(tag, rec)
            out.append(rec)
    finally:
        srv.close()
    return out


def run_plain_mix(h, boxes, pts, Ks, n, seed):
    """The serving mix one request at a time through the per-query traversal with the exact refine
    (the reference's front end serves requests one by one): queries/s and service-time percentiles."""
    rng = np.random.default_rng(seed)
    D = h.D
    lat = []
    t0 = time.perf_counter()
    for i in range(n):
        j = int(rng.integers(len(boxes)))
        s = time.perf_counter()
        if rng.random() < 0.7:
            r = h.range(boxes[j][:D], boxes[j][D:], save_fine_mesh=False)
        else:
            r = h.knn(pts[j], int(Ks[j]), save_fine_mesh=False)
        r.close()
        lat.append(time.perf_counter() - s)
    wall = time.perf_counter() - t0
    lat = np.array(lat) * 1e6
    return {"requests": n, "qps": round(n / wall, 1), "p50_us": round(float(np.percentile(lat, 50)), 1),
            "p99_us": round(float(np.percentile(lat, 99)), 1)}


BATCH_KINDS = (("range_prune", 8), ("range_final", 9), ("range_rows", 20), ("group_scan", 10),
               ("group_scatter", 11), ("radix_sort", 12), ("knn_levels", 13), ("knn_last_passes", 17),
               ("knn_select", 18), ("knn_refine", 15), ("knn_rows", 21), ("knn_merge", 16))


def batch_kernels(h, boxes, pts, Ks, reps, log):
    """Per-kernel times (ms, HIP events on the library's stream, plain launches, mean of `reps`
    runs after one warm-up) of one range batch of `boxes` and one KNN batch of `pts` with K `Ks`,
    on the handler's batched engine with its current launch parameters."""
    import tblaeq

    timing = tblaeq.Timing(h.ctx())
    h.batch_use_graph(False)
    out = {}
    try:
        for kind, args in (("range", (boxes,)), ("knn", (pts, Ks))):
            h.batch_set(kind, *args)
            res = h.batch_run()
            acc = {name: 0.0 for name, _ in BATCH_KINDS}
            wall = []
            for _ in range(reps):
                timing.enable(True)
                timing.reset()
                t0 = time.perf_counter()
                h.batch_run()
                wall.append(time.perf_counter() - t0)
                for name, k in BATCH_KINDS:
                    acc[name] += timing.get(k)[1]
            timing.enable(False)
            out[kind] = {"queries": len(args[0]), "batch_ms": round(float(np.mean(wall)) * 1e3, 3),
                         "results": int(res), "work": h.batch_work(),
                         "kernels_ms": {k: round(v / reps, 3) for k, v in acc.items() if v > 0}}
            log("batch kernels", kind, out[kind])
    finally:
        h.batch_use_graph(True)
    return out


def run_c5(n, rates, seconds, plain_n, log, tune=True):
    """The serving configuration: N FP64 points (500M by default), D = 8, height 5, ratios
    100,50,20,10, every level resident under policy L0, built by the parallel synthetic builder
    (the generator's shape from per-chunk streams).  The serving mix comes from a 1M-row sample of
    the data; the plain per-query traversal serves a sample of it one request at a time, then the
    server takes the open-loop Poisson sweep."""
    import tblaeq

    cfg = dict(N=n, D=8, vmin=1.0, vmax=100.0, is_int=False, ratios=(100, 50, 20, 10), seed=20260330,
               sigma_div=2.2)
    log("C5: building the synthetic index on host threads", cfg)
    t0 = time.perf_counter()
    h = tblaeq.Handler.synthetic(cfg["N"], cfg["D"], cfg["vmin"], cfg["vmax"], cfg["is_int"], cfg["ratios"],
                                 cfg["seed"], cfg["sigma_div"], threads=0)
    try:
        build_s = time.perf_counter() - t0
        log("C5: built in %.1fs, mesh sizes %s" % (build_s, h.mesh_sizes))
        t0 = time.perf_counter()
        h.prepare("L0")
        prep_s = time.perf_counter() - t0
        log("C5: resident under L0 in %.1fs" % prep_s)
        rng = np.random.default_rng(cfg["seed"] + 20)
        sample = h.data_rows(np.sort(rng.integers(0, n, min(n, 1_000_000))))
        boxes, pts, Ks = serving_mix(sample, cfg["seed"], h.mesh_sizes, cfg["vmin"], cfg["vmax"], n=1024,
                                     subsample=sample.shape[0])
        log("C5: request mix ready")
        plain = run_plain_mix(h, boxes, pts, Ks, plain_n, cfg["seed"] + 23)
        log("C5: plain per-query service", plain)
        # launch parameters tuned on the request mix (untimed set-up); the server's engines take them
        ns = 512  # sample of the mix: half the serving batches' size, within the sample engine's capacities
        h.batch_init(ns, 1 << 27, 1 << 27)
        sel = np.nonzero(Ks[:ns] <= 200)[0]  # within the batched KNN kernels' limit
        tuned = h.batch_autotune(boxes[:ns], pts[sel], Ks[sel].astype(np.uint64), reps=3) if tune else None
        params = {name: h.batch_get_param(pid) for name, pid in PARAMS}
        log("C5: autotune", tuned, params)
        # one 1024-request range batch and one KNN batch of the mix (K capped at 200, the batched
        # kernels' range), per kernel, with the tuned parameters
        h.batch_init(len(boxes), 1 << 27, 1 << 27)  # a new engine: the tuned values are set again
        for name, pid in PARAMS:
            h.batch_set_param(pid, params[name])
        kernels = batch_kernels(h, boxes, pts, np.minimum(Ks, 200).astype(np.uint64), 3, log)
        runs = run_serving(h, boxes, pts, Ks, rates, seconds, log, tag="C5 serving")
        out = {"workload": "synthetic hierarchy N=%d D=8 height 5 ratios 100,50,20,10 sigma/2.2 seed %d (parallel "
                           "builder), policy L0; open-loop Poisson stream, 70%% range (selectivity log-uniform "
                           "1e-6..1e-3), 30%% KNN (K log-uniform 1..256)" % (n, cfg["seed"]),
               "build_s": round(build_s, 1), "prepare_s": round(prep_s, 1), "plain": plain,
               "batch_kernels": kernels, "runs": runs,
               "launch_params": params,
               "autotune_range_ms_before_after": tuned and [round(float(x), 3) for x in tuned["range_ms"]],
               "autotune_knn_ms_before_after": tuned and [round(float(x), 3) for x in tuned["knn_ms"]],
               "note": "the server runs the batched engine with the seed method (no learned radii)"}
        log("C5", out)
        return out
    finally:
        h.close()


FP32_MFMA_PEAK_TFLOPS = 157.3


def run_build(n, D, ratios, cpu_n, log):
    """GPU index build (k-means per level on the matrix cores + device CSC) of synthetic clustered
    data, and the host build of a subsample beside the GPU build of the same subsample."""
    import tblaeq

    def data_of(m, seed):
        rng = np.random.default_rng(seed)
        k0 = max(1, m // ratios[0])
        cent = rng.uniform(1.0, 100.0, (k0, D))
        return np.ascontiguousarray(cent[rng.integers(0, k0, m)] + rng.normal(0.0, 1.5, (m, D)))

    x = data_of(n, 11)
    t0 = time.perf_counter()
    hb = tblaeq.Handler.build(x, ratios, "bench", force_cpu=False)
    gpu_s = time.perf_counter() - t0
    st = hb.build_stats()
    hb.close()
    dp = 4
    while dp < D:
        dp *= 2
    ms = float(st[0, 3])
    flops = float(st[0, 2]) * 2 * dp
    lvl = [{"iterations": int(r[0]), "rescored_points": int(r[1]), "score_evals": int(r[2]),
            "assign_ms": int(r[3])} for r in st]
    xs = data_of(cpu_n, 12)
    t0 = time.perf_counter()
    hg = tblaeq.Handler.build(xs, ratios, "bench-gpu", force_cpu=False)
    gpu_small_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    hc = tblaeq.Handler.build(xs, ratios, "bench-cpu", force_cpu=True)
    cpu_small_s = time.perf_counter() - t0
    # same hierarchy from both builds (the tests check labels and centroids bit for bit)
    same = hg.mesh_sizes == hc.mesh_sizes
    hg.close()
    hc.close()
    out = {"workload": "k-means index build of N=%d D=%d clustered points (%d centres, sigma 1.5), ratios %s, "
                       "k-means++ seed 42, at most 7 Lloyd iterations per level" % (n, D, max(1, n // ratios[0]),
                                                                                 ",".join(map(str, ratios))),
           "gpu_s": round(gpu_s, 3), "levels": lvl,
           "finest_level_ranking": {"bound": "mfma", "achieved": round(flops / (ms / 1e3) / 1e12, 2) if ms else None,
                                    "peak": FP32_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                                    "frac": round(flops / (ms / 1e3) / 1e12 / FP32_MFMA_PEAK_TFLOPS, 4) if ms else None,
                                    "note": "FP32 matrix-core score ||c||^2 - 2x.c (2 x padded D flops per point x "
                                            "centre), assignment + FP64 re-score time"},
           "rescored_fraction": round(float(st[0, 1]) / max(1.0, float(n) * float(st[0, 0])), 6),
           "cpu_sample": {"N": cpu_n, "cpu_s": round(cpu_small_s, 3), "gpu_s": round(gpu_small_s, 3),
                          "speedup": round(cpu_small_s / gpu_small_s, 1),
                          "same_mesh_sizes": same}}
    log("build", out)
    return out


def run_c4(n, D, ratios, iters, seed, sub_n, log):
    """The index-build configuration: N points of the generator's hierarchy shape (the parallel
    synthetic builder, D dimensions, the given ratios) as FP32 values, built on the GPU with k-means||
    seeding and `iters` Lloyd iterations from `seed` (FP32 matrix-core assignment with the exact FP64
    re-score of near ties, CSC P-tensors on the device).  Beside it, the first sub_n points are built
    on the GPU with the CPU build's k-means++ and on the host with the same settings (their inertia
    must agree), and on the GPU with k-means|| (its inertia against k-means++'s)."""
    import tblaeq

    log("C4: generating %d x %d points" % (n, D))
    syn = tblaeq.Handler.synthetic(n, D, 1.0, 100.0, False, ratios, seed=20260331, sigma_div=3.0, threads=0)
    x = syn.data().astype(np.float32)
    syn.close()
    log("C4: data ready (FP32, %.1f GB)" % (x.nbytes / 1e9))
    tblaeq.set_progress(lambda m: log("C4:", m))
    try:
        t0 = time.perf_counter()
        hb = tblaeq.Handler.build(x, ratios, "c4", max_iter=iters, seed=seed, seeding=1)
        build_s = time.perf_counter() - t0
    finally:
        tblaeq.set_progress(None)
    st = hb.build_stats()
    inertia = hb.build_inertia()
    hb.close()
    dp = 4
    while dp < D:
        dp *= 2
    lv = []
    for r, iv in zip(st, inertia):
        ms = float(r[3])
        fl = float(r[2]) * 2 * dp
        lv.append({"iterations": int(r[0]), "rescored_points": int(r[1]), "score_evals": int(r[2]),
                   "assign_ms": int(ms), "candidates": int(r[5]), "inertia": float(iv),
                   "mfma_tflops": round(fl / (ms / 1e3) / 1e12, 2) if ms else None,
                   "mfma_frac": round(fl / (ms / 1e3) / 1e12 / FP32_MFMA_PEAK_TFLOPS, 4) if ms else None})
    xs = np.ascontiguousarray(x[:sub_n])
    t0 = time.perf_counter()
    g = tblaeq.Handler.build(xs, ratios, "c4-sub", max_iter=iters, seed=seed)
    g_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    c = tblaeq.Handler.build(xs, ratios, "c4-sub", force_cpu=True, max_iter=iters, seed=seed)
    c_s = time.perf_counter() - t0
    p = tblaeq.Handler.build(xs, ratios, "c4-sub", max_iter=iters, seed=seed, seeding=1)
    gi, ci, pi = g.build_inertia()[0], c.build_inertia()[0], p.build_inertia()[0]
    same_iters = bool(np.array_equal(g.build_stats()[:, 0], c.build_stats()[:, 0]))
    for hh in (g, c, p):
        hh.close()
    out = {"workload": "k-means index build of N=%d FP32 points, D=%d (the generator's hierarchy shape, parallel "
                       "builder), ratios %s, %d Lloyd iterations from seed %d, k-means|| seeding (oversampling "
                       "0.5 k, 4 rounds)" % (n, D, ",".join(map(str, ratios)), iters, seed),
           "gpu_s": round(build_s, 1), "levels": lv,
           "finest_level_ranking": {"bound": "mfma", "achieved": lv[0]["mfma_tflops"], "peak": FP32_MFMA_PEAK_TFLOPS,
                                    "unit": "TFLOP/s", "frac": lv[0]["mfma_frac"],
                                    "note": "FP32 matrix-core score ||c||^2 - 2x.c (2 x padded D flops per point x "
                                            "centre) over the Lloyd iterations, assignment + FP64 re-score time"},
           "subsample": {"N": sub_n, "gpu_kmeanspp_s": round(g_s, 2), "cpu_s": round(c_s, 2),
                         "relative_inertia_gpu_vs_cpu": abs(gi - ci) / ci, "same_iterations": same_iters,
                         "kmeans_parallel_over_kmeanspp_inertia": round(pi / gi, 5)}}
    log("C4", out)
    return out


def run_learned(h, data, cfg, knn_q, Ks, n_traces, epochs, plain_sample, log):
    """Learned KNN radius on this index: traces from the batched engine, BF16-MFMA
    training on the GPU, then the per-query traversal with and without the
    learned radius on a sample (exact answers either way) and the batched engine
    with learned radii against the seed method."""
    import tblaeq
    from tblaeq import learn

    t0 = time.perf_counter()
    X, T = learn.traces(h, n_traces, cfg["seed"] + 7, cfg["vmin"], cfg["vmax"])
    t1 = time.perf_counter()
    mlp = tblaeq.Mlp(h, seed=7)
    mlp.set_norm(cfg["vmin"], cfg["vmax"] - cfg["vmin"])
    losses = learn.fit(mlp, X, T, epochs=epochs, batch=1024, lr=2e-3)
    t2 = time.perf_counter()
    Y = mlp.forward(X[:20000])
    resid = T[:20000, 0] - Y[:, 0]
    factor = float(np.exp(np.quantile(resid, 0.9)))  # first radius covers 90% of the traces
    out = {"traces": n_traces, "trace_s": round(t1 - t0, 2), "train_s": round(t2 - t1, 2), "epochs": epochs,
           "loss_first_last": [round(float(losses[0]), 5), round(float(losses[-1]), 5)],
           "log_dk_residual_rms": round(float(np.sqrt(np.mean(resid ** 2))), 4), "radius_factor": round(factor, 4)}
    # per-query path, K = 10, sample of the benchmark's KNN points
    P = plain_sample
    pts = knn_q[:P]
    radii = learn.propose_radii(h, mlp, pts, np.full(P, 10), cfg["vmin"], cfg["vmax"], factor=factor)
    h.query("knn", pts[:4], K=10, refine=2)
    h.query_learned(pts[:4], 10, radii[:4])
    t0 = time.perf_counter()
    plain = h.query("knn", pts, K=10, refine=True)
    t1 = time.perf_counter()
    learned, fb = h.query_learned(pts, 10, radii)
    t2 = time.perf_counter()
    same = all(np.array_equal(plain.refined(i)[0], learned.refined(i)[0]) for i in range(P))
    fp = sum(plain.info(i)["fine"] for i in range(P))
    fl = sum(learned.info(i)["fine"] for i in range(P))
    out["per_query"] = {"queries": P, "K": 10, "plain_qps": round(P / (t1 - t0), 1),
                        "learned_qps": round(P / (t2 - t1), 1), "speedup": round((t1 - t0) / (t2 - t1), 2),
                        "retries_per_query": round(fb / P, 4), "prolonged_points_plain": int(fp),
                        "prolonged_points_learned": int(fl), "answers_identical": bool(same)}
    # batched engine: learned radii vs the seed method on the benchmark's KNN batch
    h.batch_set("knn", knn_q, Ks)
    h.batch_run()
    ref = h.batch_results()
    tb = []
    for _ in range(3):
        t0 = time.perf_counter()
        h.batch_run()
        tb.append(time.perf_counter() - t0)
    tl = []
    fbb = 0
    for _ in range(3):
        t0 = time.perf_counter()
        fbb = h.batch_knn_learned(mlp, factor=factor)
        tl.append(time.perf_counter() - t0)
    res = h.batch_results()
    same_b = all(np.array_equal(a[0], b[0]) for a, b in zip(ref, res))
    out["batched"] = {"queries": len(knn_q), "seed_ms": round(min(tb) * 1e3, 3),
                      "learned_ms": round(min(tl) * 1e3, 3), "fallback_rate": round(fbb / len(knn_q), 4),
                      "answers_identical": bool(same_b)}
    mlp.close()
    log("learned", out)
    return out


def run_c3(n_q, n_traces, epochs, plain_sample, log):
    """The learned-radius configuration: its own index (N = 10M FP64 points, D = 16, height 4,
    ratios 100,50,20), n_q KNN points (K 10 | 100) at N(0, 0.01 spacing) from data points, the
    16+32 -> 64 -> 64 -> 4 BF16-MFMA network trained on n_traces batched-engine traces.  Reports
    the miss rate (first proposed radius below the K-th neighbour distance), the fallback rates,
    and the speedups against the plain per-query traversal and the batched seed method; every
    learned answer is compared with the plain one."""
    import tblaeq
    from tblaeq import learn
    from tblaeq import workload as W

    cfg = W.C3
    log("C3: generating index", cfg)
    t0 = time.perf_counter()
    h = tblaeq.Handler.random(cfg["N"], cfg["D"], cfg["vmin"], cfg["vmax"], cfg["is_int"], cfg["ratios"],
                              cfg["seed"], cfg["sigma_div"])
    mlp = None
    try:
        h.prepare("L0")
        gen_s = time.perf_counter() - t0
        log("C3: index ready in %.1fs, mesh sizes %s" % (gen_s, h.mesh_sizes))
        data = h.data()
        pts = W.knn_points(data, n_q, cfg["seed"] + 12, h.mesh_sizes, cfg["vmin"], cfg["vmax"], noise=0.01)
        Ks = np.where(np.arange(n_q) % 2 == 0, 10, 100).astype(np.uint64)
        # at D = 16 a query keeps about a third of the last level's clusters: 2^29 candidate
        # records and pairs (about 20 GB of engine scratch) hold a whole 8192-point batch, where
        # 2^27 overflowed and split it into reruns (158 ms against 54 ms per batch)
        h.batch_init(max(n_q, 4096), 1 << 29, 1 << 29)
        t0 = time.perf_counter()
        X, T = learn.traces(h, n_traces, cfg["seed"] + 13, cfg["vmin"], cfg["vmax"])
        t1 = time.perf_counter()
        log("C3: %d traces in %.1fs" % (n_traces, t1 - t0))
        mlp = tblaeq.Mlp(h, seed=7)
        mlp.set_norm(cfg["vmin"], cfg["vmax"] - cfg["vmin"])
        losses = learn.fit(mlp, X, T, epochs=epochs, batch=1024, lr=2e-3)
        t2 = time.perf_counter()
        log("C3: trained %d epochs in %.1fs" % (epochs, t2 - t1))
        Y = mlp.forward(X[:20000])
        factor = float(np.exp(np.quantile(T[:20000, 0] - Y[:, 0], 0.9)))
        # held-out traces: radius and survivor-budget errors of the network
        Xh, Th = learn.traces(h, 8192, cfg["seed"] + 14, cfg["vmin"], cfg["vmax"])
        Yh = mlp.forward(Xh)
        out = {"workload": "hierarchy N=%d D=%d ratios %s sigma/%.1f seed %d; %d KNN points (K 10|100, "
                           "noise 0.01 spacing)" % (cfg["N"], cfg["D"], ",".join(map(str, cfg["ratios"])),
                                                    cfg["sigma_div"], cfg["seed"], n_q),
               "index_s": round(gen_s, 1), "traces": n_traces, "trace_s": round(t1 - t0, 2),
               "train_s": round(t2 - t1, 2), "epochs": epochs,
               "loss_first_last": [round(float(losses[0]), 5), round(float(losses[-1]), 5)],
               "radius_factor": round(factor, 4),
               "heldout_log_dk_rms": round(float(np.sqrt(np.mean((Th[:, 0] - Yh[:, 0]) ** 2))), 4),
               "heldout_budget_rms": [round(float(v), 4) for v in np.sqrt(np.mean((Th[:, 1:4] - Yh[:, 1:4]) ** 2, 0))],
               "budgets_note": "outputs 1..3 (log1p survivors of levels 0..2, /10) are trained and reported; "
                               "no traversal reads them"}
        # batched engine: seed method vs learned radii, the same answers
        h.batch_set("knn", pts, Ks)
        h.batch_run()
        ref = h.batch_results()
        tb, tl = [], []
        for _ in range(3):
            t0 = time.perf_counter()
            h.batch_run()
            tb.append(time.perf_counter() - t0)
        fbb = 0
        for _ in range(3):
            t0 = time.perf_counter()
            fbb = h.batch_knn_learned(mlp, factor=factor)
            tl.append(time.perf_counter() - t0)
        res = h.batch_results()
        same_b = all(np.array_equal(a[0], b[0]) and np.array_equal(a[1].view(np.uint64), b[1].view(np.uint64))
                     for a, b in zip(ref, res))
        radii = learn.propose_radii(h, mlp, pts, Ks, cfg["vmin"], cfg["vmax"], factor=factor)
        dk = np.array([np.sqrt(r[1][-1]) if len(r[1]) else 0.0 for r in ref])
        miss = float(np.mean(dk > radii))
        # per-query path: plain vs learned on a sample (half K=10, half K=100); the
        # learned traversal tries the radius x 1, 1.25, 1.25^2, then the full traversal
        P = plain_sample
        tp = tq = 0.0
        retries = 0
        same_p = True
        for K, sl in ((10, slice(0, 2 * P, 2)), (100, slice(1, 2 * P, 2))):
            sp, sr = pts[sl], radii[sl]
            h.query("knn", sp[:2], K=K, refine=2)
            t0 = time.perf_counter()
            plain = h.query("knn", sp, K=K, refine=True)
            t1 = time.perf_counter()
            lr, fb = h.query_learned(sp, K, sr)
            t2 = time.perf_counter()
            tp += t1 - t0
            tq += t2 - t1
            retries += fb
            same_p &= all(np.array_equal(plain.refined(i)[0], lr.refined(i)[0]) for i in range(len(sp)))
            plain.close()
            lr.close()
        plain_qps = 2 * P / tp
        learned_b_qps = n_q / min(tl)
        out["batched"] = {"queries": n_q, "seed_ms": round(min(tb) * 1e3, 3), "learned_ms": round(min(tl) * 1e3, 3),
                          "fallback_rate": round(fbb / n_q, 4), "answers_identical": bool(same_b),
                          "learned_qps": round(learned_b_qps, 1), "seed_qps": round(n_q / min(tb), 1)}
        out["miss_rate"] = round(miss, 4)
        out["per_query"] = {"queries": 2 * P, "plain_qps": round(plain_qps, 1), "learned_qps": round(2 * P / tq, 1),
                            "speedup": round(tp / tq, 2), "retries_per_query": round(retries / (2 * P), 4),
                            "full_traversal_rate": round(float(np.mean(dk[:2 * P] > radii[:2 * P] * 1.5625)), 4),
                            "answers_identical": bool(same_p)}
        out["speedup_vs_plain"] = {"batched_learned": round(learned_b_qps / plain_qps, 1),
                                   "per_query_learned": round(tp / tq, 2)}
        log("C3", out)
        return out
    finally:
        if mlp is not None:
            mlp.close()
        h.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nq", type=int, default=4096, help="range queries and KNN queries per batch (each)")
    ap.add_argument("--plain-sample", type=int, default=96, help="queries of each type for the plain path")
    ap.add_argument("--cpu-range", type=int, default=48)
    ap.add_argument("--cpu-knn", type=int, default=24)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--pair-cap", type=int, default=1 << 28)
    ap.add_argument("--hit-cap", type=int, default=1 << 28)
    ap.add_argument("--small", action="store_true", help="N=500k D=5 hierarchy (quick check)")
    ap.add_argument("--serve-rates", default="20000,100000",
                    help="offered request rates (req/s) of the serving run; empty to skip")
    ap.add_argument("--serve-seconds", type=float, default=1.0)
    ap.add_argument("--learn-traces", type=int, default=200000, help="0 skips the learned-radius run")
    ap.add_argument("--learn-epochs", type=int, default=20)
    ap.add_argument("--c3-queries", type=int, default=8192,
                    help="KNN points of the learned-radius run on its own N=10M D=16 index (0 skips it)")
    ap.add_argument("--c3-traces", type=int, default=200000)
    ap.add_argument("--c5-n", type=int, default=0,
                    help="points of the serving run on its own height-5 index (500000000 for the serving "
                         "configuration; 0 skips it: it takes several minutes)")
    ap.add_argument("--c5-rates", default="1000,5000,20000,50000,100000,200000")
    ap.add_argument("--c5-seconds", type=float, default=2.0)
    ap.add_argument("--c5-only", action="store_true", help="run only the serving configuration")
    ap.add_argument("--c4-n", type=int, default=0,
                    help="points of the index-build configuration's run (50000000 for it; 0 skips it: minutes)")
    ap.add_argument("--c4-iters", type=int, default=20)
    ap.add_argument("--c4-sub", type=int, default=20000, help="subsample built on the GPU and on the host")
    ap.add_argument("--c4-only", action="store_true", help="run only the index-build configuration")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch parameters (profiling runs)")
    ap.add_argument("--build-n", type=int, default=1_000_000, help="points of the GPU index build run (0 skips it)")
    ap.add_argument("--build-d", type=int, default=32)
    ap.add_argument("--build-cpu-n", type=int, default=50_000, help="subsample built on the CPU beside the GPU")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="after the run, copy a buffer of this many MiB once on the GPU (a kernel of known "
                         "traffic, to scale rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (queries would shard across "
              "replicated indexes; see the sharding notes). Run with --gpus 1.", file=sys.stderr)
        return 2

    import tblaeq
    from tblaeq import workload as W

    c5_rates = [float(x) for x in args.c5_rates.split(",") if x]
    if args.c4_only:
        c4 = run_c4(args.c4_n or 50_000_000, 32, (100, 50, 20), args.c4_iters, 7, args.c4_sub, log)
        print(json.dumps({"build_c4": c4}, default=_json_default), flush=True)
        return 0
    if args.c5_only:
        c5 = run_c5(args.c5_n or 500_000_000, c5_rates, args.c5_seconds, 64, log, tune=not args.no_tune)
        print(json.dumps({"serving_c5": c5}), flush=True)
        return 0

    cfg = dict(W.C1 if args.small else W.C2)
    log("generating index", cfg)
    t0 = time.perf_counter()
    h = tblaeq.Handler.random(cfg["N"], cfg["D"], cfg["vmin"], cfg["vmax"], cfg["is_int"], cfg["ratios"], cfg["seed"],
                              cfg["sigma_div"])
    h.prepare("L0")
    log("index ready in %.1fs, mesh sizes %s" % (time.perf_counter() - t0, h.mesh_sizes))
    data = h.data()
    boxes, sel = W.range_boxes(data, args.nq, cfg["seed"] + 1, progress=log)
    knn_q = W.knn_points(data, args.nq, cfg["seed"] + 2, h.mesh_sizes, cfg["vmin"], cfg["vmax"])
    Ks = np.where(np.arange(args.nq) < args.nq // 2, 10, 100)
    log("queries ready")

    timing = tblaeq.Timing(h.ctx())
    timing.enable(True)
    P = min(args.plain_sample, args.nq)
    half = P // 2
    pk = np.concatenate([knn_q[:half], knn_q[args.nq // 2: args.nq // 2 + (P - half)]])
    pK = np.concatenate([Ks[:half], Ks[args.nq // 2: args.nq // 2 + (P - half)]])
    nq, wall, us, bytes_tot = run_plain(h, boxes[:P], pk, pK, args.steps, args.warmup, timing)
    plain_qps = nq / wall
    kstats = {}
    for kind in ("compact_range", "compact_knn", "compact_box", "prolong", "tilesort", "merge", "step"):
        n, ms = timing.get(kind)
        kstats[kind] = (n, ms)
    log("plain: %d queries in %.3fs -> %.1f q/s; kernels %s" % (nq, wall, plain_qps, kstats))

    # roofline of the dominant HBM kernel family over the timed region
    fam = {"prolong": ("prolong", bytes_tot["prolong"]),
           "compact": ("compact_range", bytes_tot["compact"]),
           "filter": ("compact_box", bytes_tot["filter"])}
    best = max(fam, key=lambda f: kstats[fam[f][0]][1])
    kname, kbytes = fam[best]
    if best == "compact":  # range and KNN compaction share the byte model
        n_l = kstats["compact_range"][0] + kstats["compact_knn"][0]
        t_ms = kstats["compact_range"][1] + kstats["compact_knn"][1]
    else:
        n_l, t_ms = kstats[kname]
    per_launch_bytes = kbytes / max(n_l, 1)
    avg_s = t_ms / 1e3 / max(n_l, 1)
    achieved = per_launch_bytes / avg_s / 1e9 if avg_s > 0 else 0.0
    plain_roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": {"prolong": "k_prolong", "compact": "k_compact", "filter": "k_compact<2>"}[best],
                "launches": n_l, "avg_launch_us": round(avg_s * 1e6, 3), "bytes_per_launch": round(per_launch_bytes)}

    # ---- batched path: the whole batch (nq range + nq KNN) per step
    batched = None
    roofline = plain_roofline
    if not args.no_batch:
        h.batch_init(args.nq, args.pair_cap, args.hit_cap)
        bK = Ks.astype(np.uint64)
        # launch parameters from the autotuner on this batch (untimed set-up)
        tuned = None if args.no_tune else h.batch_autotune(boxes, knn_q, bK, reps=3)
        tuned_params = {name: h.batch_get_param(pid) for name, pid in PARAMS}
        log("autotune", tuned, tuned_params)
        # a second engine (batch slot 1, its own stream) with the same launch
        # parameters holds the KNN batch, slot 0 the range batch
        h.batch_slot(1)
        h.batch_init(args.nq, args.pair_cap, args.hit_cap)
        for name, pid in PARAMS:
            h.batch_set_param(pid, tuned_params[name])
        h.batch_set("knn", knn_q, bK)
        h.batch_slot(0)
        h.batch_set("range", boxes)
        # timed steps: both batches of a step replayed as their captured
        # hipGraphs, concurrently on the two streams (no per-kernel events)
        h.batch_use_graph(True)
        timing.enable(False)
        t_s = []
        tot_hits = tot_nn = 0
        for it in range(args.warmup + args.steps):
            t0 = time.perf_counter()
            tot_hits, tot_nn = h.batch_run_slots((0, 1))
            t1 = time.perf_counter()
            if it >= args.warmup:
                t_s.append(t1 - t0)
            log("batched step %d: %.2f ms (%d hits, %d neighbours)" % (it, (t1 - t0) * 1e3, tot_hits, tot_nn))
        # the same graphs one kind at a time (per-kind times)
        t_r, t_k = [], []
        for it in range(1 + args.steps):
            h.batch_slot(0)
            t0 = time.perf_counter()
            hr = h.batch_run()
            t1 = time.perf_counter()
            h.batch_slot(1)
            t2 = time.perf_counter()
            hk = h.batch_run()
            t3 = time.perf_counter()
            if hr != tot_hits or hk != tot_nn:
                raise RuntimeError("sequential and concurrent runs disagree")
            if it >= 1:
                t_r.append(t1 - t0)
                t_k.append(t3 - t2)
        h.batch_slot(0)
        # the same batches with plain launches, every kernel bracketed by HIP
        # events on the library's stream: the per-kernel times and the roofline
        h.batch_use_graph(False)
        timing.enable(True)
        tp_r, tp_k = [], []
        kinds = (("range_prune", 8), ("range_final", 9), ("group_scan", 10), ("group_scatter", 11),
                 ("radix_sort", 12), ("knn_levels", 13), ("knn_last_passes", 17), ("knn_select", 18),
                 ("knn_refine", 15), ("knn_merge", 16), ("range_rows", 20), ("knn_rows", 21))
        kacc = {n: [0, 0.0] for n, _ in kinds}
        work_r = work_k = (0, 0)
        leaf_tests = 0
        for it in range(1 + args.steps):
            h.batch_set("range", boxes)
            timing.reset()
            t0 = time.perf_counter()
            hits_p = h.batch_run()
            t1 = time.perf_counter()
            work_r = h.batch_work()
            leaf_tests = h.batch_leaf_tests()
            h.batch_set("knn", knn_q, bK)
            t2 = time.perf_counter()
            nn_p = h.batch_run()
            t3 = time.perf_counter()
            work_k = h.batch_work()
            if hits_p != tot_hits or nn_p != tot_nn:
                raise RuntimeError("plain launches and the captured graphs disagree")
            if it >= 1:
                tp_r.append(t1 - t0)
                tp_k.append(t3 - t2)
                for n, k in kinds:
                    c, ms = timing.get(k)
                    kacc[n][0] += c
                    kacc[n][1] += ms
        h.batch_use_graph(True)
        step_s = float(np.mean(t_s))
        per_step = {n: round(v[1] / args.steps, 3) for n, v in kacc.items()}
        batched = {"qps": round(2 * args.nq / step_s, 3), "step_ms": round(step_s * 1e3, 3),
                   "range_ms": round(float(np.mean(t_r)) * 1e3, 3),
                   "knn_ms": round(float(np.mean(t_k)) * 1e3, 3), "hits": int(tot_hits), "neighbours": int(tot_nn),
                   "timed_path": "the range and KNN batches of a step replayed as captured hipGraphs, concurrently "
                                 "on two engines (batch slots 0 and 1, one stream each)",
                   "sequential_step_ms": round((float(np.mean(t_r)) + float(np.mean(t_k))) * 1e3, 3),
                   "range_plain_ms": round(float(np.mean(tp_r)) * 1e3, 3),
                   "knn_plain_ms": round(float(np.mean(tp_k)) * 1e3, 3),
                   "kernels_ms_per_step": per_step,
                   "final_filter_tests": work_r[0], "final_filter_rows": work_r[1],
                   "final_filter_leaf_tests": leaf_tests,
                   "knn_refine_evals": work_k[0], "knn_refine_rows": work_k[1],
                   "speedup_vs_plain": round(2 * args.nq / step_s / plain_qps, 2),
                   "autotune": None if tuned is None else {
                       "range_ms_before_after": [round(float(x), 3) for x in tuned["range_ms"]],
                       "knn_ms_before_after": [round(float(x), 3) for x in tuned["knn_ms"]],
                       "params": tuned_params}}
        log("batched", batched)
        # rooflines of the two largest batched kernels, per launch (one launch
        # per range batch, averaged over the timed plain-launch steps):
        #  - the exact box filter: bounded by VALU compare issue (each compare
        #    is one VALU lane-operation; 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz
        #    = 39.3 T lane-ops/s for FP64 and FP32 alike), with its algorithmic
        #    HBM bytes (candidate rows read once per work item, hits written)
        #    beside it;
        #  - the hit sort: HBM-bound, algorithmic bytes = a key read for the
        #    counts plus a key read and write per pass.
        # The dominant one (largest time per step) is the bench line's roofline.
        D = cfg["D"]
        rl = {}
        # the exact filter is one query-major launch per batch plus, when clusters kept by few
        # queries go row-major, one row-major launch: both are the filter's time per batch
        n_f, ms_f = kacc["range_final"]
        ms_f += kacc["range_rows"][1]
        if n_f:
            avg_s = ms_f / 1e3 / n_f
            ops = 2.0 * D * (work_r[0] + leaf_tests)
            ach = ops / avg_s / 1e12
            alg = work_r[1] * (8 * D + 8) + 8 * tot_hits
            rl["range_final"] = {
                "bound": "fp64-valu", "achieved": round(ach, 3), "peak": VALU_PEAK_TOPS, "unit": "TOP/s",
                "frac": round(ach / VALU_PEAK_TOPS, 4), "traffic": final_traffic(cfg),
                "kernel": "k_range_final<%d> + k_range_rows<%d> (exact box filter of the batched range path, kd "
                          "leaves; query-major, and row-major for clusters kept by few queries)" % (D, D),
                "launches": n_f, "avg_launch_us": round(avg_s * 1e6, 3), "ops_per_launch": ops,
                "row_tests": int(work_r[0]), "leaf_tests": int(leaf_tests), "rows": int(work_r[1]),
                "hits": int(tot_hits), "algorithmic_bytes": int(alg),
                "hbm_frac_algorithmic": round(alg / avg_s / 1e9 / HBM_PEAK_GBS, 4),
                "note": "compare lane-operations (2D per row x box test in FP64, 2D per leaf x box test in FP32, "
                        "v_cmpx chains) against the VALU issue rate; HBM is not its bound"}
        n_s, ms_s = kacc["radix_sort"]
        if n_s:
            avg_s = ms_s / 1e3 / n_s
            id_bits = max(1, int(cfg["N"] - 1).bit_length())
            q_bits = max(1, int(args.nq - 1).bit_length())
            nbits = id_bits + q_bits
            rb = 9 if (nbits + 8) // 9 < (nbits + 7) // 8 else 8
            passes = (nbits + rb - 1) // rb
            alg = tot_hits * (8 + 16 * passes)
            ach = alg / avg_s / 1e9
            rl["radix_sort"] = {
                "bound": "hbm", "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": sort_traffic(cfg, passes),
                "kernel": "hit sort (k_rs_counts + %d x k_rs_onesweep<%d>, %d-bit keys)" % (passes, rb, nbits),
                "launches": n_s, "avg_launch_us": round(avg_s * 1e6, 3), "bytes_per_launch": int(alg),
                "keys": int(tot_hits)}
        if rl:
            dom = max(rl, key=lambda k: rl[k]["avg_launch_us"])
            roofline = dict(rl[dom])
            roofline["others"] = {k: v for k, v in rl.items() if k != dom}
            roofline["plain"] = plain_roofline

    # serving first after the batched step: it measures latency, so it runs before the legs that
    # leave large host and device state behind.  Slot 1's engine goes first: its stream would
    # share one of the process's four hardware queues with a server worker's stream.
    if batched:
        h.batch_slot(1)
        h.batch_free()
        h.batch_slot(0)
    serving = None
    if args.serve_rates and not args.no_batch:
        rates = [float(x) for x in args.serve_rates.split(",") if x]
        serving = {"workload": "open-loop Poisson stream on the same index: 70% range (selectivity log-uniform "
                               "1e-6..1e-3), 30% KNN (K log-uniform 1..256, K above the batched kernels' limit per query); continuous batching, admission by "
                               "free HBM; latency = submit to result on the host",
                   "runs": run_serving(h, *serving_mix(data, cfg["seed"], h.mesh_sizes, cfg["vmin"], cfg["vmax"]),
                                       rates, args.serve_seconds, log)}

    learned = None
    if args.learn_traces and not args.no_batch:
        learned = run_learned(h, data, cfg, knn_q, Ks.astype(np.uint64), args.learn_traces, args.learn_epochs,
                              min(args.plain_sample, 96), log)

    c3 = None
    if args.c3_queries and not args.no_batch:
        c3 = run_c3(args.c3_queries, args.c3_traces, args.learn_epochs, min(args.plain_sample, 48), log)

    build = None
    if args.build_n:
        build = run_build(args.build_n, args.build_d, (100, 50, 20), args.build_cpu_n, log)

    c4 = None
    if args.c4_n:
        c4 = run_c4(args.c4_n, 32, (100, 50, 20), args.c4_iters, 7, args.c4_sub, log)

    c5 = None
    if args.c5_n and not args.no_batch:
        c5 = run_c5(args.c5_n, c5_rates, args.c5_seconds, 64, log)

    cpu = None
    if not args.no_cpu:
        import oracle_py as O

        log("CPU baseline: generating oracle index")
        o = O.Index.generate(cfg["N"], cfg["D"], cfg["vmin"], cfg["vmax"], cfg["is_int"], cfg["ratios"], cfg["seed"],
                             cfg["sigma_div"])
        nr, nk = args.cpu_range, args.cpu_knn
        w1, us1, _, th = o.cpu_batch(0, boxes[:nr])
        w2, us2, _, _ = o.cpu_batch(1, knn_q[:nk], K=10)
        w3, us3, _, _ = o.cpu_batch(1, knn_q[args.nq // 2: args.nq // 2 + nk], K=100)
        tot_q = nr + 2 * nk
        cpu = {"value": round(tot_q / (w1 + w2 + w3), 3), "unit": "queries/s", "cores": th, "kind": "port",
               "sample": "%d range + %d KNN(K=10) + %d KNN(K=100) queries of the same batch, OpenMP over queries, "
                         "traversal + exact refine" % (nr, nk, nk),
               "p50_us": round(float(np.percentile(np.concatenate([us1, us2, us3]), 50)), 1),
               "p99_us": round(float(np.percentile(np.concatenate([us1, us2, us3]), 99)), 1)}
        log("cpu baseline", cpu)

    value = batched["qps"] if batched else plain_qps
    ms_per_step = batched["step_ms"] if batched else wall / args.steps * 1e3
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "hierarchy N=%d D=%d ratios %s sigma/%.1f seed %d; %d range boxes (selectivity "
                               "log-uniform 1e-5..1e-2) + %d KNN points (K 10|100); exact answers"
                               % (cfg["N"], cfg["D"], ",".join(map(str, cfg["ratios"])), cfg["sigma_div"],
                                  cfg["seed"], args.nq, args.nq),
                   "path": ("batched (all %d queries per step)" % (2 * args.nq)) if batched else
                   ("plain per-query (sample of %d queries per step)" % nq), "N": cfg["N"], "D": cfg["D"]},
        "plain": {"qps": round(plain_qps, 3), "p50_us": round(float(np.percentile(us, 50)), 1),
                  "p99_us": round(float(np.percentile(us, 99)), 1), "queries": nq},
        "batched": batched,
        "learned": learned,
        "learned_d16": c3,
        "serving": serving,
        "serving_c5": c5,
        "build_c4": c4,
        "build": build,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out, default=_json_default), flush=True)
    if args.calib_copy_mb:
        # device-to-device hipMemcpy: the runtime's copy kernel, n bytes read and n written
        import ctypes

        hip = ctypes.CDLL("libamdhip64.so")
        n = args.calib_copy_mb * (1 << 20)
        src, dst = ctypes.c_void_p(), ctypes.c_void_p()
        if hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(n)) or hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(n)):
            raise RuntimeError("hipMalloc failed for the calibration copy")
        hip.hipMemset(src, 1, ctypes.c_size_t(n))
        hip.hipDeviceSynchronize()
        if hip.hipMemcpy(dst, src, ctypes.c_size_t(n), 3):
            raise RuntimeError("hipMemcpy failed for the calibration copy")
        hip.hipDeviceSynchronize()
        hip.hipFree(src)
        hip.hipFree(dst)
        log("calibration copy: %d bytes read, %d bytes written" % (n, n))
    return 0


if __name__ == "__main__":
    sys.exit(main())
