"""Benchmark: exact kNN queries/s on one MI355X, mixed (bf16 MFMA + exact rerank
+ certificate) beside the plain FP32 MFMA form.

Workload (default): database 1,000,000 x 128 f32 uniform [0,1), 100,000 queries
from the same distribution, k = 10, Euclidean.  One step = one kNN pass of all
queries against the resident database (inputs already in HBM).  Prints one JSON
line with the throughput of --mode, the same for the FP32 form and their ratio,
the dominant kernel's roofline fraction (timed live with HIP events on its
launch stream; HBM traffic from the committed rocprofv3 summary) and the CPU
baseline (the oracle's restatement of fuel's AVX2 linear scan on the host
cores, on a bounded sample).  Every mode returns fuel's exact answer; a sample
of queries is checked against the oracle bit for bit.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode bf16|fp32|exact]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAKS = {  # MI355X published dense peaks
    "bf16": (2516.6e12, "TFLOP/s", "mfma"),
    "bf16r": (2516.6e12, "TFLOP/s", "mfma"),
    "fp32": (157.3e12, "TFLOP/s", "mfma"),
    "exact": (157.3e12, "TFLOP/s", "mfma"),
    "hbm": (8.0e12, "GB/s", "hbm"),
}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--mode", default="bf16", choices=["bf16", "bf16r", "fp32", "exact"])
    p.add_argument("--n", type=int, default=1_000_000)
    p.add_argument("--m", type=int, default=100_000)
    p.add_argument("--d", type=int, default=128)
    p.add_argument("--k", type=int, default=10)
    p.add_argument("--widen", type=int, default=4, help="candidates kept per query = widen * k (<= 128)")
    p.add_argument("--seed", type=int, default=1, help="seeded first-pass thresholds of the plain MFMA scans (0: off)")
    p.add_argument("--verify", type=int, default=64, help="random queries checked against the oracle")
    p.add_argument("--verify-per-block", type=int, default=6,
                   help="queries checked from every block of 256 queries (beside every query whose first "
                        "certificate failed)")
    p.add_argument("--c3-full", type=int, default=1,
                   help="headline line: GPU LOF of the full 70,000 x 784 workload against the oracle (0: skip)")
    p.add_argument("--c4-line", type=int, default=1,
                   help="headline line: the 512 batched problems (plan, passes, parity on the smallest third and on the largest problem of at most 25,000 rows of each dimension)")
    p.add_argument("--c4-steps", type=int, default=3, help="timed passes of the batched problems")
    p.add_argument("--tune-cache", default=None,
                   help="file keeping the batch autotuner's choices across runs (read, then appended to)")
    p.add_argument("--calibrate", type=int, default=1,
                   help="headline line: achievable bf16 rate of a large torch matmul beside the nominal peak")
    p.add_argument("--cpu-sample", type=int, default=1000, help="queries timed on the CPU baseline (0: skip)")
    p.add_argument("--fp32-steps", type=int, default=2,
                   help="timed steps of the plain FP32 MFMA form run beside the mixed one (0: skip)")
    p.add_argument("--workload", default="c2", choices=["c2", "c1", "c3", "c4", "c5", "c5scan", "dbscan", "kmeans"],
                   help="c2: the headline kNN line; the others print their own line (not the headline): "
                        "c1 self-kNN 10k x 16 + pairwise 2000^2, c3 LOF 70k x 784 k=20, c4 512 uneven self-kNN "
                        "problems (batched), c5 serving 10M x 96 "
                        "under Poisson arrivals (c5scan: only 20 serving-size launches of 64 queries, for "
                        "profiling the serving scan), dbscan on c1, kmeans distance matrix 1M x 64 vs 100 centres")
    p.add_argument("--batch-mode", default="auto", choices=["auto", "bf16", "bf16r", "fp32", "exact"],
                   help="configuration of the batched problems (c4): autotuned or fixed")
    p.add_argument("--serve-mode", default="bf16r", choices=["bf16", "bf16r", "fp8r", "fp32", "exact"],
                   help="candidate mode of the served index (c5)")
    p.add_argument("--serve-seconds", type=float, default=4.0, help="seconds of arrivals per rate (c5)")
    p.add_argument("--serve-fp8-seconds", type=float, default=2.0,
                   help="seconds of serving the FP8 E4M3 copy beside the serving mode in the headline line (0: skip)")
    p.add_argument("--serve-p99-seconds", type=float, default=3.0,
                   help="seconds of serving at 10,000 queries/s measured for the headline line's p99 (0: skip)")
    p.add_argument("--serve-rows", type=int, default=10_000_000, help="rows of the served index (c5)")
    p.add_argument("--serve-workers", type=int, default=2, help="serving worker threads (c5 and the headline p99)")
    p.add_argument("--serve-verify", type=int, default=240,
                   help="served queries of the headline serving run checked against the oracle")
    p.add_argument("--host-path", action="store_true",
                   help="only time the host-pointer entry point (queries and results cross PCIe each call)")
    p.add_argument("--copy-probe", type=int, default=0,
                   help="only run a plain device copy of this many bytes 3 times (calibrates HBM counters)")
    return p.parse_args()


def profiled_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload`, from the rocprofv3
    FETCH_SIZE / WRITE_SIZE summary committed under profiles/ (calibrated by
    the copy probe below), or None when no summary covers this pair."""
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None
    for name in sorted(os.listdir(pdir)):
        if not name.endswith((".md", ".txt")):
            continue
        fields = {}
        with open(os.path.join(pdir, name)) as f:
            for line in f:
                line = line.strip()
                if line.startswith("- ") and ": " in line:
                    key, val = line[2:].split(": ", 1)
                    fields.setdefault(key.strip(), val.strip())
        if fields.get("kernel") == kernel and fields.get("workload") == workload:
            try:
                return int(fields["traffic_bytes_per_launch"])
            except (KeyError, ValueError):
                return None
    return None


def copy_probe(nbytes):
    """A plain device-to-device copy of known size: its FETCH_SIZE / WRITE_SIZE
    readings per dispatch convert the counters of the kNN kernels into bytes."""
    import torch
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes}), flush=True)


METRIC = "kNN queries/s, 1M\u00d7128 k=10: mixed vs FP32 at exact recall; p99 serve ms"


def run_mode(_lib, Xd, Qd, n, m, d, k, mode, widen, steps, warmup, stream, out_i, out_d, seed=True):
    """Warm up, then time `steps` kNN passes of all queries on a resident index.
    Returns wall time per step, the dominant scan kernel's HIP-event time and the
    index statistics of the last pass."""
    import torch
    ix = _lib.Index.from_device(Xd.data_ptr(), n, d, d, np.float32, "euclidean", mode=mode, widen=widen,
                                seed=seed)

    def step():
        ix.knn_device(Qd.data_ptr(), m, d, k, _lib.EXCLUDE_NONE, out_i.data_ptr(), out_d.data_ptr(), stream)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ix.timing_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kernel_ms, launches, kname = ix.timing_read()
    ix.timing_enable(False)
    res = {"ms_per_step": (t1 - t0) / steps * 1e3, "qps": m * steps / (t1 - t0),
           "kernel_s": (kernel_ms / max(launches, 1)) / 1e3, "kernel": kname, "stats": ix.stats(),
           "fail_rows": ix.fallback_rows()}
    ix.close()
    return res


def stratified_sample(m, per_block, block=256, seed=0):
    """per_block random queries from every block of `block` consecutive query
    slots (the scan's query blocks), so that every workgroup's queries are seen."""
    rng = np.random.default_rng(seed)
    out = []
    for b0 in range(0, m, block):
        b1 = min(m, b0 + block)
        out.append(rng.choice(np.arange(b0, b1), size=min(per_block, b1 - b0), replace=False))
    return np.sort(np.concatenate(out)) if out else np.zeros(0, np.int64)


def verify(oracle, X, Q, gi, gd, k, nverify, extra=None, per_block=0):
    """Queries checked against the oracle: identical indices and f32 bits,
    recall, and the largest relative error against an f64 oracle.  The set is
    a random sample of nverify queries, per_block queries of every 256-query
    block, and `extra` (e.g. every query whose first certificate failed)."""
    m = Q.shape[0]
    rng = np.random.default_rng(0)
    parts = [rng.choice(m, size=min(nverify, m), replace=False)]
    if per_block:
        parts.append(stratified_sample(m, per_block))
    if extra is not None and len(extra):
        parts.append(np.asarray(extra, np.int64))
    sample = np.unique(np.concatenate(parts))
    oi, od = oracle.knn(X, Q[sample], k)
    exact_rows = int(np.sum(np.all(gi[sample] == oi, axis=1) & np.all(gd[sample] == od, axis=1)))
    recall = float(np.mean([len(set(gi[s]) & set(oi[j])) / k for j, s in enumerate(sample)]))
    qi = np.repeat(sample, k)
    xi = gi[sample].ravel()
    ref64 = np.sqrt(oracle.sqdist_fp64(X, Q, qi, xi))
    max_rel_err = float(np.max(np.abs(gd[sample].ravel().astype(np.float64) - ref64) / np.maximum(ref64, 1e-300)))
    out = {"queries": int(len(sample)), "exact_rows": exact_rows, "recall": recall,
           "max_rel_err_vs_fp64": max_rel_err}
    if extra is not None:
        ex = np.asarray(extra, np.int64)
        pos = np.searchsorted(sample, ex)
        ok = np.all(gi[ex] == oi[pos], axis=1) & np.all(gd[ex] == od[pos], axis=1) if len(ex) else np.zeros(0, bool)
        out["second_pass_queries"] = int(len(ex))
        out["second_pass_exact_rows"] = int(ok.sum())
    if per_block:
        out["query_blocks_covered"] = int(len(np.unique(sample // 256)))
    return out


VALU_F32_PEAK = 256 * 4 * 16 * 2 * 2.4e9  # FMA lanes/s x 2 flop, no packed FP32 (78.6 TFLOP/s)


def timed(fn, reps):
    """Median wall time of `reps` calls (each synchronous)."""
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def serve_run(srv, pool, lam, seconds, rng, k=10, mean_batch=64, cap=4096, keep=0, keep_queries=0):
    """Open-loop arrivals: requests arrive as a Poisson process of rate
    lam / mean_batch per second with geometric sizes (mean mean_batch, capped),
    so queries arrive at lam per second on average.  Returns the server's
    latency statistics over the run and the first answers: `keep` requests,
    or as many as hold `keep_queries` queries."""
    import queue
    import threading
    srv.stats(reset=True)
    tickets = queue.Queue()
    kept = []

    def waiter():
        while True:
            item = tickets.get()
            if item is None:
                return
            t, off, m = item
            idx, dist, _, _ = srv.wait(t)
            if len(kept) < keep or sum(x[1] for x in kept) < keep_queries:
                kept.append((off, m, idx, dist))

    th = threading.Thread(target=waiter)
    th.start()
    t_start = time.perf_counter()
    t_next, off, nq, nreq = t_start, 0, 0, 0
    while True:
        t_next += rng.exponential(mean_batch / lam)
        if t_next - t_start > seconds:
            break
        m = int(min(cap, rng.geometric(1.0 / mean_batch)))
        now = time.perf_counter()
        if t_next > now:
            time.sleep(t_next - now)
        if off + m > pool.shape[0]:
            off = 0
        tickets.put((srv.submit(pool[off:off + m], k), off, m))
        off += m
        nq += m
        nreq += 1
    tickets.put(None)
    th.join()
    elapsed = time.perf_counter() - t_start
    st = srv.stats(reset=True)
    return {"offered_qps": lam, "achieved_qps": round(st["queries"] / elapsed, 1), "requests": nreq,
            "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3), "max_ms": round(st["max_ms"], 3),
            "launches": st["batches"], "mean_launch_queries": round(st["queries"] / max(1, st["batches"]), 1),
            "max_launch_queries": st["max_batch_queries"], "gpu_busy_frac": round(st["busy_s"] / elapsed, 3)}, kept


def c5_data(args):
    """The serving index rows and a pool of queries from the same mixture."""
    from fuel_amd import workloads
    X, means = workloads.c5(n=args.serve_rows)
    rng = np.random.default_rng(66)
    P = 65536
    pool = (means[rng.integers(0, means.shape[0], P)] + rng.normal(0.0, 1.0, (P, X.shape[1]))).astype(np.float32)
    return X, pool, rng


def serve_scan_roofline(ix, pool, X, mode, reps=20, m=64):
    """The serving scan alone: `reps` launches of m queries (one serving batch)
    through the synchronous entry point, the level-0 scan timed by HIP events
    on its stream; HBM roofline against the algorithmic bytes of one pass over
    the resident copy, N d s + N 8 + Q d 4 + Q k 12 (s = 1 byte for the E4M3
    copy, 2 for bf16; 8 bytes per row of norm and scale terms)."""
    n, d = X.shape
    s_db = {"fp8r": 1, "bf16r": 2, "bf16": 2, "fp32": 4}.get(mode, 4)
    q = np.ascontiguousarray(pool[:m])
    ix.knn(q, 10)  # warm (workspaces)
    ix.timing_enable(True)
    for r in range(reps):
        ix.knn(np.ascontiguousarray(pool[(r + 1) * m:(r + 2) * m]), 10)
    kernel_ms, launches, kname = ix.timing_read()
    ix.timing_enable(False)
    t = kernel_ms / max(launches, 1) / 1e3
    alg = n * d * s_db + n * 8 + m * d * 4 + m * 10 * 12
    return {"kernel": kname, "launch_queries": m, "kernel_ms": round(t * 1e3, 4), "launches": launches,
            "algorithmic_bytes": alg, "achieved_GBps": round(alg / t / 1e9, 1),
            "roofline": {"bound": "hbm", "achieved": round(alg / t / 1e9, 1), "peak": 8000.0, "unit": "GB/s",
                         "frac": round(alg / t / 8.0e12, 4)}}


def serve_headline(args):
    """p50/p99 request latency of the serving loop on the serving configuration
    at one arrival rate (10,000 queries/s), for the headline line, in the serving
    mode and (--serve-fp8-seconds) with the FP8 E4M3 copy; every row of the
    first requests of each run is checked against the oracle."""
    from fuel_amd import Server, _lib
    import oracle
    X, pool, rng = c5_data(args)
    out = {}
    runs = [(args.serve_mode, args.serve_p99_seconds)]
    if args.serve_fp8_seconds > 0 and args.serve_mode != "fp8r":
        runs.append(("fp8r", args.serve_fp8_seconds))
    for mode, seconds in runs:
        ix = _lib.Index(X, "euclidean", mode=mode)
        with Server(ix, workers=args.serve_workers, max_batch=4096) as srv:
            serve_run(srv, pool, 2000, 0.5, rng)  # warm-up
            r, kept = serve_run(srv, pool, 10000, seconds, rng, keep_queries=args.serve_verify)
        scan = serve_scan_roofline(ix, pool, X, mode)
        ix.close()
        r = dict(r)
        r["scan"] = scan
        t0 = time.perf_counter()
        same, nchk, nreq = True, 0, 0
        for off, m, idx, dist in kept:
            oi, od = oracle.knn(X, pool[off:off + m], 10, "euclidean", "none", threads=oracle.default_threads())
            same = same and bool(np.array_equal(oi, idx) and od.tobytes() == dist.tobytes())
            nchk += m
            nreq += 1
        print(f"# serve {mode}: {nchk} served queries of {nreq} requests checked against the oracle in "
              f"{time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
        r.update({"workload": f"serving: resident index {X.shape[0]:,} x {X.shape[1]} f32 (50-component mixture), "
                              "k=10, Poisson arrivals, geometric batch sizes (mean 64, cap 4096)",
                  "mode": mode, "seconds": seconds,
                  "verify": {"requests": nreq, "queries": nchk, "identical": same}})
        out[mode] = r
    main = out[args.serve_mode]
    if "fp8r" in out and args.serve_mode != "fp8r":
        main["fp8r"] = out["fp8r"]
    return main


def c3_full(threads):
    """LOF of the full 70,000 x 784 workload (k=20) on the GPU against the
    oracle's restatement of fuel's for_each_knn + LOF, bit for bit."""
    from fuel_amd import _lib, workloads
    import oracle
    t0 = time.time()
    X = workloads.c3()
    ix = _lib.Index(X, "euclidean")
    g = ix.lof(20)
    t_g = timed(lambda: ix.lof(20), 2)
    st = ix.stats()
    ix.close()
    print(f"# c3: GPU LOF in {t_g * 1e3:.1f} ms; oracle on the full problem ...", file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    o = oracle.lof(X, 20, threads=threads)
    t_c = time.perf_counter() - t1
    print(f"# c3: oracle LOF in {t_c:.1f}s ({time.time() - t0:.0f}s with data)", file=sys.stderr, flush=True)
    return {"workload": "LOF n=70000 d=784 integers 0..255 (80% zero) k=20", "gpu_ms": round(t_g * 1e3, 3),
            "knn_stats": {k: st[k] for k in ("mode_used", "fallback_queries", "exact_queries")},
            "identical": bool(g.tobytes() == o.tobytes()),
            "cpu_baseline": {"value": round(t_c * 1e3, 1), "unit": "ms", "cores": threads, "kind": "port",
                             "sample": "the full problem (oracle for_each_knn + LOF)"}}


def c4_run(args, threads, steps, warmup):
    """The 512 batched uneven self-kNN problems: plan (autotuned), graph
    replays timed, answers of the smallest third and of each dimension's largest
    problem of at most 25,000 rows checked against the oracle, and a hash of
    every problem's answer."""
    import hashlib

    import torch

    from fuel_amd import BatchKnn, workloads
    import oracle
    t0 = time.time()
    probs = workloads.c4()
    print(f"# c4: {len(probs)} problems generated in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    t0 = time.time()
    b = BatchKnn(probs, 15, mode=args.batch_mode, tune_cache=args.tune_cache)
    tune_s = time.time() - t0
    print(f"# c4: planned in {tune_s:.1f}s", file=sys.stderr, flush=True)
    for _ in range(warmup):
        b.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.run()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    print(f"# c4: {ms:.1f} ms per pass", file=sys.stderr, flush=True)
    info = b.describe()
    flop = float(sum(2.0 * X.shape[0] ** 2 * X.shape[1] for X in probs))
    achieved = flop / (ms * 1e-3)
    res = b.results()
    h = hashlib.sha256()
    for gi, gd in res:
        h.update(gi.tobytes())
        h.update(gd.tobytes())
    order = sorted(range(len(probs)), key=lambda i: probs[i].shape[0])
    checked = order[:len(probs) // 3]  # the smallest third, in full
    # and per dimension the largest problem of at most 25,000 rows (seeded
    # thresholds start at about 8,000 rows; the oracle takes seconds for these)
    for dd in sorted({X.shape[1] for X in probs}):
        cand = [i for i in order if probs[i].shape[1] == dd and probs[i].shape[0] <= 25000]
        if cand and cand[-1] not in checked:
            checked.append(cand[-1])
    t0 = time.perf_counter()
    same = True
    for i in checked:
        oi, od = oracle.knn(probs[i], probs[i], 15, "euclidean", "for_each_knn", threads=threads)
        same = same and bool(np.array_equal(oi, res[i][0]) and od.tobytes() == res[i][1].tobytes())
    print(f"# c4: {len(checked)} problems checked against the oracle in {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)
    out = {"workload": "512 self-kNN problems, n log-uniform in [1e3, 2e5], d in {8,32,64,256}, 3-20 Gaussian "
                       "components (means U[-5,5]^d), k=15, for_each_knn exclusion, autotuned, hipGraph replay",
           "rows": int(sum(X.shape[0] for X in probs)), "algorithmic_flop": flop, "ms_per_pass": round(ms, 2),
           "steps": steps, "warmup": warmup, "plan_seconds": round(tune_s, 1),
           "plan": [{k: g[k] for k in ("d", "mode", "widen", "fill") if k in g} for g in info["groups"]],
           "fallback_queries": info["last_fallback_queries"], "exact_queries": info["last_exact_queries"],
           "last_pass_split_ms": {"level0": info["last_level0_ms"],
                                  "chain": round(info["last_ms"] - info["last_level0_ms"], 3)},
           "roofline": {"bound": "mfma", "achieved": round(achieved / 1e12, 2), "peak": PEAKS["bf16"][0] / 1e12,
                        "unit": "TFLOP/s", "frac": round(achieved / PEAKS["bf16"][0], 4),
                        "note": "2 n^2 d per problem over the whole pass; bf16 dense peak"},
           "answers_sha256": h.hexdigest(), "checked_problems": len(checked),
           "checked_max_rows": int(max(probs[i].shape[0] for i in checked)), "identical": same}
    b.close()
    return out, probs, info


def calibration():
    """Achievable bf16 matrix-core rate on this box: the fastest of a few large
    torch bf16 matmuls (hipBLASLt), beside the nominal dense peak."""
    import torch
    best, shapes = 0.0, {}
    for (m, n, k) in ((8192, 8192, 8192), (16384, 16384, 8192), (8192, 8192, 16384)):
        a = torch.randn((m, k), dtype=torch.bfloat16, device="cuda")
        b = torch.randn((k, n), dtype=torch.bfloat16, device="cuda")
        for _ in range(3):
            torch.matmul(a, b)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 10
        e0.record()
        for _ in range(reps):
            torch.matmul(a, b)
        e1.record()
        torch.cuda.synchronize()
        rate = 2.0 * m * n * k / (e0.elapsed_time(e1) / reps / 1e3)
        shapes[f"{m}x{n}x{k}"] = round(rate / 1e12, 1)
        best = max(best, rate)
        del a, b
    torch.cuda.empty_cache()
    return {"bf16_gemm_tflops": round(best / 1e12, 1), "shapes_tflops": shapes,
            "shape": "fastest of the torch.matmul bf16 shapes", "frac_of_nominal": round(best / PEAKS["bf16"][0], 4)}


def secondary(args):
    """Lines for the other configurations and the widened rows: GPU time beside
    the oracle's CPU time on the same inputs, parity checked on the way."""
    from fuel_amd import _lib, workloads
    import fuel_amd
    import oracle
    threads = oracle.default_threads()
    line = {"workload": args.workload, "n_gpus": 1, "data": "synthetic", "cpu_threads": threads}
    if args.workload == "c1":
        X = workloads.c1()
        ix = _lib.Index(X, "euclidean")
        knn = lambda: ix.knn(X, 10, _lib.EXCLUDE_PY_LINEAR)  # noqa: E731
        knn()
        t_g = timed(knn, 5)
        gi, gd = knn()
        t0 = time.perf_counter()
        oi, od = oracle.knn(X, X, 10, "euclidean", "python_linear", threads=threads)
        t_c = time.perf_counter() - t0
        P = X[:2000]
        pw = lambda: fuel_amd.pairwise_distances(P, "sqeuclidean", form="square")  # noqa: E731
        pw()
        t_pg = timed(pw, 5)
        t0 = time.perf_counter()
        po = oracle.pairwise(P, "sqeuclidean", "square", threads=threads)
        t_pc = time.perf_counter() - t0
        line.update({"config": "self-kNN n=10000 d=16 k=10 exclude_self (host arrays) + pairwise 2000^2 sqeuclidean",
                     "knn_gpu_ms": round(t_g * 1e3, 3), "knn_cpu_ms": round(t_c * 1e3, 3),
                     "knn_identical": bool(np.array_equal(gi, oi) and gd.tobytes() == od.tobytes()),
                     "pairwise_gpu_ms": round(t_pg * 1e3, 3), "pairwise_cpu_ms": round(t_pc * 1e3, 3),
                     "pairwise_identical": bool(pw().tobytes() == po.tobytes())})
    elif args.workload == "c3":
        X = workloads.c3()
        ix = _lib.Index(X, "euclidean")
        lof = lambda: ix.lof(20)  # noqa: E731
        lof()
        t_g = timed(lof, 3)
        st = ix.stats()
        S = 500
        t0 = time.perf_counter()
        oracle.knn(X, X[:S], 21, "euclidean", "for_each_knn", threads=threads)
        t_s = time.perf_counter() - t0
        line.update({"config": "LOF n=70000 d=784 integers 0..255 (80% zero) k=20 (self-kNN + LOF, host arrays)",
                     "lof_gpu_ms": round(t_g * 1e3, 3), "knn_stats": st,
                     "cpu_knn_ms_extrapolated": round(t_s / S * X.shape[0] * 1e3, 1),
                     "cpu_sample": f"{S} self-kNN queries of the oracle in {t_s:.2f}s, scaled to 70000",
                     "algorithmic_flop": 2.0 * X.shape[0] ** 2 * X.shape[1]})
        # parity on a sub-problem the oracle finishes in seconds
        Xs = X[:6000]
        g = _lib.Index(Xs, "euclidean").lof(20)
        line["lof_identical_6000_subset"] = bool(g.tobytes() == oracle.lof(Xs, 20, threads=threads).tobytes())
    elif args.workload == "c4":
        c4, probs, info = c4_run(args, threads, args.steps, args.warmup)
        # CPU baseline: the oracle's for_each_knn on two mid-size problems, rate in flop/s
        mid = sorted(range(len(probs)), key=lambda i: probs[i].shape[0])[len(probs) // 2:len(probs) // 2 + 2]
        t0 = time.perf_counter()
        f_s = 0.0
        for i in mid:
            oracle.knn(probs[i], probs[i], 15, "euclidean", "for_each_knn", threads=threads)
            f_s += 2.0 * probs[i].shape[0] ** 2 * probs[i].shape[1]
        t_c = time.perf_counter() - t0
        timed = [t for t in info["tuning"] if "sample" in t]  # cached groups carry no timing
        c4["tuning_best_ms"] = {str(dd): {m: min(t["ms"] for t in timed if t["d"] == dd and t["mode"] == m)
                                          for m in {t["mode"] for t in timed if t["d"] == dd}}
                                for dd in sorted({t["d"] for t in timed})}
        # the finalists, timed on the whole group
        c4["tuning_group_ms"] = [{k: t[k] for k in ("d", "mode", "widen", "fill", "ms", "fallback")}
                                 for t in info["tuning"] if "group" in t]
        c4["cpu_baseline"] = {"value": round(c4["algorithmic_flop"] / (f_s / t_c) * 1e3, 1),
                              "unit": "ms per pass (extrapolated)", "cores": threads, "kind": "port",
                              "sample": f"oracle for_each_knn on 2 mid-size problems ({t_c:.1f}s), scaled by 2 n^2 d"}
        line.update(c4)
    elif args.workload == "c5":
        from fuel_amd import Server
        t0 = time.time()
        X, pool, rng = c5_data(args)
        ix = _lib.Index(X, "euclidean", mode=args.serve_mode)
        print(f"# c5 data + index in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
        rates, kept = [], []
        with Server(ix, workers=args.serve_workers, max_batch=4096) as srv:
            serve_run(srv, pool, 2000, 0.5, rng)  # warm-up (workspaces, operand copy)
            for lam in (2000, 5000, 10000, 20000):
                r, kp = serve_run(srv, pool, lam, args.serve_seconds, rng, keep=2)
                rates.append(r)
                kept += kp
                print(f"# rate {lam}: {r}", file=sys.stderr, flush=True)
            # saturation: 64 requests of 64 queries submitted at once
            srv.stats(reset=True)
            t1 = time.perf_counter()
            ts = [srv.submit(pool[i * 64:(i + 1) * 64], 10) for i in range(64)]
            for t in ts:
                srv.wait(t)
            sat = 64 * 64 / (time.perf_counter() - t1)
            sst = srv.stats(reset=True)
        scan = serve_scan_roofline(ix, pool, X, args.serve_mode)
        print(f"# serving scan: {scan}", file=sys.stderr, flush=True)
        # exactness of served answers: a sample of rows against the oracle
        rows = [(off, min(m, 12), idx[:12], dist[:12]) for off, m, idx, dist in kept]
        t0 = time.perf_counter()
        same, nchk = True, 0
        for off, m, idx, dist in rows:
            oi, od = oracle.knn(X, pool[off:off + m], 10, "euclidean", "none", threads=threads)
            same = same and bool(np.array_equal(oi, idx) and od.tobytes() == dist.tobytes())
            nchk += m
        t_c = time.perf_counter() - t0
        line.update({"config": f"serving: resident index {X.shape[0]:,} x {X.shape[1]} f32 (50-component mixture), "
                               "k=10, Poisson arrivals, geometric batch sizes (mean 64, cap 4096)",
                     "mode": args.serve_mode, "workers": args.serve_workers, "rates": rates,
                     "saturated_qps": round(sat, 1), "saturated_p99_ms": round(sst["p99_ms"], 3), "scan": scan,
                     "identical_sample": same, "checked_queries": nchk,
                     "cpu_baseline": {"value": round(nchk / t_c, 2), "unit": "queries/s", "cores": threads,
                                      "kind": "port", "sample": f"{nchk} served queries re-answered by the oracle"}})
    elif args.workload == "c5scan":
        X, pool, rng = c5_data(args)
        ix = _lib.Index(X, "euclidean", mode=args.serve_mode)
        scan = serve_scan_roofline(ix, pool, X, args.serve_mode)
        st = ix.stats()
        ix.close()
        line.update({"config": f"serving scan: resident index {X.shape[0]:,} x {X.shape[1]} f32 (50-component "
                               "mixture), 20 launches of 64 queries, k=10", "mode": args.serve_mode, "scan": scan,
                     "kp": st["kp"], "splits": st["splits"], "fallback_queries_last": st["fallback_queries"]})
    elif args.workload == "dbscan":
        from fuel_amd.cluster import dbscan, optics
        X = workloads.c1()
        for variant in ("dbscan", "parallel"):
            f = lambda: dbscan(X, 4.0, 10, variant=variant)  # noqa: E731
            f()
            t_g = timed(f, 3)
            t0 = time.perf_counter()
            o = oracle.dbscan(X, 4.0, 10, variant=variant, threads=threads)
            t_c = time.perf_counter() - t0
            line[variant] = {"gpu_ms": round(t_g * 1e3, 3), "cpu_ms": round(t_c * 1e3, 3),
                             "identical": bool(np.array_equal(f(), o)), "clusters": int(o.max() + 1)}
        f = lambda: optics(X, 4.0, 10)  # noqa: E731
        f()
        t_g = timed(f, 3)
        t0 = time.perf_counter()
        o = oracle.optics(X, 4.0, 10)
        t_c = time.perf_counter() - t0
        line["optics"] = {"gpu_ms": round(t_g * 1e3, 3), "cpu_ms": round(t_c * 1e3, 3),
                          "identical": bool(np.array_equal(f().ordering(), o[0]))}
        line["config"] = "c1 data (10000 x 16, 6 Gaussians), eps=4.0, min_points=10; range on the GPU, walk on the host"
    elif args.workload == "kmeans":
        import ctypes

        import torch
        rng = np.random.default_rng(8)
        n, k, d = 1_000_000, 100, 64
        pts = rng.standard_normal((n, d)).astype(np.float32)
        cen = pts[rng.choice(n, k, replace=False)].copy()
        dev = torch.device("cuda:0")
        P = torch.from_numpy(pts).to(dev)
        C = torch.from_numpy(cen).to(dev)
        out = torch.empty((k, n), dtype=torch.float32, device=dev)
        st = torch.cuda.current_stream().cuda_stream
        lib = _lib.load()
        call = lambda: _lib.check(lib.fuel_amd_pairwise_sqdist_device(  # noqa: E731
            C.data_ptr(), k, d, P.data_ptr(), n, d, d, _lib.F32, out.data_ptr(), n, ctypes.c_void_p(st)))
        call()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(args.steps):
            call()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        terms = float(n) * k * d
        S = 20_000
        t0 = time.perf_counter()
        ref = oracle.kmeans_pairwise(cen, pts[:S], threads=threads)
        t_c = time.perf_counter() - t0
        same = bool(out[:, :S].cpu().numpy().tobytes() == ref.tobytes())
        line.update({"config": "k-means distance matrix: 100 centres x 1,000,000 points, d=64 (device buffers)",
                     "gpu_ms": round(ms, 3), "identical_first_20000_columns": same,
                     "roofline": {"bound": "valu", "achieved": round(2 * terms / (ms * 1e-3) / 1e12, 3),
                                  "peak": round(VALU_F32_PEAK / 1e12, 1), "unit": "TFLOP/s",
                                  "frac": round(2 * terms / (ms * 1e-3) / VALU_F32_PEAK, 4),
                                  "note": "one sub + one fma per term (2 flop counted), non-packed FP32 FMA peak"},
                     "output_bytes": 4 * n * k,
                     "cpu_ms_extrapolated": round(t_c * n / S * 1e3, 1),
                     "cpu_sample": f"oracle on {S} of the points, scaled"})
    print(json.dumps(line), flush=True)


def main():
    args = parse()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)
    if args.workload != "c2":
        secondary(args)
        return
    import torch

    from fuel_amd import _lib, workloads

    if _lib.device_count() < 1:
        raise SystemExit("no GPU visible")
    t_gen = time.time()
    X, Q = workloads.c2(n=args.n, m=args.m, d=args.d)
    print(f"# generated {X.shape} db and {Q.shape} queries in {time.time() - t_gen:.1f}s", file=sys.stderr,
          flush=True)
    dev = torch.device("cuda:0")
    Xd = torch.from_numpy(X).to(dev)
    Qd = torch.from_numpy(Q).to(dev)
    n, m, d, k = X.shape[0], Q.shape[0], Q.shape[1], args.k
    out_i = torch.empty((m, k), dtype=torch.int64, device=dev)
    out_d = torch.empty((m, k), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream().cuda_stream

    import oracle
    if args.host_path:
        # fuel_amd_knn with host arrays: H2D of the queries, kNN, D2H of the results
        ix = _lib.Index.from_device(Xd.data_ptr(), n, d, d, np.float32, "euclidean", mode=args.mode,
                                    widen=args.widen)
        for _ in range(args.warmup):
            hi, hd = ix.knn(Q, k, _lib.EXCLUDE_NONE)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            hi, hd = ix.knn(Q, k, _lib.EXCLUDE_NONE)
        t1 = time.perf_counter()
        checks = verify(oracle, X, Q, hi, hd, k, args.verify)
        print(json.dumps({"host_path_queries_per_s": round(m * args.steps / (t1 - t0), 1),
                          "ms_per_step": round((t1 - t0) / args.steps * 1e3, 3), "mode": args.mode,
                          "verify": checks}), flush=True)
        ix.close()
        return
    main_run = run_mode(_lib, Xd, Qd, n, m, d, k, args.mode, args.widen, args.steps, args.warmup, stream, out_i,
                        out_d, seed=bool(args.seed))
    print(f"# {args.mode}: {main_run['ms_per_step']:.2f} ms/step, {main_run['qps']:.0f} q/s", file=sys.stderr,
          flush=True)
    t_v = time.time()
    checks = verify(oracle, X, Q, out_i.cpu().numpy(), out_d.cpu().numpy(), k, args.verify,
                    extra=main_run["fail_rows"], per_block=args.verify_per_block)
    print(f"# verified {checks['queries']} queries ({checks.get('second_pass_queries', 0)} of them from the "
          f"certification chain) in {time.time() - t_v:.1f}s: {checks['exact_rows']} identical", file=sys.stderr,
          flush=True)

    # the plain FP32 MFMA form beside it (the metric is mixed vs FP32 at exact recall)
    fp32 = None
    if args.mode != "fp32" and args.fp32_steps > 0:
        r = run_mode(_lib, Xd, Qd, n, m, d, k, "fp32", args.widen, args.fp32_steps, 1, stream, out_i, out_d,
                     seed=bool(args.seed))
        print(f"# fp32: {r['ms_per_step']:.2f} ms/step, {r['qps']:.0f} q/s", file=sys.stderr, flush=True)
        c32 = verify(oracle, X, Q, out_i.cpu().numpy(), out_d.cpu().numpy(), k, args.verify)
        fp32 = {"value": round(r["qps"], 1), "ms_per_step": round(r["ms_per_step"], 3), "steps": args.fp32_steps,
                "kernel": r["kernel"], "kernel_ms": round(r["kernel_s"] * 1e3, 3),
                "roofline_frac": round(2.0 * m * n * d / r["kernel_s"] / PEAKS["fp32"][0], 4),
                "exact_rows": c32["exact_rows"], "recall": c32["recall"],
                "mixed_over_fp32": round(main_run["qps"] / r["qps"], 3)}

    flops = 2.0 * m * n * d
    peak, unit, bound = PEAKS[args.mode]
    avg_kernel_s = main_run["kernel_s"]
    achieved = flops / avg_kernel_s if avg_kernel_s > 0 else 0.0
    kname = main_run["kernel"]
    workload = f"kNN db {n}x{d} f32 uniform[0,1), {m} queries, k={k}, euclidean"
    roofline = {"bound": bound, "achieved": round(achieved / 1e12, 3), "peak": round(peak / 1e12, 1), "unit": unit,
                "frac": round(achieved / peak, 4), "traffic": profiled_traffic(kname, workload), "kernel": kname,
                "kernel_ms": round(avg_kernel_s * 1e3, 3), "algorithmic_flop_per_launch": flops}

    cpu = None
    if args.cpu_sample > 0:
        threads = oracle.default_threads()
        S = min(args.cpu_sample, m)
        oracle.knn(X, Q[:8], k, threads=threads)  # warm the page cache
        c0 = time.perf_counter()
        oracle.knn(X, Q[:S], k, threads=threads)
        c1 = time.perf_counter()
        cpu = {"value": round(S / (c1 - c0), 2), "unit": "queries/s", "cores": threads, "kind": "port",
               "sample": f"{S} of the {m} queries against the full {n}x{d} database, k={k}, "
                         f"{c1 - c0:.1f}s; restated AVX2 kernel + KNNHeap, one thread per query batch"}

    calib = None
    if args.calibrate:
        calib = calibration()
        calib["scan_frac_of_achievable"] = round(achieved / (calib["bf16_gemm_tflops"] * 1e12), 4) \
            if args.mode in ("bf16", "bf16r") else None
        print(f"# calibration: {calib}", file=sys.stderr, flush=True)

    del Xd, Qd
    torch.cuda.empty_cache()
    serve = None
    if args.serve_p99_seconds > 0:
        serve = serve_headline(args)
        print(f"# serve: {serve}", file=sys.stderr, flush=True)
    torch.cuda.empty_cache()

    c3 = None
    if args.c3_full:
        c3 = c3_full(oracle.default_threads())
        torch.cuda.empty_cache()

    c4 = None
    if args.c4_line:
        c4, _, _ = c4_run(args, oracle.default_threads(), args.c4_steps, 1)
        torch.cuda.empty_cache()

    stats = main_run["stats"]
    line = {
        "metric": METRIC,
        "value": round(main_run["qps"], 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_run["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"bf16": "bf16", "bf16r": "bf16", "fp32": "f32", "exact": "f32"}[args.mode],
        "data": "synthetic",
        "config": {"workload": workload,
                   "mode": args.mode, "widen": args.widen, "splits": stats["splits"], "kp": stats["kp"],
                   "fallback_queries": stats["fallback_queries"], "exact_rescans": stats["exact_queries"]},
        "verify": checks,
        "fp32": fp32,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "p99_serve_ms": serve["p99_ms"] if serve else None,
        "serve": serve,
        "calibration": calib,
        "c3_full_identical": c3["identical"] if c3 else None,
        "c3": c3,
        "c4": c4,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
