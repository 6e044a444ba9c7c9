"""Benchmark: exact kNN queries/s on one MI355X.

Workload (default): database 1,000,000 x 128 f32 uniform [0,1), 100,000 queries
from the same distribution, k = 10, Euclidean.  One step = one kNN pass of all
queries against the resident database (inputs already in HBM).  Prints one JSON
line with the throughput, the dominant kernel's roofline fraction (timed live
with HIP events on its launch stream) and the CPU baseline (the oracle's
restatement of fuel's AVX2 linear scan on the host cores, on a bounded sample).

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
    "fp32": (157.3e12, "TFLOP/s", "mfma"),
    "exact": (157.3e12, "TFLOP/s", "mfma"),
    "hbm": (8.0e12, "GB/s", "hbm"),
}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--mode", default="bf16", choices=["bf16", "fp32", "exact"])
    p.add_argument("--n", type=int, default=1_000_000)
    p.add_argument("--m", type=int, default=100_000)
    p.add_argument("--d", type=int, default=128)
    p.add_argument("--k", type=int, default=10)
    p.add_argument("--widen", type=int, default=4, help="candidates kept per query = widen * k (<= 128)")
    p.add_argument("--verify", type=int, default=64, help="queries checked against the oracle")
    p.add_argument("--cpu-sample", type=int, default=1000, help="queries timed on the CPU baseline (0: skip)")
    p.add_argument("--copy-probe", type=int, default=0,
                   help="only run a plain device copy of this many bytes 3 times (calibrates HBM counters)")
    return p.parse_args()


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


def main():
    args = parse()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)
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
    m, d, k = Q.shape[0], Q.shape[1], args.k
    out_i = torch.empty((m, k), dtype=torch.int64, device=dev)
    out_d = torch.empty((m, k), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    ix = _lib.Index.from_device(Xd.data_ptr(), X.shape[0], d, d, np.float32, "euclidean", mode=args.mode,
                                widen=args.widen)
    stream = torch.cuda.current_stream().cuda_stream

    def step():
        ix.knn_device(Qd.data_ptr(), m, d, k, _lib.EXCLUDE_NONE, out_i.data_ptr(), out_d.data_ptr(), stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    print("# warmup done", file=sys.stderr, flush=True)
    ix.timing_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kernel_ms, launches, kname = ix.timing_read()
    ix.timing_enable(False)
    stats = ix.stats()
    elapsed = t1 - t0
    ms_per_step = elapsed / args.steps * 1e3
    qps = m * args.steps / elapsed
    print(f"# timed {args.steps} steps: {ms_per_step:.2f} ms/step, {qps:.0f} q/s", file=sys.stderr, flush=True)

    # verification on a sample of queries against the oracle (exact indices and f32 bits)
    import oracle
    gi = out_i.cpu().numpy()
    gd = out_d.cpu().numpy()
    rng = np.random.default_rng(0)
    sample = np.sort(rng.choice(m, size=min(args.verify, m), replace=False))
    oi, od = oracle.knn(X, Q[sample], k)
    exact_rows = int(np.sum(np.all(gi[sample] == oi, axis=1) & np.all(gd[sample] == od, axis=1)))
    recall = float(np.mean([len(set(gi[s]) & set(oi[j])) / k for j, s in enumerate(sample)]))
    qi = np.repeat(sample, k)
    xi = gi[sample].ravel()
    ref64 = np.sqrt(oracle.sqdist_fp64(X, Q, qi, xi))
    max_rel_err = float(np.max(np.abs(gd[sample].ravel().astype(np.float64) - ref64) / np.maximum(ref64, 1e-300)))

    flops = 2.0 * m * X.shape[0] * d
    peak, unit, bound = PEAKS[args.mode]
    avg_kernel_s = (kernel_ms / max(launches, 1)) / 1e3
    achieved = flops / avg_kernel_s if avg_kernel_s > 0 else 0.0
    roofline = {"bound": bound, "achieved": round(achieved / 1e12, 3), "peak": round(peak / 1e12, 1), "unit": unit,
                "frac": round(achieved / peak, 4), "traffic": None, "kernel": kname,
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
               "sample": f"{S} of the {m} queries against the full {X.shape[0]}x{d} database, k={k}, "
                         f"{c1 - c0:.1f}s; restated AVX2 kernel + KNNHeap, one thread per query batch"}

    line = {
        "metric": "kNN queries/s, 1M x 128 f32, k=10, exact (recall 1.0 vs fuel's linear scan)",
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"bf16": "bf16", "fp32": "f32", "exact": "f32"}[args.mode],
        "data": "synthetic",
        "config": {"workload": f"kNN db {X.shape[0]}x{d} f32 uniform[0,1), {m} queries, k={k}, euclidean",
                   "mode": args.mode, "splits": stats["splits"], "kp": stats["kp"],
                   "fallback_queries": stats["fallback_queries"], "exact_rescans": stats["exact_queries"]},
        "verify": {"queries": int(len(sample)), "exact_rows": exact_rows, "recall": recall,
                   "max_rel_err_vs_fp64": max_rel_err},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
