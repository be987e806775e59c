#!/usr/bin/env python3
"""Benchmark of the MI355X hash aggregate (headline) on one GPU.

Workload: 100M-row composite GROUP BY (a int32 Zipf(1.1) over 1M values,
b int32 uniform over 16) with SUM(price DECIMAL(15,2)) into Int128, COUNT(*),
MIN/MAX/AVG(f float64), WHERE f > -0.5 fused as a predicate. Inputs are
generated into HBM before timing. One step = reset + consume + finalize.

Prints ONE JSON line. `roofline` uses the consume kernel's device time (HIP
events on the library stream) and its algorithmic bytes (24 B/row + 56 B/group);
`cpu_baseline` times the C restatement (oracle/) on this host's cores on a
bounded prefix of the same rows.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(cols, n_rows, sample_rows, threads):
    import numpy as np

    import oracle as O
    from wadjet_amd import workloads as W
    m = min(n_rows, sample_rows)
    host = [(cols[0].data.to_host(np.int32, m), O.INT32), (cols[1].data.to_host(np.int32, m), O.INT32),
            (cols[2].data.to_host(np.int64, m), O.DECIMAL, None, 2), (cols[3].data.to_host(np.float64, m), O.FLOAT64)]
    aggs = [(a.func, a.col) for a in W.CFG2_AGGS]
    best = None
    for rep in range(2):
        t0 = time.perf_counter()
        r = O.hash_agg(host, W.CFG2_KEYS, aggs, pred=(3, O.GT, 0, -0.5), threads=threads, max_groups=m)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
        log(f"[cpu] rep {rep}: {m} rows, {r['n']} groups, {dt:.2f} s on {threads} threads")
    return dict(value=m / best, unit="rows/s", cores=threads, kind="port",
                sample=f"first {m} rows of the same 100M-row input (C restatement, {threads} threads: "
                       f"per-thread partial hash aggregates merged in thread order, like CloneSink/MergeSink)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--rows", type=int, default=100_000_000)
    ap.add_argument("--group-hint", type=int, default=8_000_000)
    ap.add_argument("--lds-slots", type=int, default=-1)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--cpu-rows", type=int, default=25_000_000)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (the multi-GPU "
              f"design is described, not built)", file=sys.stderr)
        sys.exit(2)

    from wadjet_amd import workloads as W
    from wadjet_amd.gpu import Context, HashAggregate

    ctx = Context(0)
    n = args.rows
    log(f"[bench] generating {n} rows on device")
    cols = W.gen_config2(ctx, n)
    agg = HashAggregate(ctx, W.CFG2_TYPES, W.CFG2_KEYS, W.CFG2_AGGS, scales=W.CFG2_SCALES,
                        group_hint=args.group_hint)
    if args.lds_slots >= 0 or args.block:
        agg.set_tuning(args.lds_slots if args.lds_slots >= 0 else 1024, args.block or 512)

    def step():
        agg.reset()
        agg.consume(cols, n, pred=W.CFG2_PRED)
        agg.finalize()

    for i in range(args.warmup):
        step()
        log(f"[bench] warmup {i} done")
    ctx.sync()
    kern_ms, passes = [], []
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        kern_ms.append(agg.kernel_time()[0])
        passes.append(agg.pass_times())
    ctx.sync()
    wall = time.perf_counter() - t0
    groups = agg.num_groups()
    st = agg.stats()
    ms_step = wall * 1000 / args.steps
    k_ms = sum(kern_ms) / len(kern_ms)
    alg_bytes = W.CFG2_BYTES_PER_ROW * n + W.CFG2_BYTES_PER_GROUP * groups
    achieved = alg_bytes / (k_ms / 1000) / 1e9
    log(f"[bench] {ms_step:.3f} ms/step, consume kernels {k_ms:.3f} ms, groups {groups}, stats {st}")
    log(f"[bench] pass times (direct, count+scan, scatter, partition-agg) last step: {passes[-1]}")

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        cpu = cpu_baseline(cols, n, args.cpu_rows, threads)

    out = {
        "metric": "rows/s & %HBM roofline for hash-agg/join probe; batched vs plain qps + p99",
        "value": n / (ms_step / 1000),
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64+int128+f64",
        "data": "synthetic (seeded device generators)",
        "config": {"workload": "hash-aggregate 100M rows: GROUP BY (a int32 Zipf1.1/1M, b int32 u16) "
                               "SUM(price DECIMAL(15,2))->Int128, COUNT(*), MIN/MAX/AVG(f f64), WHERE f > -0.5",
                   "rows": n, "groups": groups, "lds_slots": agg_tuning(agg), "parallelism": "1 GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "agg_consume_kernel", "kernel_ms": k_ms,
                     "alg_bytes": alg_bytes},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def agg_tuning(agg):
    return None


if __name__ == "__main__":
    main()
