#!/usr/bin/env python3
"""Benchmark of the MI355X hash aggregate / hash join on one GPU.

Headline (the JSON line's `value`): the 100M-row composite GROUP BY
(a int32 Zipf(1.1) over 1M values, b int32 uniform over 16) with
SUM(price DECIMAL(15,2)) into Int128, COUNT(*), MIN/MAX/AVG(f float64) and
WHERE f > -0.5 fused. Inputs are generated into HBM before timing; one step =
reset + consume + finalize + the output columns written on the device.

Beside it, in the same line:
  join     10M-row build, 200M-row probe, fused JOIN fk = key GROUP BY p1
           SUM(qty * p2) (exact Int128), probe rows/s
  batched  256 GROUP BY queries of uneven shapes: one persistent graph-replayed
           launch (autotuned) vs one launch per query: q/s and latency
  served   request loop over a resident flow table: Poisson arrivals at fixed
           rates, continuous batching vs one request per launch: p50/p99, q/s
`roofline` is the headline's consume kernels: algorithmic bytes
(24 B/row + 56 B/group) over their device time (HIP events on the library
stream). `cpu_baseline` times the C restatement (oracle/) on this host's cores
over the same rows (`join.cpu_baseline` likewise for the join).
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
# calibrated rocprofv3 FETCH_SIZE + WRITE_SIZE of the consume kernels (see the file's header)
AGG_PMC_SUMMARY = ROOT / "profiles" / "groupby_zipf_100M_pmc.txt"
JOIN_PMC_SUMMARY = ROOT / "profiles" / "join_probe_agg_200M_pmc.txt"


def is_consume_kernel(name):
    return any(s in name for s in ("part_l1", "part_tot", "part_l2", "part_agg", "rocprim",
                                   "agg_consume_kernel", "hot_"))


def is_join_agg_kernel(name):
    return "join_agg" in name


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def pct(xs, p):
    import numpy as np
    return float(np.percentile(np.asarray(xs), p)) if len(xs) else None


# ----------------------------------------------- Zipf composite GROUP BY, 100M rows

def bench_agg(ctx, args):
    from wadjet_amd import workloads as W
    from wadjet_amd.gpu import HashAggregate
    n = args.rows
    log(f"[agg] generating {n} rows on device")
    cols = W.gen_zipf_groupby(ctx, n)
    agg = HashAggregate(ctx, W.ZIPF_TYPES, W.ZIPF_KEYS, W.ZIPF_AGGS, scales=W.ZIPF_SCALES,
                        group_hint=args.group_hint)
    if args.strategy >= 0 or args.part_lds_slots or args.part_block:
        agg.set_strategy(max(args.strategy, 0), part_lds_slots=args.part_lds_slots or 2048,
                         part_block=args.part_block or 1024)
    if args.hot_slots >= 0:
        agg.set_heavy_hitters(args.hot_slots)

    # the step ends with every output column on the device (keys, finalized
    # values, validity): the Next drain, without the copy to the host
    agg.reset()
    agg.consume(cols, n, pred=W.ZIPF_PRED)
    agg.finalize()
    out = agg.output_buffers(agg.num_groups() + (agg.num_groups() >> 4) + 1024)

    def step():
        agg.reset()
        agg.consume(cols, n, pred=W.ZIPF_PRED)
        agg.finalize()
        agg.fetch_device(*out)

    for i in range(args.warmup):
        step()
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
    ms_step = wall * 1000 / args.steps
    k_ms = sum(kern_ms) / len(kern_ms)
    alg = W.ZIPF_BYTES_PER_ROW * n + W.ZIPF_BYTES_PER_GROUP * groups
    cold = agg.stats()["global_rows"]  # rows the sorting passes wrote to the partitions
    log(f"[agg] {ms_step:.3f} ms/step, consume kernels {k_ms:.3f} ms, groups {groups}; "
        f"passes (direct, sort level 1, sort level 2, partition-agg) {passes[-1]}; partitioned rows {cold}")
    # each pass against its own algorithmic bytes: level 1 reads the 24 B row
    # and writes a 24 B record (key, price, f) per cold row; level 2 reads and
    # writes each record once; the partition aggregate reads the records and
    # writes 56 B per group
    p = passes[-1]
    per_pass = {}
    for name, ms, b in (("sort_level1", p[1], 24 * n + 24 * cold), ("sort_level2", p[2], 48 * cold),
                        ("partition_agg", p[3], 24 * cold + W.ZIPF_BYTES_PER_GROUP * groups)):
        if ms > 0:
            per_pass[name] = dict(ms=ms, alg_bytes=b, achieved=b / (ms / 1000) / 1e9,
                                  frac=b / (ms / 1000) / 1e9 / HBM_PEAK_GBS)
    return dict(cols=cols, n=n, groups=groups, ms_step=ms_step, kernel_ms=k_ms, alg_bytes=alg,
                passes=passes[-1], per_pass=per_pass, cold_rows=cold, agg=agg)


def cpu_threads(requested=0):
    """Host threads for the CPU baselines: the cores this process may run on,
    at most 16 (the share of the host a one-GPU job gets)."""
    avail = len(os.sched_getaffinity(0))
    return (requested or min(avail, 16)), avail


def reps_summary(times, rows):
    import statistics
    return dict(reps=len(times), min_s=min(times), median_s=statistics.median(times), rows=rows)


def cpu_agg(cols, n, sample_rows, threads, host_cpus, reps):
    """The C restatement of HashAggregate (oracle/) on `threads` host threads:
    each aggregates a contiguous row range (a morsel worker's partial), the
    partials are merged in worker order (CloneSink / MergeSink), over the same
    rows the GPU step reads. Generation excluded; the merge included."""
    import numpy as np

    import oracle as O
    from wadjet_amd import workloads as W
    m = min(n, sample_rows)
    host = [(cols[0].data.to_host(np.int32, m), O.INT32), (cols[1].data.to_host(np.int32, m), O.INT32),
            (cols[2].data.to_host(np.int64, m), O.DECIMAL, None, 2), (cols[3].data.to_host(np.float64, m), O.FLOAT64)]
    aggs = [(a.func, a.col) for a in W.ZIPF_AGGS]
    times = []
    for rep in range(reps):
        t0 = time.perf_counter()
        r = O.hash_agg(host, W.ZIPF_KEYS, aggs, pred=(3, O.GT, 0, -0.5), threads=threads,
                       max_groups=min(m, 8_000_000))
        times.append(time.perf_counter() - t0)
        log(f"[cpu] agg rep {rep}: {m} rows, {r['n']} groups, {times[-1]:.2f} s on {threads} threads")
    st = reps_summary(times, m)
    return dict(value=m / st["min_s"], unit="rows/s", cores=threads, host_cpus=host_cpus, kind="port", **st,
                sample=f"{'all' if m == n else 'the first'} {m} rows of the same input; C restatement of "
                       f"HashAggregate on {threads} threads, each aggregating a contiguous row range, partials "
                       f"merged in worker order (CloneSink/MergeSink); value = rows / fastest of {reps} reps; "
                       f"generation excluded, merge included")


# ------------------------------------- join 10M build x 200M probe, then GROUP BY

def bench_join(ctx, args):
    from wadjet_amd import _lib as L
    from wadjet_amd import workloads as W
    from wadjet_amd.gpu import AggSpec, HashAggregate, HashJoin
    cfg = W.FkJoin(n_build=args.join_build, n_probe=args.join_probe)
    log(f"[join] generating build {cfg.n_build} / probe {cfg.n_probe}")
    bcols = W.gen_fk_join_build(ctx, cfg)
    pcols = W.gen_fk_join_probe(ctx, cfg)
    j = HashJoin(ctx, W.FKJ_BUILD_TYPES, 0, inline=[1, 2], scales=W.FKJ_BUILD_SCALES, build_hint=cfg.n_build)
    ctx.timer_start()
    j.build(bcols, cfg.n_build)
    j.finish()
    build_ms = ctx.timer_stop()
    types = W.FKJ_PROBE_TYPES + W.FKJ_BUILD_TYPES
    aggs = [AggSpec(L.SUM, 1, 4), AggSpec(L.COUNT, -1)]
    agg = HashAggregate(ctx, types, [3], aggs, scales=[0, 0, 0, 0, 2], group_hint=64)

    def step():
        agg.reset()
        j.probe_agg(pcols, 0, cfg.n_probe, agg)
        agg.finalize()

    for _ in range(args.warmup):
        step()
    ctx.sync()
    kms = []
    t0 = time.perf_counter()
    steps = max(1, args.steps)
    for _ in range(steps):
        step()
        kms.append(agg.kernel_time()[0])
    ctx.sync()
    ms = (time.perf_counter() - t0) * 1000 / steps
    k = sum(kms) / len(kms)
    alg = W.FKJ_BYTES_PER_PROBE_ROW * cfg.n_probe
    from wadjet_amd.profiling import load_traffic
    traffic = load_traffic(str(JOIN_PMC_SUMMARY), is_join_agg_kernel)
    floor = line_floor(ctx, cfg.n_probe)
    log(f"[join] build {build_ms:.2f} ms, probe+agg {ms:.3f} ms/step (kernel {k:.3f} ms), groups {agg.num_groups()}; "
        f"random-line floor for {cfg.n_probe} lookups {floor['ms']:.3f} ms")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_join(bcols, pcols, cfg, *cpu_threads(args.cpu_threads), args.cpu_reps)
    return dict(rows_per_s=cfg.n_probe / (ms / 1000), ms_per_step=ms, kernel_ms=k, build_ms=build_ms,
                build_rows_per_s=cfg.n_build / (build_ms / 1000), groups=agg.num_groups(),
                roofline=dict(bound="hbm", achieved=alg / (k / 1000) / 1e9, peak=HBM_PEAK_GBS, unit="GB/s",
                              frac=alg / (k / 1000) / 1e9 / HBM_PEAK_GBS, alg_bytes=alg,
                              kernel="join_agg_lean_kernel (fused probe -> GROUP BY)", traffic=traffic,
                              traffic_source=JOIN_PMC_SUMMARY.name if traffic else None,
                              line_floor=floor, of_line_floor=floor["ms"] / k),
                cpu_baseline=cpu, _keep=(j, agg, bcols, pcols))


def line_floor(ctx, n_lookups):
    """The random-line rate that bounds a hash probe beyond the L2: the
    library's gather (gatherVector) of n_lookups uniformly random rows of a
    1 GiB int64 column (the size of the join's slot array), each row one
    128-byte line request, with a 4-byte index read and an 8-byte value
    written per row."""
    from wadjet_amd import _lib as L
    from wadjet_amd.gpu import Col, gather
    rows = 1 << 27
    col = ctx.gen(L.GEN_UNIFORM_I64, "int64", rows, 91, 0, 1 << 40)
    idx = ctx.gen(L.GEN_UNIFORM_I32, "int32", n_lookups, 92, 0, rows - 1)
    out = ctx.alloc(n_lookups * 8)
    gather(ctx, Col(col, L.INT64), idx, n_lookups, out)
    ctx.sync()
    best = None
    for _ in range(3):
        ctx.timer_start()
        gather(ctx, Col(col, L.INT64), idx, n_lookups, out)
        t = ctx.timer_stop()
        best = t if best is None else min(best, t)
    for b in (col, idx, out):
        b.free()
    return dict(ms=best, lookups=n_lookups, lines_per_s=n_lookups / (best / 1000), table_bytes=rows * 8,
                kernel="gather_kernel (wg_gather)")


def cpu_join(bcols, pcols, cfg, threads, host_cpus, reps):
    """The C restatement of HashJoin.Build + the probe feeding GROUP BY p1
    SUM(qty * p2) (oracle join_agg): the build serial, the 200M probe rows over
    `threads` threads (morsel workers), partial groups merged in worker order."""
    import numpy as np

    import oracle as O
    fk = pcols[0].data.to_host(np.int64, cfg.n_probe)
    qty = pcols[1].data.to_host(np.int32, cfg.n_probe)
    key = bcols[0].data.to_host(np.int64, cfg.n_build)
    p1 = bcols[1].data.to_host(np.int32, cfg.n_build)
    p2 = bcols[2].data.to_host(np.int64, cfg.n_build)
    times = []
    for rep in range(reps):
        t0 = time.perf_counter()
        r = O.join_agg((key, O.INT64), (p1, O.INT32), (p2, O.DECIMAL, None, 2), (fk, O.INT64), (qty, O.INT32),
                       threads=threads)
        times.append(time.perf_counter() - t0)
        log(f"[cpu] join rep {rep}: {cfg.n_probe} probe rows, {r['n']} groups, {times[-1]:.2f} s on {threads} threads")
    st = reps_summary(times, cfg.n_probe)
    return dict(value=cfg.n_probe / st["min_s"], unit="probe rows/s", cores=threads, host_cpus=host_cpus,
                kind="port", **st,
                sample=f"all {cfg.n_build} build and {cfg.n_probe} probe rows; C restatement of HashJoin.Build "
                       f"(serial) and the probe feeding the GROUP BY on {threads} threads; value = probe rows / "
                       f"fastest of {reps} reps, build included")


# -------------------------------------------- join output: views vs eager gathers

def bench_join_output(ctx, args):
    """The materializing inner join of the fused-join workload (build 10M, probe
    200M, ~190M matches): the probe emitting its (probe row, build row) arrays
    as the view columns' indices (late materialization, emitViewOutput) vs the
    same probe followed by the eager gather of all five output columns (fk,
    qty, key, p1, p2) with their validity (wg_view_flatten)."""
    from wadjet_amd import workloads as W
    from wadjet_amd.gpu import HashJoin, view_flatten
    cfg = W.FkJoin(n_build=args.join_build, n_probe=args.join_probe)
    log(f"[join_output] generating build {cfg.n_build} / probe {cfg.n_probe}")
    bcols = W.gen_fk_join_build(ctx, cfg)
    pcols = W.gen_fk_join_probe(ctx, cfg)
    j = HashJoin(ctx, W.FKJ_BUILD_TYPES, 0, inline=[1, 2], scales=W.FKJ_BUILD_SCALES, build_hint=cfg.n_build)
    j.build(bcols, cfg.n_build)
    j.finish()
    cap = cfg.n_probe + 1024  # unique build keys: at most one match per probe row
    outp, outb = ctx.alloc(cap * 4), ctx.alloc(cap * 4)
    bases = [(c, cfg.n_probe, outp) for c in pcols] + [(j.build_column(i), cfg.n_build, outb) for i in range(3)]
    from wadjet_amd.gpu import TYPE_BYTES
    outs = [(ctx.alloc(cap * TYPE_BYTES[c.type]), ctx.alloc(((cap + 63) // 64) * 8)) for c, _, _ in bases]

    def probe():
        got, cur = j.probe(pcols[0], cfg.n_probe, outp, outb, cap)
        assert cur.done == 1
        return got

    def eager():
        got = probe()
        for (col, rows, idx), (o, ov) in zip(bases, outs):
            view_flatten(ctx, col, rows, idx, got, out=o, out_valid=ov)
        return got

    res = {}
    for name, fn in (("views", probe), ("eager", eager)):
        for _ in range(max(args.warmup, 1)):
            fn()
        ctx.sync()
        t0 = time.perf_counter()
        steps = max(1, args.steps)
        for _ in range(steps):
            got = fn()
        ctx.sync()
        res[name] = (time.perf_counter() - t0) * 1000 / steps
    gathered = got * (8 + 4 + 8 + 4 + 8)
    for anti in (False, True):  # semi / anti probe of the same rows (ordered selection vectors)
        j.semi(pcols[0], cfg.n_probe, outp, anti=anti)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(max(1, args.steps)):
            kept = j.semi(pcols[0], cfg.n_probe, outp, anti=anti)
        ctx.sync()
        res["anti" if anti else "semi"] = (time.perf_counter() - t0) * 1000 / max(1, args.steps)
        res["anti_rows" if anti else "semi_rows"] = kept
    for b in [outp, outb] + [x for pair in outs for x in pair]:
        b.free()
    log(f"[join_output] {got} pairs: views {res['views']:.3f} ms, eager gathers {res['eager']:.3f} ms "
        f"({gathered / 1e9:.2f} GB of gathered columns deferred); semi {res['semi']:.3f} ms "
        f"({res['semi_rows']} rows), anti {res['anti']:.3f} ms ({res['anti_rows']} rows)")
    return dict(pairs=got, views_ms=res["views"], eager_ms=res["eager"], deferred_bytes=gathered,
                semi_ms=res["semi"], anti_ms=res["anti"], semi_rows=res["semi_rows"], anti_rows=res["anti_rows"],
                pairs_per_s_views=got / (res["views"] / 1000), pairs_per_s_eager=got / (res["eager"] / 1000))


# -------------------------------------------- shuffle routing of the GROUP BY rows

def bench_shuffle(ctx, args, cols=None):
    """Partitioned-shuffle routing (wg_shuffle_partition) of the 100M-row
    (a int32, b int32) keys into 256 partitions: FNV-1a per row, stable
    per-partition row lists. Algorithmic bytes: the 8 key bytes read, the
    4-byte partition id and the 4-byte row id written."""
    from wadjet_amd import workloads as W
    from wadjet_amd.gpu import shuffle_partition
    n = args.rows
    if cols is None:
        log(f"[shuffle] generating {n} rows on device")
        cols = W.gen_zipf_groupby(ctx, n)
    parts_n = args.shuffle_parts
    pbuf, rbuf = ctx.alloc(n * 4), ctx.alloc(n * 4)
    for _ in range(max(args.warmup, 1)):
        shuffle_partition(ctx, cols[:2], [0, 1], parts_n, n, parts=pbuf, rows=rbuf)
    ctx.sync()
    ms = []
    for _ in range(max(1, args.steps)):
        ctx.timer_start()
        shuffle_partition(ctx, cols[:2], [0, 1], parts_n, n, parts=pbuf, rows=rbuf)
        ms.append(ctx.timer_stop())
    pbuf.free()
    rbuf.free()
    m = sum(ms) / len(ms)
    alg = 16 * n
    log(f"[shuffle] {n} rows -> {parts_n} partitions: {m:.3f} ms ({n / m / 1e6:.1f} G rows/s)")
    return dict(rows=n, partitions=parts_n, ms=m, rows_per_s=n / (m / 1000), alg_bytes=alg,
                achieved_gbs=alg / (m / 1000) / 1e9, frac=alg / (m / 1000) / 1e9 / HBM_PEAK_GBS)


# -------------------------------------------- GROUP BY on a 128-bit key

def bench_wide(ctx, args):
    """GROUP BY (a int64 uniform over 1M, b int64 uniform over 4) SUM(v),
    COUNT(*) over 100M rows: a two-int64 key takes the reference's 128-bit
    {lo, hi} layout, dictionary-encoded on the device."""
    from wadjet_amd import _lib as L
    from wadjet_amd.gpu import AggSpec, Col, HashAggregate
    n = args.rows
    log(f"[wide] generating {n} rows on device")
    a = ctx.gen(L.GEN_UNIFORM_I64, "int64", n, 21, 0, 999_999)
    b = ctx.gen(L.GEN_UNIFORM_I64, "int64", n, 22, 0, 3)
    v = ctx.gen(L.GEN_UNIFORM_I64, "int64", n, 23, -10**9, 10**9)
    cols = [Col(a, L.INT64), Col(b, L.INT64), Col(v, L.INT64)]
    agg = HashAggregate(ctx, [L.INT64, L.INT64, L.INT64], [0, 1], [AggSpec(L.SUM, 2), AggSpec(L.COUNT, -1)],
                        group_hint=4_000_000)

    def step():
        agg.reset()
        agg.consume(cols, n)
        agg.finalize()

    for _ in range(max(args.warmup, 1)):
        step()
    ctx.sync()
    t0 = time.perf_counter()
    steps = max(1, args.steps)
    for _ in range(steps):
        step()
    ctx.sync()
    ms = (time.perf_counter() - t0) * 1000 / steps
    g = agg.num_groups()
    log(f"[wide] {n} rows, {g} groups: {ms:.3f} ms/step ({n / ms / 1e6:.1f} G rows/s)")
    agg.close()
    # algorithmic bytes: the two key columns and the value column read once,
    # each group's 128-bit key, sum and count written once
    alg = n * 24 + g * 32
    gbs = alg / (ms / 1000) / 1e9
    return dict(rows=n, groups=g, ms_per_step=ms, rows_per_s=n / (ms / 1000),
                roofline=dict(bound="hbm", achieved=gbs, peak=HBM_PEAK_GBS, unit="GB/s", frac=gbs / HBM_PEAK_GBS,
                              alg_bytes=alg, kernel="whole step (dictionary claim pass + partitioned consume)"))


# -------------------------------------------- Parquet pages decoded on the device

def bench_parquet(ctx, args):
    """A 100M-row Parquet file (pyarrow's default writer, snappy, 4M-row row
    groups) of k int64 uniform over 1M (dictionary, falling back to PLAIN as
    the dictionary outgrows its page), g int32 uniform over 16 (dictionary,
    4-bit indices), v float64 with 10 % NULLs (PLAIN, definition levels):
    every page of every chunk decompressed on the host and staged in HBM once,
    then the timed step decodes the three columns on the device
    (wg_parquet_decode_pages, one call per column). Algorithmic bytes: the
    decompressed page streams read once, the values and validity written."""
    import tempfile

    import numpy as np
    import pyarrow as pa
    import pyarrow.parquet as pq

    from wadjet_amd import parquet as PQ
    n = args.rows
    rng = np.random.default_rng(7)
    t = pa.table({"k": pa.array(rng.integers(0, 1_000_000, n), pa.int64()),
                  "g": pa.array(rng.integers(0, 16, n).astype(np.int32), pa.int32()),
                  "v": pa.array(rng.normal(size=n), pa.float64(), mask=rng.random(n) < 0.1)})
    d = tempfile.mkdtemp(prefix="wg_pq_")
    path = os.path.join(d, "bench.parquet")
    log(f"[parquet] writing {n} rows")
    pq.write_table(t, path, row_group_size=4_000_000)
    del t
    t0 = time.perf_counter()
    staged = [PQ.StagedColumn(ctx, path, c) for c in ("k", "g", "v")]
    stage_s = time.perf_counter() - t0
    outs = [ctx.alloc(n * st.width) for st in staged]
    valids = [ctx.alloc(((n + 63) // 64) * 8) if st.max_def > 0 else None for st in staged]

    def step():
        for st, o, vb in zip(staged, outs, valids):
            if vb is not None:
                ctx.memset(vb, 0)
            st.decode(o, vb, 0)

    for _ in range(max(args.warmup, 1)):
        step()
    ctx.sync()
    steps = max(1, args.steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    ctx.sync()
    ms = (time.perf_counter() - t0) * 1000 / steps
    page_bytes = sum(st.page_bytes for st in staged)
    out_bytes = sum(n * st.width + (((n + 63) // 64) * 8 if st.max_def > 0 else 0) for st in staged)
    alg = page_bytes + out_bytes
    t0 = time.perf_counter()
    pq.read_table(path, columns=["k", "g", "v"])
    arrow_ms = (time.perf_counter() - t0) * 1000
    pages = sum(st.n_pages for st in staged)
    for st in staged:
        st.free()
    for b in outs + [v for v in valids if v is not None]:
        b.free()
    os.remove(path)
    os.rmdir(d)
    gbs = alg / (ms / 1000) / 1e9
    log(f"[parquet] {n} rows x 3 columns, {pages} pages: {ms:.3f} ms/step ({n / ms / 1e6:.1f} G rows/s, "
        f"{gbs:.0f} GB/s); host staging {stage_s:.2f} s; pyarrow read_table {arrow_ms:.0f} ms")
    return dict(rows=n, columns=3, pages=pages, ms_per_step=ms, rows_per_s=n / (ms / 1000),
                page_bytes=page_bytes, out_bytes=out_bytes, achieved_gbs=gbs, frac=gbs / HBM_PEAK_GBS,
                host_stage_s=stage_s, pyarrow_read_ms=arrow_ms)


# -------------------------------------------- 256 concurrent queries, one launch

def bench_batched(ctx, args):
    """256 GROUP BY / JOIN dim GROUP BY payload queries of uneven shapes. The
    autotuner picks its tuning on one seeded mix; a different mix (another
    seed, same distributions) is timed: one graph-replayed persistent launch
    vs one launch sequence per query."""
    from wadjet_amd import _lib as L
    from wadjet_amd import workloads as W
    from wadjet_amd.gpu import TYPE_BYTES, Batch
    tune_plan = W.plan_query_mix(n_queries=args.batch_queries, seed=4)
    plan = W.plan_query_mix(n_queries=args.batch_queries, seed=5)
    dim = W.gen_query_dim(ctx)
    tune_data = W.gen_query_mix(ctx, tune_plan, dim)
    data = W.gen_query_mix(ctx, plan, dim)
    total_rows = sum(q["rows"] for q in plan)
    qarr = Batch.queries(data)
    b = Batch(ctx, args.batch_queries, max(sum(q["groups"] for q in plan), sum(q["groups"] for q in tune_plan)))
    t0 = time.perf_counter()
    if args.batch_tuning:  # a fixed tuning (profiling runs): no autotuner launches
        b.set_tuning(*[int(x) for x in args.batch_tuning.split(",")])
    else:
        b.autotune(Batch.queries(tune_data), len(tune_plan))
    tune_s = time.perf_counter() - t0
    del tune_data
    tuning = b.tuning()
    b.run(qarr, len(plan), use_graph=True)  # warm the graph
    reps = max(3, args.steps)
    bt, dev = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        ctx.timer_start()
        b.run(qarr, len(plan), use_graph=True)
        dev.append(ctx.timer_stop())
        bt.append(time.perf_counter() - t0)
    choices = [b.choice(q) for q in range(len(plan))]
    n_lds = sum(1 for c in choices if c[0] == 1)
    n_cache = sum(1 for c in choices if c[0] == 2)
    n_fan = sum(1 for c in choices if c[2] > 1)
    if args.batch_no_plain:  # profiling runs: the batched launches only
        return dict(queries=len(plan), batched_ms=min(bt) * 1000, device_ms=min(dev), tuning=b.tuning())
    # one launch per query, serialised; latency of query i = its completion time
    b.run(qarr, len(plan), plain=True)  # warm
    pt = []
    for _ in range(max(1, reps // 2)):
        t0 = time.perf_counter()
        b.run(qarr, len(plan), plain=True)
        pt.append(time.perf_counter() - t0)
    # per-query plain latencies (all submitted at t=0, run one after another)
    lat = []
    t0 = time.perf_counter()
    for q in range(len(plan)):
        b.run(Batch.queries([data[q]]), 1, plain=True)
        lat.append((time.perf_counter() - t0) * 1000)
    batched_s = min(bt)
    plain_s = min(pt)
    dev_ms = min(dev)
    # algorithmic bytes: every query's key and value columns read once
    alg = sum(n * (TYPE_BYTES[kc.type] + TYPE_BYTES[vc.type]) for kc, vc, n, g, *x in data)
    n_join = sum(1 for q in plan if q["kind"] == "join")
    log(f"[batched] {len(plan)} queries ({n_join} joins), {total_rows} rows: batched {batched_s*1000:.2f} ms "
        f"(device {dev_ms:.2f} ms), plain {plain_s*1000:.2f} ms, tuning {tuning} ({tune_s:.1f} s, on another mix); "
        f"{n_lds} queries on LDS tables ({n_fan} fanned out), {n_cache} on global tables behind an LDS cache")
    return dict(queries=len(plan), join_queries=n_join, rows=total_rows, batched_ms=batched_s * 1000,
                device_ms=dev_ms, plain_ms=plain_s * 1000,
                batched_qps=len(plan) / batched_s, plain_qps=len(plan) / plain_s,
                batched_p50_ms=batched_s * 1000, batched_p99_ms=batched_s * 1000,
                latency_note="every query of the batch completes with its one launch, so the batched p50 and p99 "
                             "are that launch's time; plain latencies are completion times of queries run one "
                             "after another from t=0",
                plain_p50_ms=pct(lat, 50), plain_p99_ms=pct(lat, 99),
                batched_rows_per_s=total_rows / batched_s, speedup=plain_s / batched_s,
                tuning=tuning, tuned_on="seed-4 mix (timed: seed-5 mix)", lds_queries=n_lds, fanned_out_queries=n_fan,
                lds_cached_queries=n_cache,
                autotune_s=tune_s,
                roofline=dict(bound="hbm", achieved=alg / (dev_ms / 1000) / 1e9, peak=HBM_PEAK_GBS, unit="GB/s",
                              frac=alg / (dev_ms / 1000) / 1e9 / HBM_PEAK_GBS, alg_bytes=alg,
                              kernel="one batched launch (fill, batch_agg_kernel, batch_compact_kernel)"))


# ------------------------------------------------ served flow table, Poisson load

def bench_served(ctx, args):
    from wadjet_amd import workloads as W
    from wadjet_amd.gpu import HashJoin, Server
    n = args.serve_rows
    log(f"[served] generating a {n}-row flow table")
    cols = W.gen_flow_table(ctx, n)
    dim = W.gen_flow_dim(ctx, 50_000)
    j = HashJoin(ctx, W.FLOW_DIM_TYPES, 0, inline=[1], build_hint=50_000)
    j.build(dim, 50_000)
    j.finish()
    served = Server(ctx, cols, n, 3, lanes=3, hbm_budget=8 << 30, max_batch=64)
    plain = Server(ctx, cols, n, 3, plain=True)
    served.add_dim(j)
    plain.add_dim(j)
    # warm both at the top offered rate with that point's request count (code
    # objects, graphs, staging buffers and the pool of result buffers that a
    # server under that load keeps), unmeasured
    top = max(args.qps)
    wq, _ = W.plan_flow_requests(n, max(50, min(args.serve_requests, int(top * args.serve_seconds))), seed=7)
    for s in (served, plain):
        s.replay(wq, qps=top, seed=3)
        s.reset()
    out = []
    for qps in args.qps:
        n_req = max(50, min(args.serve_requests, int(qps * args.serve_seconds)))
        qs, shapes = W.plan_flow_requests(n, n_req, seed=int(qps))
        res = {"qps_offered": qps, "requests": n_req}
        for name, s in (("served", served), ("plain", plain)):
            s.reset()
            lat, _ = s.replay(qs, qps=qps, seed=1)
            st = s.stats()
            h = len(lat) // 2  # a p99 that holds over the run, not one that grows with it
            res[name] = dict(p50_ms=pct(lat, 50), p99_ms=pct(lat, 99), qps=st["qps"], launches=st["batches"],
                             max_in_flight=st["max_in_flight"], p99_first_half_ms=pct(lat[:h], 99),
                             p99_second_half_ms=pct(lat[h:], 99), requests_per_launch=len(lat) / max(st["batches"], 1),
                             device_busy_frac=st["device_busy_s"] / max(st["wall_s"], 1e-9),
                             lane_overlap=st["lane_busy_s"] / max(st["device_busy_s"], 1e-9),
                             device_ms_per_launch=st["lane_busy_s"] * 1000 / max(st["batches"], 1),
                             host_submit_ms_per_launch=st["host_submit_s"] * 1000 / max(st["batches"], 1),
                             host_retire_ms_per_launch=st["host_retire_s"] * 1000 / max(st["batches"], 1),
                             host_deliver_ms_per_launch=st["host_deliver_s"] * 1000 / max(st["batches"], 1),
                             engine_busy_frac=(st["host_submit_s"] + st["host_retire_s"] + st["host_deliver_s"])
                             / max(st["wall_s"], 1e-9))
        sv = res["served"]
        log(f"[served] {qps} q/s offered: served p50 {sv['p50_ms']:.2f} p99 {sv['p99_ms']:.2f} "
            f"ms ({sv['qps']:.0f} q/s; device busy {sv['device_busy_frac']:.0%}, "
            f"{sv['device_ms_per_launch']:.2f} ms device / {sv['host_submit_ms_per_launch']:.3f} ms submit / "
            f"{sv['host_retire_ms_per_launch']:.3f} ms retire per launch), plain p50 {res['plain']['p50_ms']:.2f} p99 "
            f"{res['plain']['p99_ms']:.2f} ms ({res['plain']['qps']:.0f} q/s)")
        out.append(res)
    served.close()
    plain.close()
    return dict(table_rows=n, sweep=out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--rows", type=int, default=100_000_000)
    ap.add_argument("--group-hint", type=int, default=6_000_000)
    ap.add_argument("--strategy", type=int, default=-1)
    ap.add_argument("--part-lds-slots", type=int, default=0,
                    help="LDS table slots of the partition aggregate (default: the library's 2048)")
    ap.add_argument("--part-block", type=int, default=0,
                    help="workgroup size of the partition aggregate (default: the library's 1024)")
    ap.add_argument("--hot-slots", type=int, default=-1,
                    help="heavy-hitter slots of the partitioned consume (default: the library's 512)")
    ap.add_argument("--cpu-rows", type=int, default=100_000_000)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--join-build", type=int, default=10_000_000)
    ap.add_argument("--join-probe", type=int, default=200_000_000)
    ap.add_argument("--batch-queries", type=int, default=256)
    ap.add_argument("--batch-tuning", default="",
                    help="block,lds_threshold,tile_rows,max_fanout,lds_cache: skip the autotuner (profiling runs)")
    ap.add_argument("--batch-no-plain", action="store_true",
                    help="skip the one-query-per-launch comparison (profiling runs)")
    ap.add_argument("--serve-rows", type=int, default=500_000_000)
    ap.add_argument("--serve-requests", type=int, default=10000)
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--qps", type=float, nargs="*", default=[100, 1000, 2500, 5000])
    ap.add_argument("--shuffle-parts", type=int, default=256)
    ap.add_argument("--only", choices=["agg", "join", "join_output", "batched", "served", "shuffle", "wide", "parquet"], default=None)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run a device copy of this many MiB (counter calibration under rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (the multi-GPU "
              f"design is described, not built)", file=sys.stderr)
        sys.exit(2)

    from wadjet_amd.gpu import Context
    ctx = Context(0)
    if args.calib_copy_mib:
        # known-size copy for calibrating HBM traffic counters (profiled runs only)
        nb = args.calib_copy_mib << 20
        a, b = ctx.alloc(nb), ctx.alloc(nb)
        ctx.copy(b, a, nb)
        ctx.sync()
        a.free()
        b.free()
    want = (lambda s: args.only in (None, s))
    A = bench_agg(ctx, args) if want("agg") else None
    extra = {}
    if want("join"):
        J = bench_join(ctx, args)
        J.pop("_keep")
        extra["join"] = J
    if want("join_output"):
        extra["join_output"] = bench_join_output(ctx, args)
    if want("shuffle"):
        extra["shuffle"] = bench_shuffle(ctx, args, A["cols"] if A is not None else None)
    if want("wide"):
        extra["wide_key"] = bench_wide(ctx, args)
    if want("parquet"):
        extra["parquet_decode"] = bench_parquet(ctx, args)
    if want("batched"):
        extra["batched"] = bench_batched(ctx, args)
    if want("served"):
        extra["served"] = bench_served(ctx, args)
    cpu = None
    if A is not None and not args.no_cpu:
        cpu = cpu_agg(A["cols"], A["n"], args.cpu_rows, *cpu_threads(args.cpu_threads), args.cpu_reps)

    if A is not None:
        achieved = A["alg_bytes"] / (A["kernel_ms"] / 1000) / 1e9
        from wadjet_amd.profiling import load_traffic
        traffic = load_traffic(str(AGG_PMC_SUMMARY), is_consume_kernel)
        out = {
            "metric": "rows/s & %HBM roofline for hash-agg/join probe; batched vs plain qps + p99",
            "value": A["n"] / (A["ms_step"] / 1000),
            "unit": "rows/s",
            "n_gpus": 1,
            "steps": args.steps,
            "warmup": args.warmup,
            "ms_per_step": A["ms_step"],
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "int64+int128+f64",
            "data": "synthetic (seeded device generators)",
            "config": {"workload": "hash-aggregate, 100M rows: GROUP BY (a int32 Zipf(1.1) over 1M, b int32 "
                                   "uniform 16) SUM(price DECIMAL(15,2)) -> Int128, COUNT(*), MIN/MAX/AVG(f "
                                   "float64) WHERE f > -0.5",
                       "rows": A["n"], "groups": A["groups"], "parallelism": "1 GPU"},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                         "traffic_source": AGG_PMC_SUMMARY.name if traffic else None,
                         "kernels": "consume kernels of one step (direct, or heavy-hitter set + sort level 1 "
                                    "+ scan, sort level 2, partition-aggregate)",
                         "kernel_ms": A["kernel_ms"], "pass_ms": A["passes"], "alg_bytes": A["alg_bytes"],
                         "per_pass": A["per_pass"], "partitioned_rows": A["cold_rows"]},
            "cpu_baseline": cpu,
            "parity": {"tests": "tests/test_fullsize_gpu.py (this workload at 100M rows and the join at 10M x 200M, "
                                "against the oracle)",
                       "exact": "group keys, COUNT, MIN/MAX, Int128 DECIMAL sums, integer sums, the join's row count",
                       "float_tolerance": "float64 SUM/AVG within 1e-12 x sum(|v|) of the group (x 1/count for AVG)"},
        }
    else:
        out = {"metric": "rows/s & %HBM roofline for hash-agg/join probe; batched vs plain qps + p99",
               "value": None, "unit": "rows/s", "n_gpus": 1}
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
