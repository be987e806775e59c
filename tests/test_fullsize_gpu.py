"""Parity at the timed sizes: the two headline workloads run exactly as the
benchmark runs them, compared with the oracle on the same rows.

* GROUP BY (a int32 Zipf(1.1) over 1M, b int32 uniform 16) SUM(price
  DECIMAL(15,2)) -> Int128, COUNT(*), MIN/MAX/AVG(f) WHERE f > -0.5 over
  100M rows (5.7M groups), with the benchmark's tuning (group hint 6M, the
  default strategy and heavy-hitter set): keys, COUNT, MIN/MAX and the Int128
  limbs bit-exact, AVG within 1e-12 of the group's sum of |f| / count.
* JOIN (10M unique build keys, 200M probe rows, 95 % matching) GROUP BY p1
  SUM(qty * p2): the fused probe's 25 Int128 sums and counts == the oracle's
  join -> aggregate on the same rows; the matched-row count == the
  generator's closed form, and the materializing and semi probes return
  that many rows.
The oracle runs on the host's cores (at most 16 threads)."""
import os

import numpy as np
import pytest

import oracle as O
from helpers import compare, fk_match_count, gpu_groups, oracle_groups
from wadjet_amd import _lib as L
from wadjet_amd import workloads as W
from wadjet_amd.gpu import AggSpec, HashAggregate, HashJoin

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def test_zipf_groupby_100M_bench_tuning(ctx):
    n = 100_000_000
    print(f"[full] Zipf GROUP BY {n} rows, oracle on {THREADS} threads", flush=True)
    cols = W.gen_zipf_groupby(ctx, n)
    agg = HashAggregate(ctx, W.ZIPF_TYPES, W.ZIPF_KEYS, W.ZIPF_AGGS, scales=W.ZIPF_SCALES, group_hint=6_000_000)
    agg.consume(cols, n, pred=W.ZIPF_PRED)
    agg.finalize()
    assert agg.pass_times()[2] > 0  # the partitioned passes ran, as in the benchmark
    got = gpu_groups(agg)
    print(f"[full] GPU: {got['n']} groups", flush=True)
    a = cols[0].data.to_host(np.int32, n)
    b = cols[1].data.to_host(np.int32, n)
    p = cols[2].data.to_host(np.int64, n)
    f = cols[3].data.to_host(np.float64, n)
    hcols = [(a, O.INT32), (b, O.INT32), (p, O.DECIMAL, None, 2), (f, O.FLOAT64), (np.abs(f), O.FLOAT64)]
    aggs = [(x.func, x.col) for x in W.ZIPF_AGGS] + [(O.SUM, 4)]
    ref = oracle_groups(O.hash_agg(hcols, W.ZIPF_KEYS, aggs, pred=(3, O.GT, 0, -0.5), threads=THREADS,
                                   max_groups=8_000_000))
    print(f"[full] oracle: {ref['n']} groups", flush=True)
    abs_f = ref["acc"][:, 5]["sum_f64"]
    ref["acc"] = ref["acc"][:, :5]
    compare(got, ref, [(x.func, x.col) for x in W.ZIPF_AGGS], W.ZIPF_TYPES, W.ZIPF_SCALES,
            abs_sums=[None, None, None, None, abs_f])
    agg.close()


def test_fk_join_aggregate_10M_x_200M(ctx):
    cfg = W.FkJoin()
    print(f"[full] fk join {cfg.n_build} x {cfg.n_probe}, oracle on {THREADS} threads", flush=True)
    bcols = W.gen_fk_join_build(ctx, cfg)
    pcols = W.gen_fk_join_probe(ctx, cfg)
    j = HashJoin(ctx, W.FKJ_BUILD_TYPES, 0, inline=[1, 2], scales=W.FKJ_BUILD_SCALES, build_hint=cfg.n_build)
    j.build(bcols, cfg.n_build)
    j.finish()
    types = W.FKJ_PROBE_TYPES + W.FKJ_BUILD_TYPES
    agg = HashAggregate(ctx, types, [3], [AggSpec(L.SUM, 1, 4), AggSpec(L.COUNT, -1)], scales=[0, 0, 0, 0, 2],
                        group_hint=64)
    j.probe_agg(pcols, 0, cfg.n_probe, agg)
    agg.finalize()
    r = agg.fetch()
    st = agg.fetch_state(0)
    o = np.argsort(r["keys"][0])
    gk, gcnt = r["keys"][0][o].astype(np.int64), r["values"][1][o]
    glo, ghi = st["sum_lo"][o], st["sum_hi"][o]
    matched = fk_match_count(cfg.n_probe, cfg.seed + 10, cfg.match_permille)
    print(f"[full] GPU: {r['n']} groups, {int(gcnt.sum())} joined rows; closed form {matched}", flush=True)
    assert int(gcnt.sum()) == matched

    fk = pcols[0].data.to_host(np.int64, cfg.n_probe)
    qty = pcols[1].data.to_host(np.int32, cfg.n_probe)
    key = bcols[0].data.to_host(np.int64, cfg.n_build)
    p1 = bcols[1].data.to_host(np.int32, cfg.n_build)
    p2 = bcols[2].data.to_host(np.int64, cfg.n_build)
    ref = O.join_agg((key, O.INT64), (p1, O.INT32), (p2, O.DECIMAL, None, 2), (fk, O.INT64), (qty, O.INT32),
                     threads=THREADS)
    print(f"[full] oracle: {ref['n']} groups", flush=True)
    assert r["n"] == ref["n"] == 25
    np.testing.assert_array_equal(gk, ref["keys"])
    np.testing.assert_array_equal(gcnt, ref["count"])
    np.testing.assert_array_equal(glo, ref["sum_lo"])
    np.testing.assert_array_equal(ghi, ref["sum_hi"])
    del fk, qty

    # the materializing probe (one call, room for every pair) and the semi probe
    cap = cfg.n_probe + 1024
    outp, outb = ctx.alloc(cap * 4), ctx.alloc(cap * 4)
    got, cur = j.probe(pcols[0], cfg.n_probe, outp, outb, cap)
    assert cur.done == 1 and got == matched
    assert j.semi(pcols[0], cfg.n_probe, outp) == matched
    assert j.semi(pcols[0], cfg.n_probe, outp, anti=True) == cfg.n_probe - matched
    outp.free()
    outb.free()
    agg.close()
    j.close()
