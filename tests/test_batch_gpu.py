"""Batched form: many GROUP BY queries of uneven shapes in one persistent,
graph-replayed launch give, per query, exactly the plain one-query-per-launch
result and the oracle's; the autotuner's choices keep them identical."""
import numpy as np
import pytest

import oracle as O
from wadjet_amd import _lib as L
from wadjet_amd import workloads as W
from wadjet_amd.gpu import Batch, Col

pytestmark = pytest.mark.gpu

NP = {L.INT32: np.int32, L.INT64: np.int64, L.UINT32: np.uint32, L.DECIMAL: np.int64}


def oracle_query(kc, vc, n, extras=None):
    """GROUP BY key SUM(val), COUNT(*) on the oracle; a join query first maps
    each key to the payload of its dimension row (unique keys: at most one)
    and drops the keys without one."""
    k = kc.data.to_host(NP[kc.type], n)
    v = vc.data.to_host(np.int64, n)
    ktype = kc.type
    if extras and "dim" in extras:
        dk = extras["dim_cols"][0].data.to_host(np.int64, W.QMIX_DIM_ROWS)
        dp = extras["dim_cols"][1 + extras["payload"]].data.to_host(np.int32, W.QMIX_DIM_ROWS)
        p, b = O.hash_join((dk, O.INT64), (k, O.INT64), cap=n + 1)
        k, v, ktype = dp[b].astype(np.int64), v[p], O.INT64
    r = O.hash_agg([(k, ktype), (v, O.DECIMAL)], [0], [(O.SUM, 1), (O.COUNT, -1)])
    o = np.argsort(r["keys"][:, 0], kind="stable")
    acc = r["acc"][o]
    lo = acc[:, 0]["sum_lo"].copy()
    hi = acc[:, 0]["sum_hi"].copy()
    if vc.type != L.DECIMAL:  # SUM(int64) wraps at 64 bits
        hi = np.where(lo.view(np.int64) < 0, -1, 0)
    return dict(n=r["n"], keys=r["keys"][o, 0], sum_lo=lo, sum_hi=hi, counts=acc[:, 1]["count"])


def same(a, b, q):
    assert a["n"] == b["n"], f"query {q}: {a['n']} vs {b['n']} groups"
    for f in ("keys", "sum_lo", "sum_hi", "counts"):
        np.testing.assert_array_equal(np.asarray(a[f]).astype(np.int64), np.asarray(b[f]).astype(np.int64),
                                      err_msg=f"query {q}: {f}")


def test_batched_equals_plain_equals_oracle(ctx):
    plan = W.plan_query_mix(n_queries=40, seed=5, max_rows=300_000, max_groups=20_000)
    data = W.gen_query_mix(ctx, plan)
    qarr = Batch.queries(data)
    b = Batch(ctx, 64, sum(q["groups"] for q in plan))
    print(f"[batch] {len(plan)} queries, {sum(q['rows'] for q in plan)} rows", flush=True)
    b.run(qarr, len(plan), use_graph=True)
    batched = [b.result(q) for q in range(len(plan))]
    b.run(qarr, len(plan), use_graph=True)  # graph replay
    replay = [b.result(q) for q in range(len(plan))]
    b.run(qarr, len(plan), use_graph=False)
    direct = [b.result(q) for q in range(len(plan))]
    b.run(qarr, len(plan), plain=True)
    plain = [b.result(q) for q in range(len(plan))]
    for q, (kc, vc, n, g, *x) in enumerate(data):
        ref = oracle_query(kc, vc, n, x[0] if x else None)
        same(batched[q], ref, q)
        same(replay[q], ref, q)
        same(direct[q], ref, q)
        same(plain[q], ref, q)


def test_graph_replay_after_the_work_queue_grows(ctx):
    """A replayed graph must never see work-queue buffers that a bigger batch
    reallocated: small batch, then a batch with many more tiles, both graphed."""
    b = Batch(ctx, 96, 200_000)
    for n_q, seed in ((4, 11), (80, 12), (8, 13)):
        plan = W.plan_query_mix(n_queries=n_q, seed=seed, max_rows=200_000, max_groups=2_000)
        data = W.gen_query_mix(ctx, plan)
        b.run(Batch.queries(data), len(plan), use_graph=True)
        for q, (kc, vc, n, g, *x) in enumerate(data):
            same(b.result(q), oracle_query(kc, vc, n, x[0] if x else None), q)


def test_autotune_keeps_results(ctx):
    plan = W.plan_query_mix(n_queries=24, seed=6, max_rows=200_000, max_groups=50_000)
    data = W.gen_query_mix(ctx, plan)
    qarr = Batch.queries(data)
    b = Batch(ctx, 32, sum(q["groups"] for q in plan))
    b.autotune(qarr, len(plan))
    t = b.tuning()
    print(f"[batch] autotuned {t}", flush=True)
    assert t["block"] in (256, 512, 1024)
    b.run(qarr, len(plan))
    for q, (kc, vc, n, g, *x) in enumerate(data):
        use_lds, blk, fan = b.choice(q)
        assert blk == t["block"]
        assert fan <= t["max_fanout"]
        est = max(g, 8)
        fits = t["lds_threshold"] > 0 and est // fan <= t["lds_threshold"]
        assert use_lds == (1 if fits else 2 if t["lds_cache"] and t["lds_threshold"] > 0 else 0)
        same(b.result(q), oracle_query(kc, vc, n, x[0] if x else None), q)


def test_underestimated_cardinality_falls_back_exactly(ctx):
    plan = W.plan_query_mix(n_queries=6, seed=8, max_rows=100_000, min_groups=5000, max_groups=20_000)
    data = W.gen_query_mix(ctx, plan)
    lying = [(k, v, n, 10, *x) for (k, v, n, g, *x) in data]  # estimate 10 groups
    b = Batch(ctx, 8, 1000)
    b.run(Batch.queries(lying), len(lying))
    for q, (kc, vc, n, g, *x) in enumerate(data):
        same(b.result(q), oracle_query(kc, vc, n, x[0] if x else None), q)


@pytest.mark.parametrize("tuning", [(512, 1024, 8192, 1, 0), (1024, 256, 32768, 16, 1), (256, 2048, 8192, 4, 0),
                                    (512, 0, 16384, 1, 0), (1024, 512, 65536, 64, 1)])
def test_mixed_groupby_and_join_batches_under_every_layout(ctx, tuning):
    """GROUP BY and JOIN dim GROUP BY payload queries in one launch, under
    tunings that put queries on LDS tables, on global tables and on
    key-partitioned work items (fan-out): every result == the oracle's, and
    the fan-out is taken where the tuning allows it."""
    plan = W.plan_query_mix(n_queries=32, seed=21, max_rows=400_000, max_groups=60_000, join_every=3)
    data = W.gen_query_mix(ctx, plan)
    assert sum(1 for q in plan if q["kind"] == "join") >= 10
    b = Batch(ctx, 64, sum(q["groups"] for q in plan))
    b.set_tuning(*tuning)
    print(f"[batch] mixed GROUP BY / join, tuning {tuning}", flush=True)
    for graph in (True, False):
        b.run(Batch.queries(data), len(plan), use_graph=graph)
        fans = []
        for q, (kc, vc, n, g, *x) in enumerate(data):
            same(b.result(q), oracle_query(kc, vc, n, x[0] if x else None), q)
            fans.append(b.choice(q)[2])
        if tuning[1] and tuning[3] > 1:
            assert max(fans) > 1  # some query was split over key partitions
        if tuning[1] and tuning[4]:
            assert any(b.choice(q)[0] == 2 for q in range(len(plan)))  # some query behind the LDS cache
        assert all(f == 1 for f, q in zip(fans, plan) if q["kind"] == "join")


def test_skewed_high_cardinality_behind_the_lds_cache(ctx):
    """Zipf-keyed queries above the LDS threshold (a hot key carrying a large
    share of the rows) through the LDS cache and without it: == the oracle."""
    plan = [dict(rows=r, groups=g, key_type=L.UINT32, val_type=v, seed=900 + i, kind="groupby")
            for i, (r, g, v) in enumerate([(600_000, 200_000, L.INT64), (1_500_000, 400_000, L.DECIMAL),
                                           (90_000, 50_000, L.INT64)])]
    data = []
    for q in plan:
        k = ctx.gen(L.GEN_ZIPF_U32, "uint32", q["rows"], q["seed"], 1_000_000, 0, 1.2)
        v = ctx.gen(L.GEN_UNIFORM_I64, "int64", q["rows"], q["seed"] + 1, -10**9, 10**9)
        data.append((Col(k, L.UINT32), Col(v, q["val_type"], None, 2 if q["val_type"] == L.DECIMAL else 0),
                     q["rows"], q["groups"]))
    b = Batch(ctx, 8, 1_000_000)
    for cache in (1, 0):
        b.set_tuning(1024, 2048, 32768, 16, cache)
        b.run(Batch.queries(data), len(data))
        for q, (kc, vc, n, g) in enumerate(data):
            assert b.choice(q)[0] == (2 if cache else 0)
            same(b.result(q), oracle_query(kc, vc, n), q)
