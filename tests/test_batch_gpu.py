"""Batched form: many GROUP BY queries of uneven shapes in one persistent,
graph-replayed launch give, per query, exactly the plain one-query-per-launch
result and the oracle's; the autotuner's choices keep them identical."""
import numpy as np
import pytest

import oracle as O
from wadjet_amd import _lib as L
from wadjet_amd import workloads as W
from wadjet_amd.gpu import Batch

pytestmark = pytest.mark.gpu

NP = {L.INT32: np.int32, L.INT64: np.int64, L.UINT32: np.uint32, L.DECIMAL: np.int64}


def oracle_query(kc, vc, n):
    k = kc.data.to_host(NP[kc.type], n)
    v = vc.data.to_host(np.int64, n)
    r = O.hash_agg([(k, kc.type), (v, O.DECIMAL)], [0], [(O.SUM, 1), (O.COUNT, -1)])
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
    for q, (kc, vc, n, g) in enumerate(data):
        ref = oracle_query(kc, vc, n)
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
        for q, (kc, vc, n, g) in enumerate(data):
            same(b.result(q), oracle_query(kc, vc, n), q)


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
    for q, (kc, vc, n, g) in enumerate(data):
        use_lds, blk = b.choice(q)
        assert blk == t["block"]
        assert use_lds == (g <= t["lds_threshold"])
        same(b.result(q), oracle_query(kc, vc, n), q)


def test_underestimated_cardinality_falls_back_exactly(ctx):
    plan = W.plan_query_mix(n_queries=6, seed=8, max_rows=100_000, min_groups=5000, max_groups=20_000)
    data = W.gen_query_mix(ctx, plan)
    lying = [(k, v, n, 10) for (k, v, n, g) in data]  # estimate 10 groups
    b = Batch(ctx, 8, 1000)
    b.run(Batch.queries(lying), len(lying))
    for q, (kc, vc, n, g) in enumerate(data):
        same(b.result(q), oracle_query(kc, vc, n), q)
