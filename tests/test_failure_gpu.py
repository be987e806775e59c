"""Failure handling under injected device-allocation failures
(wg_ctx_inject_alloc_failure), the paths the reference's own tests force with
a small memory.Tracker budget and FatalEvalPanic (pipeline.go:45-57).

For every allocation a call sequence makes, that allocation is made to fail:
the failing call must return WG_ERR_OOM and the handles stay valid. An
aggregate then either holds exactly the input of the calls that succeeded —
the failed call is retried and the result equals the oracle's — or, when the
call failed after it had changed the aggregate, it says so and refuses every
call but reset, after which the whole input gives the oracle's result again.
Each case is run with one failing allocation and with every allocation from
the failing one on failing (sticky), as under real exhaustion."""
import numpy as np
import pytest

import oracle as O
from helpers import compare, gpu_groups, oracle_groups
from wadjet_amd import _lib as L
from wadjet_amd.gpu import AggSpec, Batch, Col, Context, HashAggregate, HashJoin

pytestmark = pytest.mark.gpu

AGGS = [AggSpec(L.SUM, 1), AggSpec(L.COUNT, -1), AggSpec(L.MIN, 1), AggSpec(L.MAX, 1)]
OAGGS = [(O.SUM, 1), (O.COUNT, -1), (O.MIN, 1), (O.MAX, 1)]
TYPES = [L.INT64, L.INT64]
POISONED = "holds part of this call's input"


@pytest.fixture(scope="module")
def fctx():
    """A context of its own: injected failures never reach the shared one."""
    c = Context(0)
    yield c
    c.inject_alloc_failure(0)
    c.close()


def batches(seed):
    rng = np.random.default_rng(seed)
    out = []
    for n, groups in ((40_000, 30_000), (60_000, 200_000)):
        k = rng.integers(0, groups, n).astype(np.int64)
        v = rng.integers(-10**12, 10**12, n).astype(np.int64)
        out.append((k, v))
    return out


def expected(bs):
    k = np.concatenate([b[0] for b in bs])
    v = np.concatenate([b[1] for b in bs])
    return oracle_groups(O.hash_agg([(k, O.INT64), (v, O.INT64)], [0], OAGGS))


def sweep(fctx, make, steps, verify, max_k=200):
    """Fail allocation k = 1, 2, ... of the step sequence until it completes
    with no failure left to inject. Returns (clean retries, poisoned)."""
    tally = {"clean": 0, "poisoned": 0}
    for sticky in (False, True):
        for k in range(1, max_k + 1):
            print(f"[fail] sticky={sticky} allocation {k}", flush=True)
            h = make()
            fctx.inject_alloc_failure(k, sticky)
            failed, err = None, None
            try:
                for i, step in enumerate(steps):
                    failed = i
                    step(h)
                failed = None
            except L.WgError as e:
                err = e
            finally:
                fctx.inject_alloc_failure(0)
            if failed is None:
                verify(h)
                break
            print(f"[fail]   step {failed}: {err}", flush=True)
            assert err.status == L.WG_ERR_OOM, f"step {failed}: {err}"
            if POISONED in str(err):
                tally["poisoned"] += 1
                with pytest.raises(L.WgError) as again:  # the failed call, repeated, is refused
                    steps[failed](h)
                assert again.value.status == L.WG_ERR_STATE
                h.reset()
                for step in steps:
                    step(h)
            else:
                tally["clean"] += 1
                for step in steps[failed:]:
                    step(h)
            verify(h)
            h.close()
        else:
            pytest.fail(f"more than {max_k} allocations in the sequence")
    print(f"[fail] {tally}", flush=True)
    return tally


@pytest.mark.parametrize("mode", ["direct", "partitioned", "host"])
def test_aggregate_survives_every_allocation_failure(fctx, mode):
    bs = batches(3)
    ref = expected(bs)
    dev = [[Col(fctx.to_device(k), L.INT64), Col(fctx.to_device(v), L.INT64)] for k, v in bs]
    host = [[Col(k, L.INT64), Col(v, L.INT64)] for k, v in bs]

    class Agg:  # the aggregate and the state a retry needs
        def __init__(self):
            self.a = HashAggregate(fctx, TYPES, [0], AGGS, group_hint=1024)
            if mode == "direct":
                self.a.set_strategy(1)
            elif mode == "partitioned":
                self.a.set_strategy(2, 256, 256)

        def reset(self):
            self.a.reset()

        def close(self):
            self.a.close()

    def consume(i):
        def step(h):
            if mode == "host":
                h.a.consume(host[i], len(bs[i][0]), device=False)
            else:
                h.a.consume(dev[i], len(bs[i][0]))
        return step

    def verify(h):
        compare(gpu_groups(h.a), ref, OAGGS, TYPES)

    t = sweep(fctx, Agg, [consume(0), consume(1), lambda h: h.a.finalize(), lambda h: h.a.fetch()], verify)
    assert t["clean"] + t["poisoned"] >= 4


def test_fused_join_aggregate_survives_every_allocation_failure(fctx):
    """Build, probe fused into a GROUP BY that must grow mid-probe (duplicate
    build keys, a group hint of 1), finalize; against the oracle's join."""
    rng = np.random.default_rng(21)
    nb, npr = 30_000, 120_000
    bk = rng.integers(0, 10_000, nb).astype(np.int64)
    bv = rng.integers(-10**9, 10**9, nb).astype(np.int64)
    pk = rng.integers(0, 12_000, npr).astype(np.int64)
    pv = rng.integers(0, 100, npr).astype(np.int32)
    bcols = [Col(fctx.to_device(bk), L.INT64), Col(fctx.to_device(bv), L.INT64)]
    pcols = [Col(fctx.to_device(pk), L.INT64), Col(fctx.to_device(pv), L.INT32)]
    types = [L.INT64, L.INT32, L.INT64, L.INT64]
    aggs = [AggSpec(L.SUM, 3), AggSpec(L.COUNT, -1), AggSpec(L.MAX, 1)]
    oaggs = [(O.SUM, 3), (O.COUNT, -1), (O.MAX, 1)]
    pp, bb = O.hash_join((bk, O.INT64), (pk, O.INT64), cap=npr * 20)
    jcols = [(pk[pp], O.INT64), (pv[pp], O.INT32), (bk[bb], O.INT64), (bv[bb], O.INT64)]
    ref = oracle_groups(O.hash_agg(jcols, [3], oaggs))

    class JA:
        def __init__(self):
            self.j = HashJoin(fctx, [L.INT64, L.INT64], 0, build_hint=1024)
            self.a = HashAggregate(fctx, types, [3], aggs, group_hint=1)
            self.built = False

        def reset(self):
            self.a.reset()

        def close(self):
            self.a.close()
            self.j.close()

    def build(h):
        if not h.built:  # a build append that failed added no row: it is simply repeated
            h.j.build(bcols, nb)
            h.j.finish()
            h.built = True

    def verify(h):
        compare(gpu_groups(h.a), ref, oaggs, types)

    steps = [build, lambda h: h.j.probe_agg(pcols, 0, npr, h.a), lambda h: h.a.finalize(), lambda h: h.a.fetch()]
    t = sweep(fctx, JA, steps, verify)
    assert t["clean"] + t["poisoned"] >= 4


def test_batch_survives_every_allocation_failure(fctx):
    """The batched runner: a failed run leaves the runner usable, and the
    next run gives the same groups as an undisturbed one."""
    from wadjet_amd import workloads as W
    plan = W.plan_query_mix(n_queries=12, seed=9, max_rows=200_000, join_every=0)
    data = W.gen_query_mix(fctx, plan)
    qs = Batch.queries(data)
    ref_b = Batch(fctx, 12, sum(q["groups"] for q in plan))
    ref_b.run(qs, len(plan))
    want = [ref_b.result(i) for i in range(len(plan))]
    ref_b.close()

    class B:
        def __init__(self):
            self.b = Batch(fctx, 12, sum(q["groups"] for q in plan))

        def reset(self):
            pass

        def close(self):
            self.b.close()

    def verify(h):
        for i in range(len(plan)):
            got = h.b.result(i)
            for f in ("n", "keys", "sum_lo", "sum_hi", "counts"):
                np.testing.assert_array_equal(got[f], want[i][f], err_msg=f"query {i} {f}")

    steps = [lambda h: h.b.run(qs, len(plan)), lambda h: h.b.run(qs, len(plan), use_graph=True)]
    t = sweep(fctx, B, steps, verify)
    assert t["poisoned"] == 0 and t["clean"] >= 2


def test_create_failures_leave_nothing_behind(fctx):
    """A create whose allocation fails returns WG_ERR_OOM and no handle."""
    for k in range(1, 6):
        fctx.inject_alloc_failure(k)
        try:
            for make in (lambda: HashAggregate(fctx, TYPES, [0], AGGS, group_hint=1 << 16),
                         lambda: HashJoin(fctx, [L.INT64, L.INT64], 0, build_hint=1 << 16)):
                try:
                    make().close()
                except L.WgError as e:
                    assert e.status == L.WG_ERR_OOM, str(e)
        finally:
            fctx.inject_alloc_failure(0)
    a = HashAggregate(fctx, TYPES, [0], AGGS, group_hint=1 << 16)  # the context is unharmed
    a.close()
