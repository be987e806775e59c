"""Served form: a request loop over a device-resident flow table with
continuous batching. Every request's result is identical whether it ran in a
batch with others (several lanes in flight) or alone (one request per launch),
and equals the oracle's."""
import numpy as np
import pytest

import oracle as O
from wadjet_amd import _lib as L
from wadjet_amd import workloads as W
from wadjet_amd.gpu import HashJoin, Server

pytestmark = pytest.mark.gpu


def oracle_request(host, dim_ip, dim_asn, shape, q):
    src, port, nbytes, ts = host
    lo, hi = q.ts_lo, q.ts_hi
    rows = np.nonzero((ts >= lo) & (ts < hi))[0].astype(np.uint32)
    if shape["kind"] == 0:
        keycol = (port, O.UINT16) if q.key_col == 1 else (src, O.UINT32)
        r = O.hash_agg([keycol, (nbytes, O.DECIMAL)], [0], [(O.SUM, 1), (O.COUNT, -1)], sel=rows)
        keys = r["keys"][:, 0]
        acc = r["acc"]
    else:
        p, b = O.hash_join((dim_ip, O.UINT32), (src, O.UINT32), probe_sel=rows, cap=len(rows) + 1)
        r = O.hash_agg([(dim_asn[b], O.INT32), (nbytes[p], O.DECIMAL)], [0], [(O.SUM, 1), (O.COUNT, -1)])
        keys = r["keys"][:, 0]
        acc = r["acc"]
    o = np.argsort(keys, kind="stable")
    lo_ = acc[o, 0]["sum_lo"]
    return dict(n=r["n"], keys=keys[o], sum_lo=lo_, sum_hi=np.where(lo_.view(np.int64) < 0, -1, 0),
                counts=acc[o, 1]["count"])


def same(a, b, i):
    assert a["n"] == b["n"], f"request {i}: {a['n']} vs {b['n']} groups"
    for f in ("keys", "sum_lo", "sum_hi", "counts"):
        np.testing.assert_array_equal(np.asarray(a[f]).astype(np.int64), np.asarray(b[f]).astype(np.int64),
                                      err_msg=f"request {i}: {f}")


def test_served_equals_alone_equals_oracle(ctx):
    n = 6_000_000
    cols = W.gen_flow_table(ctx, n)
    dim = W.gen_flow_dim(ctx, 50_000)
    j = HashJoin(ctx, W.FLOW_DIM_TYPES, 0, inline=[1], build_hint=50_000)
    j.build(dim, 50_000)
    j.finish()
    qs, shapes = W.plan_flow_requests(n, 80, seed=3, scale=0.1)
    print(f"[serve] {len(qs)} requests over {n} rows", flush=True)
    served = Server(ctx, cols, n, 3, lanes=3, hbm_budget=1 << 30, max_batch=16)
    served.add_dim(j)
    lat_s, ids_s = served.replay(qs, qps=3000, seed=1)
    st = served.stats()
    print(f"[serve] served stats {st}", flush=True)
    plain = Server(ctx, cols, n, 3, plain=True)
    plain.add_dim(j)
    lat_p, ids_p = plain.replay(qs, qps=3000, seed=1)
    print(f"[serve] plain stats {plain.stats()}", flush=True)
    assert st["completed"] == len(qs)
    host = [c.data.to_host({L.UINT32: np.uint32, L.UINT16: np.uint16, L.INT64: np.int64}[c.type], n) for c in cols]
    dim_ip = dim[0].data.to_host(np.uint32, 50_000)
    dim_asn = dim[1].data.to_host(np.int32, 50_000)
    for i in range(len(qs)):
        a = served.result(int(ids_s[i]))
        b = plain.result(int(ids_p[i]))
        same(a, b, i)
        if i % 8 == 0:
            same(a, oracle_request(host, dim_ip, dim_asn, shapes[i], qs[i]), i)
    served.close()
    plain.close()


def test_results_after_reset_reuse_buffers(ctx):
    """A reset returns the requests' result buffers to the server's pool; the
    next runs deliver into those buffers (in another order, so sizes differ
    from the first owner's) and every result equals the first run's."""
    n = 2_000_000
    cols = W.gen_flow_table(ctx, n, seed=11)
    dim = W.gen_flow_dim(ctx, 20_000)
    j = HashJoin(ctx, W.FLOW_DIM_TYPES, 0, inline=[1], build_hint=20_000)
    j.build(dim, 20_000)
    j.finish()
    qs, _ = W.plan_flow_requests(n, 40, seed=8, scale=0.05)
    print(f"[serve] {len(qs)} requests, three runs with resets", flush=True)
    s = Server(ctx, cols, n, 3, lanes=3, hbm_budget=1 << 30, max_batch=8)
    s.add_dim(j)
    _, ids = s.replay(qs, qps=3000, seed=1)
    first = [s.result(int(i)) for i in ids]
    for rep in range(2):
        s.reset()
        order = list(range(len(qs)))[::-1] if rep == 0 else list(np.random.default_rng(rep).permutation(len(qs)))
        _, ids2 = s.replay([qs[k] for k in order], qps=3000, seed=2 + rep)
        for pos, k in enumerate(order):
            same(s.result(int(ids2[pos])), first[k], k)
    s.close()


def test_admission_respects_the_budget(ctx):
    """A tiny HBM budget forces requests through one at a time; results stay exact."""
    n = 1_000_000
    cols = W.gen_flow_table(ctx, n, seed=9)
    qs, shapes = W.plan_flow_requests(n, 12, seed=4, scale=0.05)
    qs = [q for q, s in zip(qs, shapes) if s["kind"] == 0]
    s1 = Server(ctx, cols, n, 3, lanes=2, hbm_budget=1, max_batch=8)
    lat, ids = s1.replay(qs, qps=5000, seed=2)
    st = s1.stats()
    assert st["completed"] == len(qs)
    assert st["batches"] == len(qs)  # every request alone: each exceeds the budget
    host = [c.data.to_host({L.UINT32: np.uint32, L.UINT16: np.uint16, L.INT64: np.int64}[c.type], n) for c in cols]
    for i, q in enumerate(qs):
        same(s1.result(int(ids[i])), oracle_request(host, None, None, dict(kind=0), q), i)
    s1.close()
