"""GPU parity of the hash join (wg_join_*) against the CPU oracle: the joined
row multiset (pairs and their digest), the bounded/resumable probe, NULL keys,
duplicate build keys, and the fused probe -> GROUP BY with exact Int128 sums."""
import numpy as np
import pytest

import oracle as O
from helpers import compare, gpu_groups, oracle_groups
from wadjet_amd import _lib as L
from wadjet_amd import workloads as W
from wadjet_amd.gpu import AggSpec, Col, HashAggregate, HashJoin, bitmap_from_mask

pytestmark = pytest.mark.gpu


def probe_all(j, key, n, cap, total_cap):
    ctx = j.ctx
    out_p = ctx.alloc(total_cap * 4)
    out_b = ctx.alloc(total_cap * 4)
    tp = ctx.alloc(cap * 4)
    tb = ctx.alloc(cap * 4)
    cur = L.ProbeCursor(0, -1, 0, 0)
    ps, bs = [], []
    calls = 0
    while not cur.done:
        k, cur = j.probe(key, n, tp, tb, cap, cur)
        assert k <= cap
        ps.append(tp.to_host(np.uint32, k))
        bs.append(tb.to_host(np.uint32, k))
        calls += 1
        assert calls < total_cap // cap * 4 + 1000
    return np.concatenate(ps), np.concatenate(bs), calls


def pairs_set(p, b):
    return np.sort(p.astype(np.uint64) << np.uint64(32) | b.astype(np.uint64))


@pytest.mark.parametrize("cap", [32768, 1000, 97])
def test_probe_multiset_with_duplicates_and_nulls(ctx, cap):
    """Duplicate build keys (chains), NULL keys on both sides, int32 probe
    key against an int64 build key; the bounded probe resumes mid-chain."""
    print(f"[join] duplicates/nulls cap={cap}", flush=True)
    rng = np.random.default_rng(cap)
    nb, npr = 20_000, 50_000
    bk = rng.integers(0, 8000, nb).astype(np.int64)
    bk[:300] = 77  # one chain longer than the smallest output bound
    bm = rng.random(nb) > 0.02
    pk = rng.integers(0, 12_000, npr).astype(np.int32)
    pk[::100] = 77
    pm = rng.random(npr) > 0.02
    j = HashJoin(ctx, [L.INT64, L.INT32], 0, inline=[1], build_hint=1000)
    half = nb // 2
    for lo, hi in ((0, half), (half, nb)):  # two build batches, one host one device
        m = bitmap_from_mask(bm[lo:hi])
        if lo == 0:
            j.build([Col(bk[lo:hi], L.INT64, m), Col(np.arange(lo, hi, dtype=np.int32), L.INT32)], hi - lo, device=False)
        else:
            j.build([Col(ctx.to_device(bk[lo:hi]), L.INT64, ctx.to_device(m)),
                     Col(ctx.to_device(np.arange(lo, hi, dtype=np.int32)), L.INT32)], hi - lo)
    j.finish()
    assert j.rows() == nb
    key = Col(ctx.to_device(pk), L.INT32, ctx.to_device(bitmap_from_mask(pm)))
    want_p, want_b = O.hash_join((bk, O.INT64, bitmap_from_mask(bm)), (pk, O.INT32, bitmap_from_mask(pm)),
                                 cap=npr * 60)
    p, b, calls = probe_all(j, key, npr, cap, len(want_p) + 1)
    np.testing.assert_array_equal(pairs_set(p, b), pairs_set(want_p, want_b))
    assert calls >= len(want_p) // cap


def test_fk_join_multiset_and_digest(ctx):
    """Unique permutation keys, 95 % matching foreign keys: pair multiset,
    device digest == oracle digest, matched count closed form."""
    cfg = W.FkJoin(n_build=200_000, n_probe=1_000_000)
    print(f"[join] fk-join shape {cfg}", flush=True)
    bcols = W.gen_fk_join_build(ctx, cfg)
    pcols = W.gen_fk_join_probe(ctx, cfg)
    j = HashJoin(ctx, W.FKJ_BUILD_TYPES, 0, inline=[1, 2], scales=W.FKJ_BUILD_SCALES, build_hint=cfg.n_build)
    j.build(bcols, cfg.n_build)
    j.finish()
    bk = bcols[0].data.to_host(np.int64)
    assert len(np.unique(bk)) == cfg.n_build  # a permutation: unique keys
    fk = pcols[0].data.to_host(np.int64)
    qty = pcols[1].data.to_host(np.int32)
    p1 = bcols[1].data.to_host(np.int32)
    p2 = bcols[2].data.to_host(np.int64)
    want_p, want_b = O.hash_join((bk, O.INT64), (fk, O.INT64), cap=cfg.n_probe)
    p, b, _ = probe_all(j, pcols[0], cfg.n_probe, cfg.n_probe, cfg.n_probe)
    assert len(p) == len(want_p)
    np.testing.assert_array_equal(pairs_set(p, b), pairs_set(want_p, want_b))
    # the matched count is the number of probe rows drawn from the build domain
    assert abs(len(p) / cfg.n_probe - 0.95) < 0.005
    dp, db = ctx.to_device(p), ctx.to_device(b)
    d_gpu = j.digest(pcols, dp, db, len(p), [0, 1], [1, 2])
    rows = np.stack([fk[want_p], qty[want_p].astype(np.int64), p1[want_b].astype(np.int64), p2[want_b]], axis=1)
    assert d_gpu == O.digest_rows(rows)
    # gathers
    out = ctx.alloc(len(b) * 8)
    j.gather_build(2, db, len(b), out)
    np.testing.assert_array_equal(out.to_host(np.int64, len(b)), p2[b])


@pytest.mark.parametrize("lds,lean", [(1024, True), (1024, False), (0, False)])
def test_fk_join_fused_probe_aggregate(ctx, lds, lean):
    """JOIN fk = key GROUP BY p1 SUM(qty * p2) (DECIMAL product -> Int128),
    COUNT(*), MIN(qty), AVG(p2): fused probe vs oracle over the pair list,
    through the streamlined kernel and the general one."""
    cfg = W.FkJoin(n_build=300_000, n_probe=2_000_000, seed=17)
    print(f"[join] fused probe-agg lds={lds} lean={lean} {cfg}", flush=True)
    bcols = W.gen_fk_join_build(ctx, cfg)
    pcols = W.gen_fk_join_probe(ctx, cfg)
    j = HashJoin(ctx, W.FKJ_BUILD_TYPES, 0, inline=[1, 2], scales=W.FKJ_BUILD_SCALES, build_hint=cfg.n_build)
    j.set_tuning(lean=lean)
    j.build(bcols, cfg.n_build)
    j.finish()
    types = W.FKJ_PROBE_TYPES + W.FKJ_BUILD_TYPES  # fk, qty, key, p1, p2
    scales = [0, 0, 0, 0, 2]
    aggs = [AggSpec(L.SUM, 1, 4), AggSpec(L.COUNT, -1), AggSpec(L.MIN, 1), AggSpec(L.AVG, 4)]
    agg = HashAggregate(ctx, types, [3], aggs, scales=scales, group_hint=64)
    agg.set_tuning(lds, 512)
    j.probe_agg(pcols, 0, cfg.n_probe, agg)
    agg.finalize()
    g = gpu_groups(agg)
    # oracle: join pairs, then aggregate the joined rows; the product is a
    # DECIMAL128 column formed exactly in Python integers
    fk = pcols[0].data.to_host(np.int64)
    qty = pcols[1].data.to_host(np.int32)
    bk = bcols[0].data.to_host(np.int64)
    p1 = bcols[1].data.to_host(np.int32)
    p2 = bcols[2].data.to_host(np.int64)
    pp, bb = O.hash_join((bk, O.INT64), (fk, O.INT64), cap=cfg.n_probe)
    prod = qty[pp].astype(np.int64) * p2[bb]  # < 50 * 10^12: exact in int64
    prod128 = np.empty(2 * len(prod), np.uint64)
    prod128[0::2] = np.where(prod < 0, np.uint64(2**64 - 1), np.uint64(0))
    prod128[1::2] = prod.view(np.uint64)
    jcols = [(fk[pp], O.INT64), (qty[pp], O.INT32), (bk[bb], O.INT64), (p1[bb], O.INT32), (p2[bb], O.DECIMAL, None, 2),
             (prod128, O.DECIMAL128, None, 2)]
    ref = oracle_groups(O.hash_agg(jcols, [3], [(O.SUM, 5), (O.COUNT, -1), (O.MIN, 1), (O.AVG, 4)]))
    assert g["n"] == ref["n"] == 25
    types_o = types + [O.DECIMAL128]
    compare(g, ref, [(O.SUM, 5), (O.COUNT, -1), (O.MIN, 1), (O.AVG, 4)], types_o, scales + [2])


def test_fused_probe_aggregate_many_to_many_with_overflow(ctx):
    """Duplicate build keys, high-cardinality groups from a tiny table hint:
    the fused kernel claims every group of a probe row before applying any,
    so replays after a grow stay exact."""
    rng = np.random.default_rng(21)
    nb, npr = 30_000, 200_000
    bk = rng.integers(0, 10_000, nb).astype(np.int64)
    bv = rng.integers(-10**9, 10**9, nb).astype(np.int64)
    pk = rng.integers(0, 12_000, npr).astype(np.int64)
    pv = rng.integers(0, 100, npr).astype(np.int32)
    j = HashJoin(ctx, [L.INT64, L.INT64], 0)
    j.build([Col(ctx.to_device(bk), L.INT64), Col(ctx.to_device(bv), L.INT64)], nb)
    j.finish()
    types = [L.INT64, L.INT32, L.INT64, L.INT64]
    aggs = [AggSpec(L.SUM, 3), AggSpec(L.COUNT, -1), AggSpec(L.MAX, 1)]
    agg = HashAggregate(ctx, types, [3], aggs, group_hint=1)  # GROUP BY the build value
    j.probe_agg([Col(ctx.to_device(pk), L.INT64), Col(ctx.to_device(pv), L.INT32)], 0, npr, agg)
    agg.finalize()
    pp, bb = O.hash_join((bk, O.INT64), (pk, O.INT64), cap=npr * 20)
    jcols = [(pk[pp], O.INT64), (pv[pp], O.INT32), (bk[bb], O.INT64), (bv[bb], O.INT64)]
    ref = oracle_groups(O.hash_agg(jcols, [3], [(O.SUM, 3), (O.COUNT, -1), (O.MAX, 1)]))
    compare(gpu_groups(agg), ref, [(O.SUM, 3), (O.COUNT, -1), (O.MAX, 1)], types)


SENTINEL = -0x7FFFFFFFFFFFFFFF  # the tables' EMPTY key word, a legal key value


@pytest.mark.parametrize("block,replicas,hint", [(512, 0, 40), (1024, 0, 40), (512, 1, 40), (512, 0, 4)])
def test_fused_probe_aggregate_streamlined_edges(ctx, block, replicas, hint):
    """The streamlined fused kernel and its hand-offs to the general one:
    duplicate build keys (chains), NULL probe keys, the sentinel value as a
    probe/build key and as a group key, a selection vector, float and int
    probe values, a group key from an inline build column and from a probe
    column, and a group hint below the real group count (LDS table full)."""
    print(f"[join] streamlined edges block={block} replicas={replicas} hint={hint}", flush=True)
    rng = np.random.default_rng(block + replicas + hint)
    nb, npr = 40_000, 300_000
    bk = rng.permutation(60_000)[:nb].astype(np.int64) - 30_000
    bk[:500] = bk[500:1000]  # 500 keys with two build rows
    bk[1000] = SENTINEL
    bg = rng.integers(0, 20, nb).astype(np.int64)  # inline group column
    bg[2000:2010] = SENTINEL
    bv = rng.integers(-10**12, 10**12, nb).astype(np.int64)  # inline DECIMAL(18,2) value
    pk = rng.integers(-31_000, 31_000, npr).astype(np.int64)
    pk[::1000] = SENTINEL
    pk[1::97] = bk[2000]  # rows that meet the sentinel group
    pvalid = rng.random(npr) > 0.03
    pq = rng.integers(1, 50, npr).astype(np.int32)
    pf = rng.normal(0, 1, npr)
    pgrp = rng.integers(0, 30, npr).astype(np.int32)
    sel = np.sort(rng.choice(npr, npr // 2, replace=False)).astype(np.uint32)
    j = HashJoin(ctx, [L.INT64, L.INT64, L.DECIMAL], 0, inline=[1, 2], scales=[0, 0, 2], build_hint=nb)
    j.set_tuning(lean=True, block=block, replicas=replicas)
    j.build([Col(ctx.to_device(bk), L.INT64), Col(ctx.to_device(bg), L.INT64),
             Col(ctx.to_device(bv), L.DECIMAL, None, 2)], nb)
    j.finish()
    pw = bitmap_from_mask(pvalid)
    pcols = [Col(ctx.to_device(pk), L.INT64, ctx.to_device(pw)), Col(ctx.to_device(pq), L.INT32),
             Col(ctx.to_device(pf), L.FLOAT64), Col(ctx.to_device(pgrp), L.INT32)]
    types = [L.INT64, L.INT32, L.FLOAT64, L.INT32, L.INT64, L.INT64, L.DECIMAL]
    scales = [0, 0, 0, 0, 0, 0, 2]
    pp, bb = O.hash_join((bk, O.INT64), (pk, O.INT64, pw), probe_sel=sel, cap=npr * 4)
    for gcol, aggs in ((5, [AggSpec(L.SUM, 1, 6), AggSpec(L.COUNT, -1), AggSpec(L.MAX, 2), AggSpec(L.SUM, 2)]),
                       (3, [AggSpec(L.SUM, 6), AggSpec(L.MIN, 2), AggSpec(L.AVG, 1), AggSpec(L.COUNT, 1)])):
        agg = HashAggregate(ctx, types, [gcol], aggs, scales=scales, group_hint=hint)
        j.probe_agg(pcols, 0, npr, agg, sel=(ctx.to_device(sel), len(sel)))
        agg.finalize()
        prod = pq[pp].astype(np.int64) * bv[bb]  # |.| < 50 * 10^12
        prod128 = np.empty(2 * len(prod), np.uint64)
        prod128[0::2] = np.where(prod < 0, np.uint64(2**64 - 1), np.uint64(0))
        prod128[1::2] = prod.view(np.uint64)
        jcols = [(pk[pp], O.INT64), (pq[pp], O.INT32), (pf[pp], O.FLOAT64), (pgrp[pp], O.INT32), (bk[bb], O.INT64),
                 (bg[bb], O.INT64), (bv[bb], O.DECIMAL, None, 2), (prod128, O.DECIMAL128, None, 2)]
        oaggs = [(a.func, 7 if a.col2 >= 0 else a.col) for a in aggs]
        res = O.hash_agg(jcols, [gcol], oaggs)
        ref = oracle_groups(res)
        # per-group sum(|f|) for the float-sum tolerance
        absf = None
        if any(a.func in (L.SUM, L.AVG) and a.col == 2 for a in aggs):
            g = np.unique(ref["keys"][:, 0])
            idx = np.searchsorted(g, bg[bb] if gcol == 5 else pgrp[pp].astype(np.int64))
            absf = np.zeros(len(g))
            np.add.at(absf, idx, np.abs(pf[pp]))
        abs_sums = [absf if (a.func in (L.SUM, L.AVG) and a.col == 2) else None for a in aggs]
        compare(gpu_groups(agg), ref, oaggs, types + [O.DECIMAL128], scales + [2], abs_sums=abs_sums)
        agg.close()
    j.close()


@pytest.mark.parametrize("anti", [False, True])
def test_semi_anti_probe(ctx, anti):
    """Semi / anti probe: ascending probe positions whose key has (semi) / has
    no (anti) build row; NULL probe keys never match; duplicate build keys,
    NULL build keys, a selection vector, int32 and int64 keys."""
    rng = np.random.default_rng(41 if anti else 42)
    nb, npr = 50_000, 400_000
    for dt, t in ((np.int64, L.INT64), (np.int32, L.INT32)):
        bk = rng.integers(-60_000, 60_000, nb).astype(dt)
        pk = rng.integers(-90_000, 90_000, npr).astype(dt)
        bvalid = rng.random(nb) > 0.05
        pvalid = rng.random(npr) > 0.05
        bw, pw = bitmap_from_mask(bvalid), bitmap_from_mask(pvalid)
        j = HashJoin(ctx, [t], 0)
        j.build([Col(ctx.to_device(bk), t, ctx.to_device(bw))], nb)
        j.finish()
        key = Col(ctx.to_device(pk), t, ctx.to_device(pw))
        sel = np.sort(rng.choice(npr, npr // 3, replace=False)).astype(np.uint32)
        dsel = ctx.to_device(sel)
        out = ctx.alloc(npr * 4)
        print(f"[join] semi anti={anti} key={dt.__name__}", flush=True)
        for s in (None, (dsel, len(sel))):
            n = j.semi(key, npr, out, anti=anti, sel=s)
            got = out.to_host(np.uint32, n)
            ref = O.semi_join((bk, t, bw), (pk, t, pw), probe_sel=None if s is None else sel, anti=anti)
            np.testing.assert_array_equal(got, ref)
        j.close()


@pytest.mark.parametrize("n_build", [0, 1, 40_000, 2_000_000])
def test_bloom_prefilter_matches_oracle(ctx, n_build):
    """Device Bloom build == the oracle's words bit for bit; the filter's
    ordered selection == the oracle's (selection vector, NULL probe keys,
    duplicate and NULL build keys, int32 and int64 probe keys)."""
    print(f"[join] bloom n_build={n_build}", flush=True)
    rng = np.random.default_rng(n_build + 7)
    bk = rng.integers(-10**12, 10**12, max(n_build, 1)).astype(np.int64)[:n_build]
    if n_build > 10:
        bk[: n_build // 10] = bk[n_build // 10: 2 * (n_build // 10)]  # duplicates
    bvalid = rng.random(n_build) > 0.02
    j = HashJoin(ctx, [L.INT64], 0, build_hint=max(n_build, 1))
    if n_build:
        j.build([Col(ctx.to_device(bk), L.INT64, ctx.to_device(bitmap_from_mask(bvalid)))], n_build)
    j.finish()
    j.bloom_build()
    want_words = O.bloom_build(bk, bvalid)
    np.testing.assert_array_equal(j.bloom_words(), want_words)
    npr = 500_000
    pk = np.where(rng.random(npr) < 0.5, bk[rng.integers(0, max(n_build, 1), npr)] if n_build else 0,
                  rng.integers(-10**12, 10**12, npr)).astype(np.int64)
    pvalid = rng.random(npr) > 0.03
    sel = np.sort(rng.choice(npr, npr // 2, replace=False)).astype(np.uint32)
    out = ctx.alloc(npr * 4)
    key = Col(ctx.to_device(pk), L.INT64, ctx.to_device(bitmap_from_mask(pvalid)))
    for s_ in (None, sel):
        n = j.bloom_filter(key, npr, out, sel=None if s_ is None else (ctx.to_device(s_), len(s_)))
        np.testing.assert_array_equal(out.to_host(np.uint32, n), O.bloom_filter(want_words, pk, pvalid, s_))
    # int32 probe keys are widened like intKeyFromVector
    pk32 = (pk % 1000).astype(np.int32)
    n = j.bloom_filter(Col(ctx.to_device(pk32), L.INT32), npr, out)
    np.testing.assert_array_equal(out.to_host(np.uint32, n), O.bloom_filter(want_words, pk32.astype(np.int64)))
    j.close()


@pytest.mark.parametrize("cap", [5000, 70_000])
def test_bounded_probe_windows_over_a_large_batch(ctx, cap):
    """A bounded probe of a 300K-row batch with a selection vector examines a
    window of positions per call (twice the room, at least 64K): every call
    returns at most `cap` pairs, calls resume across windows and chains, and
    the pairs are the oracle's multiset."""
    print(f"[join] windows cap={cap}", flush=True)
    rng = np.random.default_rng(cap + 1)
    nb, npr = 30_000, 300_000
    bk = rng.integers(0, 40_000, nb).astype(np.int64)
    bk[:2000] = 5  # a chain longer than the smaller bound's room
    pk = rng.integers(0, 60_000, npr).astype(np.int64)
    pk[::997] = 5
    sel = np.sort(rng.choice(npr, 250_000, replace=False)).astype(np.uint32)
    j = HashJoin(ctx, [L.INT64], 0, build_hint=nb)
    j.build([Col(ctx.to_device(bk), L.INT64)], nb)
    j.finish()
    want_p, want_b = O.hash_join((bk, O.INT64), (pk, O.INT64), probe_sel=sel, cap=npr * 40)
    tp, tb = ctx.alloc(cap * 4), ctx.alloc(cap * 4)
    key = Col(ctx.to_device(pk), L.INT64)
    dsel = (ctx.to_device(sel), len(sel))
    cur, calls = None, 0
    ps, bs = [], []
    while True:
        k, cur = j.probe(key, npr, tp, tb, cap, cur, sel=dsel)
        assert k <= cap
        calls += 1
        ps.append(tp.to_host(np.uint32, k))
        bs.append(tb.to_host(np.uint32, k))
        if cur.done:
            break
        assert calls < 10_000
    p, b = np.concatenate(ps), np.concatenate(bs)
    np.testing.assert_array_equal(pairs_set(p, b), pairs_set(want_p, want_b))
