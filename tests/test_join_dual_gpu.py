"""GPU parity of joins on two int-class key columns (wg_join_create_dual +
wg_join_dual_probe_key; dualIntKeyFromVectors and the verifying chain walk,
internal/engine/exec/join.go:1621-1648, :2848-3130) against the oracle: the
joined pair multiset, semi/anti selections and the fused probe -> GROUP BY,
bit-exact."""
import numpy as np
import pytest

import oracle as O
from test_join_gpu import pairs_set, probe_all
from wadjet_amd import _lib as L
from wadjet_amd.gpu import AggSpec, Col, HashAggregate, HashJoin, bitmap_from_mask

pytestmark = pytest.mark.gpu


def dev(ctx, arr, t, mask=None):
    return Col(ctx.to_device(arr), t, None if mask is None else ctx.to_device(bitmap_from_mask(mask)))


def test_reference_cases(ctx):
    """TestHashJoin_DualIntKey / _SemiAnti (dual_int_test.go:11-140)."""
    print("[dual] reference cases", flush=True)
    j = HashJoin(ctx, [L.INT64, L.INT64, L.FLOAT64], (0, 1))
    j.build([dev(ctx, np.array([1, 2, 3], np.int64), L.INT64), dev(ctx, np.array([10, 20, 30], np.int64), L.INT64),
             dev(ctx, np.array([100.0, 200.0, 300.0]), L.FLOAT64)], 3)
    j.finish()
    key = j.dual_probe_key(dev(ctx, np.array([1, 2, 99], np.int64), L.INT64),
                           dev(ctx, np.array([10, 20, 99], np.int64), L.INT64), 3)
    p, b, _ = probe_all(j, key, 3, 32768, 16)
    assert sorted(zip(p.tolist(), b.tolist())) == [(0, 0), (1, 1)]
    j2 = HashJoin(ctx, [L.INT64, L.INT64], (0, 1))
    j2.build([dev(ctx, np.array([1, 2], np.int64), L.INT64), dev(ctx, np.array([10, 20], np.int64), L.INT64)], 2)
    j2.finish()
    key = j2.dual_probe_key(dev(ctx, np.array([1, 3, 2], np.int64), L.INT64),
                            dev(ctx, np.array([10, 30, 20], np.int64), L.INT64), 3)
    out = ctx.alloc(64)
    assert j2.semi(key, 3, out) == 2 and out.to_host(np.uint32, 2).tolist() == [0, 2]
    assert j2.semi(key, 3, out, anti=True) == 1 and out.to_host(np.uint32, 1).tolist() == [1]


@pytest.mark.parametrize("cap", [32768, 777])
def test_multiset_duplicates_nulls_mixed_widths(ctx, cap):
    """Duplicate build pairs (chains), pairs equal in one key only, NULLs in
    either key on both sides, int32 probe keys against int64 build keys, two
    build batches (host, device), a probe Sel; the bounded probe resumes."""
    print(f"[dual] multiset cap={cap}", flush=True)
    rng = np.random.default_rng(cap)
    nb, npr = 60_000, 150_000
    ba = rng.integers(-500, 500, nb).astype(np.int64)
    bb = rng.integers(0, 40, nb).astype(np.int64)
    ba[:200], bb[:200] = 7, 3  # one long chain
    bm = (rng.random(nb) > 0.03, rng.random(nb) > 0.03)
    pa = rng.integers(-600, 600, npr).astype(np.int32)
    pb = rng.integers(0, 45, npr).astype(np.int32)
    pa[::97], pb[::97] = 7, 3
    pm = (rng.random(npr) > 0.03, rng.random(npr) > 0.03)
    j = HashJoin(ctx, [L.INT64, L.INT32, L.INT64], (0, 2), build_hint=1000)
    half = nb // 2
    for lo, hi in ((0, half), (half, nb)):
        cols = [(ba[lo:hi], L.INT64, bm[0][lo:hi]), (np.arange(lo, hi, dtype=np.int32), L.INT32, None),
                (bb[lo:hi], L.INT64, bm[1][lo:hi])]
        if lo == 0:
            j.build([Col(a, t, None if m is None else bitmap_from_mask(m)) for a, t, m in cols], hi - lo, device=False)
        else:
            j.build([dev(ctx, a, t, m) for a, t, m in cols], hi - lo)
    j.finish()
    sel = np.sort(rng.choice(npr, npr // 2, replace=False)).astype(np.uint32)
    dsel = (ctx.to_device(sel), len(sel))
    key = j.dual_probe_key(dev(ctx, pa, L.INT32, pm[0]), dev(ctx, pb, L.INT32, pm[1]), npr, sel=dsel)
    # bounded probe over the Sel
    tp, tb = ctx.alloc(cap * 4), ctx.alloc(cap * 4)
    cur = L.ProbeCursor(0, -1, 0, 0)
    ps, bs = [], []
    while not cur.done:
        k, cur = j.probe(key, npr, tp, tb, cap, cur, sel=dsel)
        ps.append(tp.to_host(np.uint32, k))
        bs.append(tb.to_host(np.uint32, k))
    p, b = np.concatenate(ps), np.concatenate(bs)
    wm = lambda m: bitmap_from_mask(m)
    ep, eb = O.hash_join2((ba, bb), (pa, pb), (wm(bm[0]), wm(bm[1])), (wm(pm[0]), wm(pm[1])), probe_sel=sel)
    assert len(p) == len(ep) and len(ep) > 0
    assert np.array_equal(pairs_set(p, b), pairs_set(ep, eb))
    out = ctx.alloc(npr * 4)
    for anti in (False, True):
        m = j.semi(key, npr, out, anti=anti, sel=dsel)
        want = O.semi_join2((ba, bb), (pa, pb), (wm(bm[0]), wm(bm[1])), (wm(pm[0]), wm(pm[1])), probe_sel=sel,
                            anti=anti)
        assert np.array_equal(out.to_host(np.uint32, m), want), f"anti={anti}"


def test_fused_probe_groupby_and_bloom(ctx):
    """The id column appended to the probe columns feeds the fused probe ->
    GROUP BY (GROUP BY a build column, SUM of a probe column and COUNT(*)); the
    Bloom pre-filter on ids keeps every matching row."""
    print("[dual] fused + bloom", flush=True)
    rng = np.random.default_rng(5)
    nb, npr = 20_000, 400_000
    ba = rng.permutation(nb).astype(np.int64)
    bb = (ba % 11).astype(np.int64)
    grp = rng.integers(0, 25, nb).astype(np.int32)
    pa = rng.integers(0, nb + nb // 10, npr).astype(np.int64)
    pb = (pa % 11).astype(np.int64)
    pb[::5] += 1  # a fifth of the rows miss on the second key
    qty = rng.integers(1, 50, npr).astype(np.int32)
    j = HashJoin(ctx, [L.INT64, L.INT64, L.INT32], (0, 1), inline=[2], build_hint=nb)
    j.build([dev(ctx, ba, L.INT64), dev(ctx, bb, L.INT64), dev(ctx, grp, L.INT32)], nb)
    j.finish()
    key = j.dual_probe_key(dev(ctx, pa, L.INT64), dev(ctx, pb, L.INT64), npr)
    pcols = [dev(ctx, pa, L.INT64), dev(ctx, pb, L.INT64), dev(ctx, qty, L.INT32), key]
    types = [L.INT64, L.INT64, L.INT32, L.INT64, L.INT64, L.INT64, L.INT32]
    agg = HashAggregate(ctx, types, [6], [AggSpec(L.SUM, 2), AggSpec(L.COUNT, -1)], group_hint=32)
    j.probe_agg(pcols, 3, npr, agg)
    agg.finalize()
    r = agg.fetch()
    ep, eb = O.hash_join2((ba, bb), (pa, pb))
    want = {}
    for pr, br in zip(ep, eb):
        g = int(grp[br])
        s, c = want.get(g, (0, 0))
        want[g] = (s + int(qty[pr]), c + 1)
    got = {int(k): (int(s), int(c)) for k, s, c in zip(r["keys"][0], agg.fetch_state(0)["sum_lo"].view(np.int64),
                                                        r["values"][1])}
    assert got == want
    j.bloom_build()
    out = ctx.alloc(npr * 4)
    m = j.bloom_filter(key, npr, out)
    kept = set(out.to_host(np.uint32, m).tolist())
    assert set(ep.tolist()) <= kept


def test_dual_key_rejections(ctx):
    from wadjet_amd import WgError
    with pytest.raises(WgError):
        HashJoin(ctx, [L.INT64, L.FLOAT64], (0, 1))
    j = HashJoin(ctx, [L.INT64], 0)
    j.build([dev(ctx, np.arange(5, dtype=np.int64), L.INT64)], 5)
    j.finish()
    with pytest.raises(WgError):
        j.dual_probe_key(dev(ctx, np.arange(5, dtype=np.int64), L.INT64), dev(ctx, np.arange(5, dtype=np.int64),
                                                                              L.INT64), 5)
