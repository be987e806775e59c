"""GPU parity of the grouped aggregate (wg_agg_*) against the CPU oracle.

Every test generates its inputs on the device from a fixed seed, copies them
back, runs the oracle on the same rows and compares sorted by key: integer,
COUNT, MIN/MAX and Int128 DECIMAL results bit-exact, float64 SUM/AVG within
FLOAT_RTOL (tests/helpers.py).
"""
import numpy as np
import pytest

import oracle as O
from helpers import compare, gpu_groups, oracle_groups
from wadjet_amd import _lib as L
from wadjet_amd import workloads as W
from wadjet_amd.gpu import AggSpec, Col, HashAggregate, Pred, bitmap_from_mask

pytestmark = pytest.mark.gpu

NP = {L.INT32: np.int32, L.INT64: np.int64, L.FLOAT64: np.float64, L.DECIMAL: np.int64, L.UINT32: np.uint32,
      L.UINT16: np.uint16, L.IPV4: np.int64}


def host_cols(cols, n):
    out = []
    for c in cols:
        arr = c.data.to_host(NP[c.type], n)
        v = None if c.validity is None else c.validity.to_host(np.uint64, (n + 63) // 64)
        out.append((arr, c.type, v, c.scale))
    return out


def run_gpu(ctx, cols, n, types, keys, aggs, scales=None, nullable=None, pred=None, sel=None, hint=0, tuning=None):
    agg = HashAggregate(ctx, types, keys, aggs, scales=scales, nullable=nullable, group_hint=hint)
    if tuning:
        agg.set_tuning(*tuning)
    agg.consume(cols, n, sel=sel, pred=pred)
    agg.finalize()
    return agg


def run_oracle(hcols, keys, aggs, pred=None, sel=None):
    p = None if pred is None else (pred.col, pred.op, pred.i64, pred.f64)
    return O.hash_agg(hcols, keys, [(a.func, a.col) for a in aggs], sel=sel, pred=p)


def abs_sums(hcols, keys, aggs, pred=None, sel=None):
    """per-group sum(|v|) for float SUM/AVG aggregates (oracle group order)."""
    out = []
    for a in aggs:
        if a.func in (L.SUM, L.AVG) and hcols[a.col][1] == L.FLOAT64:
            v = hcols[a.col][2] if len(hcols[a.col]) > 2 else None
            h2 = list(hcols) + [(np.abs(hcols[a.col][0]), L.FLOAT64, v)]
            r = O.hash_agg(h2, keys, [(O.SUM, len(hcols))], sel=sel,
                           pred=None if pred is None else (pred.col, pred.op, pred.i64, pred.f64))
            out.append(oracle_groups(r)["acc"][:, 0]["sum_f64"])
        else:
            out.append(None)
    return out


def check(ctx, cols, n, types, keys, aggs, scales=None, nullable=None, pred=None, sel=None, hint=0, tuning=None):
    print(f"[agg] n={n} keys={keys} types={types} pred={pred} sel={'yes' if sel is not None else 'no'} "
          f"hint={hint} tuning={tuning}", flush=True)
    agg = run_gpu(ctx, cols, n, types, keys, aggs, scales, nullable, pred, sel, hint, tuning)
    hcols = host_cols(cols, n)
    hsel = None if sel is None else (sel.to_host(np.uint32) if hasattr(sel, "to_host") else sel)
    ref = oracle_groups(run_oracle(hcols, keys, aggs, pred, hsel))
    compare(gpu_groups(agg), ref, [(a.func, a.col) for a in aggs], types, scales, abs_sums(hcols, keys, aggs, pred, hsel))
    return agg


def test_kv1k_shape(ctx):
    """k int64 over 1000 values: SUM(v), COUNT(*), MIN(v), MAX(v) GROUP BY k."""
    cols = W.gen_kv1k(ctx, 1_000_000)
    agg = check(ctx, cols, 1_000_000, W.KV1K_TYPES, [0], W.KV1K_AGGS)
    assert agg.num_groups() == 1000


@pytest.mark.parametrize("tuning", [None, (0, 256), (256, 1024), (2048, 512)])
def test_zipf_groupby_small(ctx, tuning):
    """Composite Zipf key, Int128 DECIMAL sum, float min/max/avg, fused filter."""
    n = 3_000_000
    cols = W.gen_zipf_groupby(ctx, n)
    check(ctx, cols, n, W.ZIPF_TYPES, W.ZIPF_KEYS, W.ZIPF_AGGS, W.ZIPF_SCALES, pred=W.ZIPF_PRED, tuning=tuning)


@pytest.mark.parametrize("chunk,hot", [(1 << 16, 512), (2048, 512), (1 << 16, 0), (2048, 0), (1 << 16, 2048)])
def test_zipf_groupby_partitioned(ctx, chunk, hot):
    """The partitioned consume (count, scatter, per-partition LDS aggregation);
    a small chunk splits the large partitions over several workgroups; the
    heavy hitters of the Zipf key aggregate in the scatter pass (hot > 0)."""
    n = 3_000_000
    cols = W.gen_zipf_groupby(ctx, n, seed=31)
    print(f"[agg] partitioned chunk={chunk} hot={hot}", flush=True)
    agg = HashAggregate(ctx, W.ZIPF_TYPES, W.ZIPF_KEYS, W.ZIPF_AGGS, scales=W.ZIPF_SCALES, group_hint=300_000)
    agg.set_strategy(2, 2048, 1024, chunk)
    agg.set_heavy_hitters(hot)
    agg.consume(cols, n, pred=W.ZIPF_PRED)
    t = agg.pass_times()
    assert t[0] == 0 and t[2] > 0, t  # the partitioned passes ran
    agg.finalize()
    hcols = host_cols(cols, n)
    ref = oracle_groups(run_oracle(hcols, W.ZIPF_KEYS, W.ZIPF_AGGS, W.ZIPF_PRED))
    compare(gpu_groups(agg), ref, [(a.func, a.col) for a in W.ZIPF_AGGS], W.ZIPF_TYPES, W.ZIPF_SCALES,
            abs_sums(hcols, W.ZIPF_KEYS, W.ZIPF_AGGS, W.ZIPF_PRED))


def test_partitioned_overflow_grows_and_replays(ctx):
    """Partitioned consume into a table far too small: items stop at the
    overflow headroom, the table grows, overflowed positions replay."""
    n = 2_000_000
    cols = W.gen_zipf_groupby(ctx, n, seed=37)
    agg = HashAggregate(ctx, W.ZIPF_TYPES, W.ZIPF_KEYS, W.ZIPF_AGGS, scales=W.ZIPF_SCALES, group_hint=1)
    agg.set_strategy(2, 1024, 512, 4096)
    agg.consume(cols, n, pred=W.ZIPF_PRED)
    agg.finalize()
    assert agg.stats()["capacity"] > 100_000
    hcols = host_cols(cols, n)
    ref = oracle_groups(run_oracle(hcols, W.ZIPF_KEYS, W.ZIPF_AGGS, W.ZIPF_PRED))
    compare(gpu_groups(agg), ref, [(a.func, a.col) for a in W.ZIPF_AGGS], W.ZIPF_TYPES, W.ZIPF_SCALES,
            abs_sums(hcols, W.ZIPF_KEYS, W.ZIPF_AGGS, W.ZIPF_PRED))


@pytest.mark.parametrize("hot", [512, 64, 0])
def test_partitioned_two_consumes_and_sentinel_keys(ctx, hot):
    """Two partitioned consumes into one aggregate (keys already in the table
    are merged by their owner), with sentinel-valued and NULL keys on the
    side path and 30 % of the rows on 50 heavy keys (hot set)."""
    rng = np.random.default_rng(41 + hot)
    n = 2_600_000
    k = rng.integers(0, 400_000, n).astype(np.int64)
    heavy = rng.random(n) < 0.3
    k[heavy] = rng.integers(0, 50, int(heavy.sum())) * 7919
    k[::1000] = -0x7FFFFFFFFFFFFFFF
    kvalid = rng.random(n) > 0.01
    v = rng.integers(-10**12, 10**12, n).astype(np.int64)
    f = rng.standard_normal(n)
    types = [L.INT64, L.INT64, L.FLOAT64]
    aggs = [AggSpec(L.SUM, 1), AggSpec(L.COUNT, -1), AggSpec(L.MAX, 2), AggSpec(L.AVG, 1), AggSpec(L.MIN, 1)]
    print(f"[agg] partitioned two consumes hot={hot}", flush=True)
    agg = HashAggregate(ctx, types, [0], aggs, nullable=[True, False, False], group_hint=400_000)
    agg.set_strategy(2, 2048, 1024, 1 << 15)
    agg.set_heavy_hitters(hot)
    h = n // 2
    for lo, hi in ((0, h), (h, n)):
        kv = ctx.to_device(bitmap_from_mask(kvalid[lo:hi]))
        agg.consume([Col(ctx.to_device(k[lo:hi]), L.INT64, kv), Col(ctx.to_device(v[lo:hi]), L.INT64),
                     Col(ctx.to_device(f[lo:hi]), L.FLOAT64)], hi - lo)
    agg.finalize()
    hcols = [(k, L.INT64, bitmap_from_mask(kvalid)), (v, L.INT64), (f, L.FLOAT64)]
    ref = oracle_groups(run_oracle(hcols, [0], aggs))
    compare(gpu_groups(agg), ref, [(a.func, a.col) for a in aggs], types, None, abs_sums(hcols, [0], aggs))


def test_table_growth_from_tiny_hint(ctx):
    """Start from the minimum table and let overflow + grow + replay do the work."""
    n = 2_000_000
    cols = W.gen_zipf_groupby(ctx, n, seed=7)
    agg = check(ctx, cols, n, W.ZIPF_TYPES, W.ZIPF_KEYS, W.ZIPF_AGGS, W.ZIPF_SCALES, pred=W.ZIPF_PRED, hint=1)
    assert agg.stats()["capacity"] > 1024


def test_selection_vector_matches_predicate(ctx):
    """A device Sel from the filter kernel gives the same groups as the fused predicate."""
    from wadjet_amd.gpu import filter_sel
    n = 1_000_000
    cols = W.gen_zipf_groupby(ctx, n, seed=11)
    sel, k = filter_sel(ctx, cols[3], n, L.GT, f64=-0.5)
    a1 = run_gpu(ctx, cols, n, W.ZIPF_TYPES, W.ZIPF_KEYS, W.ZIPF_AGGS, W.ZIPF_SCALES, sel=(sel, k))
    a2 = run_gpu(ctx, cols, n, W.ZIPF_TYPES, W.ZIPF_KEYS, W.ZIPF_AGGS, W.ZIPF_SCALES, pred=W.ZIPF_PRED)
    g1, g2 = gpu_groups(a1), gpu_groups(a2)
    assert g1["n"] == g2["n"]
    np.testing.assert_array_equal(g1["keys"], g2["keys"])
    for a in range(len(W.ZIPF_AGGS)):
        if a == 4:
            np.testing.assert_allclose(g1["values"][a], g2["values"][a], rtol=1e-12)
        else:
            np.testing.assert_array_equal(g1["values"][a], g2["values"][a])


def test_all_types_and_nulls(ctx):
    """Every column type, NULLs in values and in a single key (the NULL group)."""
    rng = np.random.default_rng(5)
    n = 200_000
    key = rng.integers(-50, 50, n).astype(np.int64)
    kvalid = rng.random(n) > 0.05
    i32 = rng.integers(-2**31, 2**31 - 1, n).astype(np.int32)
    i64 = rng.integers(-2**62, 2**62, n).astype(np.int64)
    f64 = rng.standard_normal(n)
    dec = rng.integers(-10**15, 10**15, n).astype(np.int64)
    u32 = rng.integers(0, 2**32, n).astype(np.uint32)
    u16 = rng.integers(0, 2**16, n).astype(np.uint16)
    ipv4 = rng.integers(0, 2**32, n).astype(np.int64)
    masks = [kvalid] + [rng.random(n) > 0.1 for _ in range(7)]
    arrays = [key, i32, i64, f64, dec, u32, u16, ipv4]
    types = [L.INT64, L.INT32, L.INT64, L.FLOAT64, L.DECIMAL, L.UINT32, L.UINT16, L.IPV4]
    scales = [0, 0, 0, 0, 3, 0, 0, 0]
    cols = []
    for arr, t, m, s in zip(arrays, types, masks, scales):
        cols.append(Col(ctx.to_device(arr), t, ctx.to_device(bitmap_from_mask(m)), s))
    for c in range(1, 8):
        aggs = [AggSpec(L.COUNT, -1)] + [AggSpec(f, c) for f in (L.SUM, L.COUNT, L.MIN, L.MAX, L.AVG)]
        check(ctx, cols, n, types, [0], aggs, scales, nullable=[True] * 8)


def test_decimal_carries_exact(ctx):
    """Int128 sums that cross 2^64 and go negative: carry-save limbs stay exact."""
    rng = np.random.default_rng(9)
    n = 500_000
    k = rng.integers(0, 8, n).astype(np.int32)
    d = rng.integers(2**62, 2**63 - 1, n).astype(np.int64)
    d[rng.random(n) < 0.3] *= -1
    cols = [Col(ctx.to_device(k), L.INT32), Col(ctx.to_device(d), L.DECIMAL, None, 4)]
    check(ctx, cols, n, [L.INT32, L.DECIMAL], [0], [AggSpec(L.SUM, 1), AggSpec(L.AVG, 1)], [0, 4])


def test_sentinel_key_and_extremes(ctx):
    """The key value equal to the table's EMPTY sentinel and int64 extremes."""
    k = np.array([-0x7FFFFFFFFFFFFFFF, 0, -2**63, 2**63 - 1, -0x7FFFFFFFFFFFFFFF, 5] * 1000, dtype=np.int64)
    v = np.arange(len(k), dtype=np.int64) * 7 - 3000
    cols = [Col(ctx.to_device(k), L.INT64), Col(ctx.to_device(v), L.INT64)]
    agg = check(ctx, cols, len(k), [L.INT64, L.INT64], [0], W.KV1K_AGGS)
    assert agg.num_groups() == 5


def test_host_memory_morsels_and_merge(ctx):
    """2048-row host morsels with Sel vectors (coalesced staging), two clone
    sinks merged (MergeSink), equals one device pass."""
    rng = np.random.default_rng(3)
    n = 300_000
    a = rng.integers(0, 5000, n).astype(np.int32)
    b = rng.integers(-3, 3, n).astype(np.int32)
    p = rng.integers(100, 10**7, n).astype(np.int64)
    f = rng.standard_normal(n)
    types = W.ZIPF_TYPES
    aggs = W.ZIPF_AGGS
    s1 = HashAggregate(ctx, types, W.ZIPF_KEYS, aggs, scales=W.ZIPF_SCALES)
    s2 = HashAggregate(ctx, types, W.ZIPF_KEYS, aggs, scales=W.ZIPF_SCALES)
    all_sel = []
    for m, start in enumerate(range(0, n, 2048)):
        end = min(n, start + 2048)
        cols = [Col(x[start:end], t, None, sc) for x, t, sc in zip((a, b, p, f), types, W.ZIPF_SCALES)]
        sel = np.nonzero(f[start:end] > -0.5)[0].astype(np.uint32)
        all_sel.append(sel + start)
        (s1 if m % 2 == 0 else s2).consume(cols, end - start, sel=sel, device=False)
    s1.merge(s2)
    s1.finalize()
    hcols = [(a, L.INT32), (b, L.INT32), (p, L.DECIMAL, None, 2), (f, L.FLOAT64)]
    sel = np.concatenate(all_sel)
    ref = oracle_groups(run_oracle(hcols, W.ZIPF_KEYS, aggs, sel=sel))
    compare(gpu_groups(s1), ref, [(x.func, x.col) for x in aggs], types, W.ZIPF_SCALES,
            abs_sums(hcols, W.ZIPF_KEYS, aggs, sel=sel))


def test_multi_key_null_is_rejected(ctx):
    from wadjet_amd import NullKeyError
    a = np.array([1, 2, 3], np.int32)
    m = bitmap_from_mask(np.array([True, False, True]))
    agg = HashAggregate(ctx, [L.INT32, L.INT32], [0, 1], [AggSpec(L.COUNT, -1)], nullable=[True, False])
    with pytest.raises(NullKeyError):
        agg.consume([Col(ctx.to_device(a), L.INT32, ctx.to_device(m)), Col(ctx.to_device(a), L.INT32)], 3)


def test_empty_input_and_reset(ctx):
    agg = HashAggregate(ctx, W.KV1K_TYPES, [0], W.KV1K_AGGS)
    cols = W.gen_kv1k(ctx, 10)
    agg.consume(cols, 0)
    agg.finalize()
    assert agg.num_groups() == 0
    agg.reset()
    agg.consume(cols, 10)
    agg.finalize()
    assert agg.num_groups() > 0


def test_dense_emission_then_merge_and_more_rows(ctx):
    """A partitioned consume into an empty aggregate writes its finished groups
    straight to the output; later work folds them back first: a MergeSink of
    two such aggregates, a direct consume after one, and a spill run of the
    result merged into a third aggregate all equal the oracle."""
    n = 3_000_000
    cols = W.gen_zipf_groupby(ctx, n, seed=43)
    hcols = host_cols(cols, n)
    aggs = [(a.func, a.col) for a in W.ZIPF_AGGS]
    ref = oracle_groups(run_oracle(hcols, W.ZIPF_KEYS, W.ZIPF_AGGS, W.ZIPF_PRED))
    sums = abs_sums(hcols, W.ZIPF_KEYS, W.ZIPF_AGGS, W.ZIPF_PRED)
    half = n // 2
    print("[agg] dense emission: merge of two dense aggregates", flush=True)
    a1 = HashAggregate(ctx, W.ZIPF_TYPES, W.ZIPF_KEYS, W.ZIPF_AGGS, scales=W.ZIPF_SCALES, group_hint=400_000)
    a2 = HashAggregate(ctx, W.ZIPF_TYPES, W.ZIPF_KEYS, W.ZIPF_AGGS, scales=W.ZIPF_SCALES, group_hint=400_000)
    for a in (a1, a2):
        a.set_strategy(2, 2048, 1024, 1 << 16)
    sel1 = ctx.to_device(np.arange(0, half, dtype=np.uint32))
    sel2 = ctx.to_device(np.arange(half, n, dtype=np.uint32))
    a1.consume(cols, n, sel=(sel1, half), pred=W.ZIPF_PRED)
    a2.consume(cols, n, sel=(sel2, n - half), pred=W.ZIPF_PRED)
    a1.merge(a2)
    a1.finalize()
    compare(gpu_groups(a1), ref, aggs, W.ZIPF_TYPES, W.ZIPF_SCALES, sums)

    print("[agg] dense emission: a direct consume after a dense one", flush=True)
    a3 = HashAggregate(ctx, W.ZIPF_TYPES, W.ZIPF_KEYS, W.ZIPF_AGGS, scales=W.ZIPF_SCALES, group_hint=400_000)
    a3.set_strategy(2, 2048, 1024, 1 << 16)
    a3.consume(cols, n, sel=(sel1, half), pred=W.ZIPF_PRED)
    a3.set_strategy(1)
    a3.consume(cols, n, sel=(sel2, n - half), pred=W.ZIPF_PRED)
    a3.finalize()
    compare(gpu_groups(a3), ref, aggs, W.ZIPF_TYPES, W.ZIPF_SCALES, sums)

    print("[agg] dense emission: reset, again, spill run merged into a fresh aggregate", flush=True)
    a3.reset()
    a3.set_strategy(2, 2048, 1024, 1 << 16)
    a3.consume(cols, n, pred=W.ZIPF_PRED)
    a3.finalize()
    compare(gpu_groups(a3), ref, aggs, W.ZIPF_TYPES, W.ZIPF_SCALES, sums)
    a4 = HashAggregate(ctx, W.ZIPF_TYPES, W.ZIPF_KEYS, W.ZIPF_AGGS, scales=W.ZIPF_SCALES, group_hint=400_000)
    a4.set_strategy(2, 2048, 1024, 1 << 16)
    a4.consume(cols, n, sel=(sel1, half), pred=W.ZIPF_PRED)  # dense rows pending when the run arrives
    a5 = HashAggregate(ctx, W.ZIPF_TYPES, W.ZIPF_KEYS, W.ZIPF_AGGS, scales=W.ZIPF_SCALES, group_hint=400_000)
    a5.set_strategy(2, 2048, 1024, 1 << 16)
    a5.consume(cols, n, sel=(sel2, n - half), pred=W.ZIPF_PRED)
    a5.finalize()
    a4.merge_run(a5.spill_run())
    a4.finalize()
    g4 = gpu_groups(a4)
    assert g4["n"] == ref["n"]
    np.testing.assert_array_equal(g4["keys"], ref["keys"])
    for x in (a1, a2, a3, a4, a5):
        x.close()


@pytest.mark.parametrize("case", ["zipf", "nulls", "wide"])
def test_fetch_to_device_columns_equals_host_fetch(ctx, case):
    """wg_agg_fetch into device columns (one pass writing every key, value and
    validity column) == the host fetch, for a composite key with Int128 and
    float aggregates, a nullable single key with the NULL group, and a 128-bit
    key through the dictionary; also for a window of the groups."""
    print(f"[agg] device fetch {case}", flush=True)
    rng = np.random.default_rng(71)
    if case == "zipf":
        n = 2_000_000
        cols = W.gen_zipf_groupby(ctx, n, seed=5)
        agg = run_gpu(ctx, cols, n, W.ZIPF_TYPES, W.ZIPF_KEYS, W.ZIPF_AGGS, W.ZIPF_SCALES, pred=W.ZIPF_PRED,
                      hint=200_000)
    elif case == "nulls":
        n = 300_000
        k = rng.integers(-50, 50, n).astype(np.int64)
        v = rng.integers(-10**9, 10**9, n).astype(np.int64)
        m = rng.random(n) > 0.05
        cols = [Col(ctx.to_device(k), L.INT64, ctx.to_device(bitmap_from_mask(m))), Col(ctx.to_device(v), L.INT64,
                                                                                        ctx.to_device(bitmap_from_mask(m)))]
        agg = run_gpu(ctx, cols, n, [L.INT64, L.INT64], [0], [AggSpec(L.SUM, 1), AggSpec(L.COUNT, 1),
                                                               AggSpec(L.MIN, 1), AggSpec(L.AVG, 1)],
                      nullable=[True, True])
    else:
        n = 500_000
        a = rng.integers(0, 40_000, n).astype(np.int64)
        b = rng.integers(-3, 3, n).astype(np.int64)
        v = rng.integers(-10**6, 10**6, n).astype(np.int64)
        cols = [Col(ctx.to_device(x), L.INT64) for x in (a, b, v)]
        agg = run_gpu(ctx, cols, n, [L.INT64] * 3, [0, 1], [AggSpec(L.SUM, 2), AggSpec(L.COUNT, -1)])
    host = agg.fetch()
    g = host["n"]
    ks, vs, vv = agg.output_buffers(g)
    agg.fetch_device(ks, vs, vv)
    for x, k in enumerate(agg.keys):
        np.testing.assert_array_equal(ks[x].to_host(host["keys"][x].dtype, g), host["keys"][x])
    for a in range(len(agg.aggs)):
        np.testing.assert_array_equal(vs[a].to_host(host["values"][a].dtype, g).view(np.uint64),
                                      host["values"][a].view(np.uint64))
        np.testing.assert_array_equal(vv[a].to_host(np.uint8, g).astype(bool), host["valid"][a])
    off, cnt = g // 3, g // 3
    agg.fetch_device([None] * len(agg.keys), vs, [None] * len(agg.aggs), offset=off, count=cnt)
    np.testing.assert_array_equal(vs[0].to_host(host["values"][0].dtype, cnt).view(np.uint64),
                                  host["values"][0][off:off + cnt].view(np.uint64))
    agg.close()


@pytest.mark.parametrize("n_vals", [1, 3, 5])
@pytest.mark.parametrize("null_key", [False, True])
def test_partitioned_record_widths(ctx, n_vals, null_key):
    """The partitioned consume with records of 2, 4 and 8 words (one, three
    and five value columns) through the typed pass 1 (non-null 8-byte
    columns) and the general one (a nullable key, an int32 value column),
    with a skewed key so the heavy-hitter set takes part, against the oracle."""
    rng = np.random.default_rng(300 + n_vals + 10 * null_key)
    # five value columns: a batch large enough for the automatic choice of
    # the partitioned consume, whose default 2048-slot LDS table the wide
    # slots (128 bytes) halve to fit
    n = 1_300_000 if n_vals < 5 else (1 << 22) + 100_000
    k = (rng.pareto(1.2, n) * 1000).astype(np.int64) % 250_000
    kvalid = rng.random(n) > 0.02
    v = rng.integers(-10**12, 10**12, n).astype(np.int64)
    f = rng.standard_normal(n)
    d = rng.integers(-10**13, 10**13, n).astype(np.int64)
    i32 = rng.integers(-2**31, 2**31 - 1, n).astype(np.int32)
    ip = rng.integers(0, 2**32, n).astype(np.int64)
    arrays = [k, v, f, d, i32, ip]
    types = [L.INT64, L.INT64, L.FLOAT64, L.DECIMAL, L.INT32, L.IPV4]
    scales = [0, 0, 0, 2, 0, 0]
    aggs = {1: [AggSpec(L.SUM, 1), AggSpec(L.COUNT, -1)],
            3: [AggSpec(L.SUM, 1), AggSpec(L.MIN, 2), AggSpec(L.SUM, 3), AggSpec(L.AVG, 1), AggSpec(L.COUNT, -1)],
            5: [AggSpec(L.SUM, 1), AggSpec(L.MAX, 2), AggSpec(L.SUM, 3), AggSpec(L.MIN, 4), AggSpec(L.MAX, 5),
                AggSpec(L.AVG, 2)]}[n_vals]
    cols = [Col(ctx.to_device(k), L.INT64, ctx.to_device(bitmap_from_mask(kvalid)) if null_key else None)]
    cols += [Col(ctx.to_device(a), t, None, s) for a, t, s in zip(arrays[1:], types[1:], scales[1:])]
    print(f"[agg] partitioned record widths n_vals={n_vals} null_key={null_key}", flush=True)
    agg = HashAggregate(ctx, types, [0], aggs, scales=scales, nullable=[null_key] + [False] * 5, group_hint=250_000)
    if n_vals < 5:
        agg.set_strategy(2, 2048, 1024, 1 << 15)
    # with three value columns the rows arrive through a selection vector
    hsel = np.flatnonzero(rng.random(n) < 0.7).astype(np.uint32) if n_vals == 3 else None
    if hsel is None:
        agg.consume(cols, n)
    else:
        agg.consume(cols, n, sel=ctx.to_device(hsel))
    agg.finalize()
    assert agg.stats()["global_rows"] > 0  # rows went through the partitions
    hcols = [(k, L.INT64, bitmap_from_mask(kvalid) if null_key else None, 0)]
    hcols += [(a, t, None, s) for a, t, s in zip(arrays[1:], types[1:], scales[1:])]
    ref = oracle_groups(run_oracle(hcols, [0], aggs, sel=hsel))
    compare(gpu_groups(agg), ref, [(a.func, a.col) for a in aggs], types, scales,
            abs_sums(hcols, [0], aggs, sel=hsel))
