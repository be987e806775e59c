"""GPU parity of GROUP BY keys wider than 64 bits: two to four int-class
columns packed into the reference's 128-bit {lo, hi} key (buildPackedLayout /
packedKeyAt, internal/engine/exec/packed_hash.go:384-472; the packed consume
loop, aggregate.go:2326-2588). Groups, keys and every accumulator against the
oracle, bit-exact for integers and Int128, float sums within FLOAT_RTOL."""
import numpy as np
import pytest

import oracle as O
from helpers import compare, gpu_groups, oracle_groups
from test_agg_gpu import abs_sums, run_oracle
from wadjet_amd import NullKeyError, UnsupportedError
from wadjet_amd import _lib as L
from wadjet_amd.gpu import AggSpec, Col, HashAggregate, bitmap_from_mask

pytestmark = pytest.mark.gpu

NP = {L.INT32: np.int32, L.INT64: np.int64, L.UINT32: np.uint32, L.UINT16: np.uint16, L.IPV4: np.int64,
      L.FLOAT64: np.float64, L.DECIMAL: np.int64}
EXTREME64 = [0, 1, -1, 2**63 - 1, -2**63, 1 << 40, -(1 << 40), 0xFFFFFFFF, 0xFFFFFFFFFFFF,
             -0x7FFFFFFFFFFFFFFF]  # the last is the int-key table's empty marker
EXTREME32 = [0, 1, -1, 2**31 - 1, -2**31, 1 << 30, -(1 << 30)]


def key_col(t, n, rng, card):
    """n values of type t drawn from `card` distinct values, extremes included."""
    if t in (L.INT64, L.IPV4):
        pool = np.array(EXTREME64 + list(rng.integers(-2**62, 2**62, card)), np.int64)
        if t == L.IPV4:
            pool = np.abs(pool) % (1 << 48)
    elif t == L.INT32:
        pool = np.array(EXTREME32 + list(rng.integers(-2**31, 2**31 - 1, card)), np.int64).astype(np.int32)
    elif t == L.UINT32:
        pool = rng.integers(0, 2**32, card + 2, dtype=np.int64).astype(np.uint32)
        pool[:2] = [0, 0xFFFFFFFF]
    else:
        pool = rng.integers(0, 65536, card + 2).astype(np.uint16)
    return pool[rng.integers(0, len(pool), n)]


VALS = [(L.INT64, None), (L.DECIMAL, 2), (L.FLOAT64, None)]
AGGS_OF = lambda k: [AggSpec(L.SUM, k), AggSpec(L.COUNT, -1), AggSpec(L.MIN, k), AggSpec(L.MAX, k),
                     AggSpec(L.SUM, k + 1), AggSpec(L.AVG, k + 2), AggSpec(L.MAX, k + 2), AggSpec(L.COUNT, k + 2)]


def make_batch(key_types, n, rng, card, fnull=0.0):
    keys = [key_col(t, n, rng, card) for t in key_types]
    v = rng.integers(-10**12, 10**12, n).astype(np.int64)
    d = rng.integers(0, 10**15, n).astype(np.int64)
    f = rng.standard_normal(n)
    fv = bitmap_from_mask(rng.random(n) >= fnull) if fnull else None
    return keys, v, d, f, fv


def check_wide(ctx, key_types, n, card, seed, hint=0, device=True, sel=None, strategy=None, fnull=0.0):
    print(f"[wide] keys={key_types} n={n} card={card} hint={hint} device={device} strategy={strategy}", flush=True)
    rng = np.random.default_rng(seed)
    keys, v, d, f, fv = make_batch(key_types, n, rng, card, fnull)
    nk = len(key_types)
    types = list(key_types) + [L.INT64, L.DECIMAL, L.FLOAT64]
    scales = [0] * nk + [0, 2, 0]
    aggs = AGGS_OF(nk)
    agg = HashAggregate(ctx, types, list(range(nk)), aggs, scales=scales, group_hint=hint,
                        nullable=[False] * nk + [False, False, fnull > 0])
    if strategy is not None:
        agg.set_strategy(strategy, part_lds_slots=1024)  # a 16-word group state
    arrs = keys + [v, d, f]
    if device:
        cols = [Col(ctx.to_device(a), t, ctx.to_device(fv) if (i == nk + 2 and fv is not None) else None, sc)
                for i, (a, t, sc) in enumerate(zip(arrs, types, scales))]
        agg.consume(cols, n, sel=None if sel is None else (ctx.to_device(sel), len(sel)))
    else:
        for start in range(0, n, 65_536):
            end = min(n, start + 65_536)
            cols = [Col(np.ascontiguousarray(a[start:end]), t, None, sc) for a, t, sc in zip(arrs, types, scales)]
            agg.consume(cols, end - start, device=False)
    agg.finalize()
    hcols = [(a, t, fv if (i == nk + 2 and fv is not None) else None, sc)
             for i, (a, t, sc) in enumerate(zip(arrs, types, scales))]
    ref = oracle_groups(run_oracle(hcols, list(range(nk)), aggs, sel=sel))
    compare(gpu_groups(agg), ref, [(a.func, a.col) for a in aggs], types, scales,
            abs_sums(hcols, list(range(nk)), aggs, sel=sel))
    return agg


# the eligible shapes of TestBuildPackedLayoutEligibility / TestPackedKeyRoundTrip
# (packed_hash_test.go:175-302) wider than 8 bytes, in this library's types
SHAPES = [
    [L.INT64, L.INT64],
    [L.INT64, L.INT32],
    [L.INT32, L.INT64],
    [L.INT32, L.INT32, L.INT32],
    [L.INT32, L.INT32, L.INT32, L.INT32],
    [L.INT32, L.INT64, L.INT32],
    [L.IPV4, L.IPV4],
    [L.UINT16, L.UINT16, L.INT32, L.INT32],
    [L.UINT32, L.INT64],
]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "-".join(map(str, s)))
def test_shapes_extremes_round_trip(ctx, shape):
    """Every column of every shape round-trips bit-exactly through the key,
    the extreme values (MinInt64, -1, 0xFFFFFFFF, the int-key empty marker)
    being ordinary groups."""
    check_wide(ctx, shape, 200_000, 40, seed=len(shape) * 7 + shape[0])


@pytest.mark.parametrize("n_ones", [20_000, 300_000])
def test_all_ones_key_words(ctx, n_ones):
    """Pairs with an all-ones word (int64 key columns at -1) never read as
    written while the claim pass runs: their rows take the verify pass, from
    the pass's row list (up to 64K rows) or over every row (more)."""
    print(f"[wide] all-ones key words in {n_ones} rows", flush=True)
    rng = np.random.default_rng(77 + n_ones)
    n = n_ones + 50_000
    a = rng.integers(-1000, 1000, n).astype(np.int64)
    a[rng.permutation(n)[:n_ones]] = -1
    b = rng.integers(0, 8, n).astype(np.int64)
    b[::7] = -1
    v = rng.integers(-10**12, 10**12, n).astype(np.int64)
    d = rng.integers(0, 10**15, n).astype(np.int64)
    f = rng.standard_normal(n)
    types = [L.INT64, L.INT64, L.INT64, L.DECIMAL, L.FLOAT64]
    scales = [0, 0, 0, 2, 0]
    aggs = AGGS_OF(2)
    agg = HashAggregate(ctx, types, [0, 1], aggs, scales=scales, group_hint=4096)
    arrs = [a, b, v, d, f]
    agg.consume([Col(ctx.to_device(x), t, None, sc) for x, t, sc in zip(arrs, types, scales)], n)
    agg.finalize()
    hcols = [(x, t, None, sc) for x, t, sc in zip(arrs, types, scales)]
    ref = oracle_groups(run_oracle(hcols, [0, 1], aggs))
    compare(gpu_groups(agg), ref, [(g.func, g.col) for g in aggs], types, scales, abs_sums(hcols, [0, 1], aggs))
    agg.close()


def test_many_groups_grow_the_dictionary(ctx):
    """~1.5M distinct pairs from a 1024 hint: the dictionary grows and the
    batch is encoded again; the table grows too."""
    check_wide(ctx, [L.INT64, L.INT64], 3_000_000, 1_500, seed=5)


@pytest.mark.parametrize("strategy", [1, 2])
def test_direct_and_partitioned_consume(ctx, strategy):
    """The id column through the direct and the partitioned consume."""
    check_wide(ctx, [L.INT32, L.INT64], 4_500_000, 600, seed=6, hint=300_000, strategy=strategy)


def test_host_batches_selection_and_nulls(ctx):
    """Host morsels through the staging path; then device rows through a
    Sel with NULL float values."""
    check_wide(ctx, [L.INT64, L.INT32, L.INT32], 300_000, 30, seed=7, device=False)
    rng = np.random.default_rng(8)
    sel = np.sort(rng.choice(500_000, 200_000, replace=False)).astype(np.uint32)
    check_wide(ctx, [L.INT64, L.INT32, L.INT32], 500_000, 30, seed=8, sel=sel, fnull=0.2)


def test_merge_of_two_dictionaries(ctx):
    """MergeSink of two wide aggregates whose dictionaries hand out different
    ids for the same pairs; then the merged state equals one pass."""
    print("[wide] merge", flush=True)
    rng = np.random.default_rng(9)
    n = 400_000
    key_types = [L.INT64, L.INT64]
    keys, v, d, f, _ = make_batch(key_types, n, rng, 300)
    types = key_types + [L.INT64, L.DECIMAL, L.FLOAT64]
    scales = [0, 0, 0, 2, 0]
    aggs = AGGS_OF(2)
    arrs = keys + [v, d, f]
    h = n // 3
    a1 = HashAggregate(ctx, types, [0, 1], aggs, scales=scales)
    a2 = HashAggregate(ctx, types, [0, 1], aggs, scales=scales)
    a1.consume([Col(ctx.to_device(a[:h]), t, None, s) for a, t, s in zip(arrs, types, scales)], h)
    a2.consume([Col(ctx.to_device(a[h:][::-1].copy()), t, None, s) for a, t, s in zip(arrs, types, scales)], n - h)
    a1.merge(a2)
    a1.finalize()
    hcols = [(a, t, None, s) for a, t, s in zip(arrs, types, scales)]
    ref = oracle_groups(run_oracle(hcols, [0, 1], aggs))
    compare(gpu_groups(a1), ref, [(x.func, x.col) for x in aggs], types, scales, abs_sums(hcols, [0, 1], aggs))


def test_spill_run_and_merge_back(ctx):
    """The partial-state run of a 128-bit-key aggregate is byte-identical to
    the oracle's; merging it into a fresh aggregate twice doubles every
    integer accumulator."""
    print("[wide] spill", flush=True)
    rng = np.random.default_rng(10)
    n = 200_000
    a = key_col(L.INT64, n, rng, 500)
    b = key_col(L.INT32, n, rng, 7)
    c = key_col(L.INT32, n, rng, 3)
    v = rng.integers(-10**12, 10**12, n).astype(np.int64)
    types = [L.INT64, L.INT32, L.INT32, L.INT64]
    aggs = [AggSpec(L.SUM, 3), AggSpec(L.COUNT, -1), AggSpec(L.MIN, 3), AggSpec(L.MAX, 3)]
    agg = HashAggregate(ctx, types, [0, 1, 2], aggs)
    agg.consume([Col(ctx.to_device(x), t) for x, t in zip((a, b, c, v), types)], n)
    agg.finalize()
    got = agg.spill_run(["a", "b", "c"])
    res = O.hash_agg([(a, L.INT64), (b, L.INT32), (c, L.INT32), (v, L.INT64)], [0, 1, 2],
                     [(x.func, x.col) for x in aggs])
    specs = [O.partial_spec(x.func, None if x.col < 0 else types[x.col]) for x in aggs]
    want = O.partial_run(res, specs, [L.INT64, L.INT32, L.INT32], names=["a", "b", "c"])
    assert got == want
    back = HashAggregate(ctx, types, [0, 1, 2], aggs)
    back.merge_run(got)
    back.merge_run(got)
    back.finalize()
    g = gpu_groups(back)
    r = oracle_groups(res)
    np.testing.assert_array_equal(g["keys"], r["keys"])
    np.testing.assert_array_equal(g["values"][1], 2 * r["acc"][:, 1]["count"])
    np.testing.assert_array_equal(g["state"][0]["sum_lo"].view(np.int64), 2 * r["acc"][:, 0]["sum_i64"])


def test_reset_and_rejections(ctx):
    """Reset drops groups and dictionary; a NULL in a wide key is handed back
    (NullKeyError, no state changed); 20 bytes of key are not on the path."""
    rng = np.random.default_rng(11)
    n = 10_000
    a = key_col(L.INT64, n, rng, 50)
    b = key_col(L.INT64, n, rng, 50)
    types = [L.INT64, L.INT64]
    agg = HashAggregate(ctx, types, [0, 1], [AggSpec(L.COUNT, -1)], nullable=[True, False])
    cols = [Col(ctx.to_device(a), L.INT64), Col(ctx.to_device(b), L.INT64)]
    agg.consume(cols, n)
    agg.finalize()
    g1 = agg.num_groups()
    agg.reset()
    agg.consume(cols, n // 2)
    agg.finalize()
    assert 0 < agg.num_groups() <= g1
    agg.reset()
    m = bitmap_from_mask(np.arange(n) != 17)
    with pytest.raises(NullKeyError):
        agg.consume([Col(ctx.to_device(a), L.INT64, ctx.to_device(m)), Col(ctx.to_device(b), L.INT64)], n)
    agg.finalize()
    assert agg.num_groups() == 0
    with pytest.raises(UnsupportedError):
        HashAggregate(ctx, [L.INT64, L.INT64, L.INT32], [0, 1, 2], [AggSpec(L.COUNT, -1)])
