"""GPU parity of the selection-vector producer (wg_filter, KernelFilter.Execute
in internal/engine/exec/filter.go:373-425 over the typed compare kernels of
kernel/compare.go:47-178) and of the predicates fused into the aggregate's
consume, against the CPU oracle (or_filter, or_hash_agg).

The selection vectors must be identical: same rows, ascending. Cases: the six
operators on every storage type the filter takes (int32, int64, uint16 port,
uint32 compact IPv4, int64 IPv4, float64), NULL rows (never pass), NaN and
signed zeros and infinities in float columns, constants at and beyond the
column's extremes, an incoming selection vector (compaction of an existing
Sel, filter.go:408-418), and the everything-filtered result (count 0)."""
import numpy as np
import pytest

import oracle as O
from helpers import compare, gpu_groups, oracle_groups
from wadjet_amd import _lib as L
from wadjet_amd.gpu import AggSpec, Col, HashAggregate, Pred, bitmap_from_mask, filter_sel

pytestmark = pytest.mark.gpu

OPS = [L.EQ, L.NE, L.LT, L.LE, L.GT, L.GE]
NP = {L.INT32: np.int32, L.INT64: np.int64, L.UINT16: np.uint16, L.UINT32: np.uint32, L.IPV4: np.int64,
      L.FLOAT64: np.float64}


def column(rng, t, n):
    """Values with many repeats of the constants the tests compare against."""
    if t == L.FLOAT64:
        v = rng.standard_normal(n)
        v[rng.random(n) < 0.05] = np.nan
        v[rng.random(n) < 0.02] = 0.0
        v[rng.random(n) < 0.02] = -0.0
        v[rng.random(n) < 0.01] = np.inf
        v[rng.random(n) < 0.01] = -np.inf
        v[rng.random(n) < 0.05] = 0.25
        return v
    if t == L.INT32:
        v = rng.integers(-1000, 1000, n).astype(np.int32)
        v[rng.random(n) < 0.01] = np.iinfo(np.int32).min
        v[rng.random(n) < 0.01] = np.iinfo(np.int32).max
        return v
    if t == L.UINT16:
        return rng.integers(0, 2**16, n).astype(np.uint16)
    if t == L.UINT32:
        v = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
        v[rng.random(n) < 0.05] = 0xC0A80001  # 192.168.0.1
        return v
    v = rng.integers(-2**40, 2**40, n).astype(np.int64)
    if t == L.IPV4:
        v = rng.integers(0, 2**32, n).astype(np.int64)
        v[rng.random(n) < 0.05] = 0xC0A80001
        return v
    v[rng.random(n) < 0.01] = np.iinfo(np.int64).min
    v[rng.random(n) < 0.01] = np.iinfo(np.int64).max
    v[rng.random(n) < 0.05] = 12345
    return v


def constants(t):
    """(i64, f64) constants: a common value, the extremes, beyond the range."""
    if t == L.FLOAT64:
        return [(0, 0.25), (0, 0.0), (0, -0.0), (0, np.nan), (0, np.inf), (0, -np.inf), (0, -0.5)]
    if t == L.INT32:
        # int32 columns compare against int32(constant) (ResolveFilterKernel,
        # compare.go:152-153): 2^32 + 7 truncates to 7
        return [(7, 0.0), (-2**31, 0.0), (2**31 - 1, 0.0), (2**32 + 7, 0.0), (0, 0.0)]
    if t == L.UINT16:
        return [(80, 0.0), (0, 0.0), (65535, 0.0), (70000, 0.0)]
    if t in (L.UINT32, L.IPV4):
        return [(0xC0A80001, 0.0), (0, 0.0), (2**32 - 1, 0.0), (-1, 0.0)]
    return [(12345, 0.0), (-2**63, 0.0), (2**63 - 1, 0.0), (0, 0.0)]


def run_both(ctx, vals, t, valid, n, op, ci, cf, sel=None):
    dcol = Col(ctx.to_device(vals), t, None if valid is None else ctx.to_device(bitmap_from_mask(valid)))
    hcol = (vals, t, None if valid is None else bitmap_from_mask(valid))
    if sel is None:
        out, k = filter_sel(ctx, dcol, n, op, ci, cf)
    else:
        dsel = ctx.to_device(sel)
        out, k = filter_sel(ctx, dcol, n, op, ci, cf, sel=dsel, n_sel=len(sel))
    got = out.to_host(np.uint32, k) if k else np.zeros(0, np.uint32)
    want = O.filter_rows(hcol, n, op, ci, cf, sel=sel)
    return got, want


@pytest.mark.parametrize("t", [L.INT32, L.INT64, L.UINT16, L.UINT32, L.IPV4, L.FLOAT64])
def test_filter_every_op_and_constant(ctx, t):
    """Six operators x typed constants, with NULL rows and an incoming Sel."""
    rng = np.random.default_rng(100 + t)
    n = 300_001  # not a multiple of any tile
    vals = column(rng, t, n)
    valid = rng.random(n) > 0.1
    sel = np.sort(rng.choice(n, n // 3, replace=False)).astype(np.uint32)
    for ci, cf in constants(t):
        for op in OPS:
            for vmask in (None, valid):
                for s in (None, sel):
                    print(f"[filter] type={t} op={op} const=({ci},{cf}) nulls={vmask is not None} "
                          f"sel={s is not None}", flush=True)
                    got, want = run_both(ctx, vals, t, vmask, n, op, ci, cf, s)
                    assert len(got) == len(want), (t, op, ci, cf, len(got), len(want))
                    np.testing.assert_array_equal(got, want)


def test_filter_everything_filtered_and_empty(ctx):
    """No row passes (count 0, filter.go returns nil), an empty Sel, zero rows,
    an all-NULL column, and a NaN constant under LT."""
    rng = np.random.default_rng(7)
    n = 70_000
    v = rng.integers(0, 100, n).astype(np.int64)
    for op, ci in ((L.LT, 0), (L.GT, 99), (L.EQ, 1000)):
        got, want = run_both(ctx, v, L.INT64, None, n, op, ci, 0.0)
        assert len(got) == 0 and len(want) == 0
    got, want = run_both(ctx, v, L.INT64, None, n, L.GE, 0, 0.0, sel=np.zeros(0, np.uint32))
    assert len(got) == 0 and len(want) == 0
    got, want = run_both(ctx, v, L.INT64, np.zeros(n, bool), n, L.NE, 5, 0.0)
    assert len(got) == 0 and len(want) == 0
    f = rng.standard_normal(n)
    got, want = run_both(ctx, f, L.FLOAT64, None, n, L.LT, 0, np.nan)
    assert len(got) == 0 and len(want) == 0
    got, want = run_both(ctx, f, L.FLOAT64, None, n, L.NE, 0, np.nan)  # NaN != x holds for every row
    assert len(got) == n and len(want) == n
    np.testing.assert_array_equal(got, want)


def test_filter_large_selective(ctx):
    """4M rows (many tiles) at selectivities 0.1 %, 69 % and 99.9 %."""
    rng = np.random.default_rng(11)
    n = 4_000_000
    f = rng.standard_normal(n)
    valid = rng.random(n) > 0.001
    for c in (-3.09, -0.5, 3.09):
        got, want = run_both(ctx, f, L.FLOAT64, valid, n, L.GT, 0, c)
        np.testing.assert_array_equal(got, want)


def _agg_case(ctx, rng, n, pred_type, op, ci, nullable_pred):
    k = rng.integers(0, 3000, n).astype(np.int64)
    v = rng.integers(-10**9, 10**9, n).astype(np.int64)
    p = column(rng, pred_type, n)
    pvalid = (rng.random(n) > 0.07) if nullable_pred else None
    types = [L.INT64, L.INT64, pred_type]
    aggs = [AggSpec(L.SUM, 1), AggSpec(L.COUNT, -1), AggSpec(L.MIN, 1), AggSpec(L.MAX, 1)]
    pm = None if pvalid is None else bitmap_from_mask(pvalid)
    cols = [Col(ctx.to_device(k), L.INT64), Col(ctx.to_device(v), L.INT64),
            Col(ctx.to_device(p), pred_type, None if pm is None else ctx.to_device(pm))]
    return k, v, p, pm, types, aggs, cols


@pytest.mark.parametrize("pred_type", [L.INT32, L.INT64, L.UINT16, L.FLOAT64])
@pytest.mark.parametrize("op", [L.LT, L.EQ, L.NE, L.GE])
@pytest.mark.parametrize("nullable_pred", [False, True])
def test_fused_predicate_in_consume(ctx, pred_type, op, nullable_pred):
    """The predicate fused into wg_agg_consume on int and float columns (the
    typed count pass and the general loaders) == or_hash_agg with the same
    predicate; NULL predicate values drop the row."""
    rng = np.random.default_rng(1000 + 10 * pred_type + op + (5 if nullable_pred else 0))
    n = 1_500_000
    k, v, p, pm, types, aggs, cols = _agg_case(ctx, rng, n, pred_type, op, 0, nullable_pred)
    ci, cf = constants(pred_type)[0]
    print(f"[fused] pred type={pred_type} op={op} const=({ci},{cf}) nullable={nullable_pred}", flush=True)
    for strategy in (1, 2):
        agg = HashAggregate(ctx, types, [0], aggs, nullable=[False, False, nullable_pred], group_hint=3000)
        agg.set_strategy(strategy)
        agg.consume(cols, n, pred=Pred(2, op, ci, cf))
        agg.finalize()
        hcols = [(k, L.INT64), (v, L.INT64), (p, pred_type, pm)]
        ref = oracle_groups(O.hash_agg(hcols, [0], [(a.func, a.col) for a in aggs], pred=(2, op, ci, cf)))
        compare(gpu_groups(agg), ref, [(a.func, a.col) for a in aggs], types)
        agg.close()


def test_two_fused_predicates_and_sel(ctx):
    """Two ANDed predicates plus an incoming Sel (a chain of KernelFilters
    through one Sel) == the oracle's filter followed by its predicate."""
    rng = np.random.default_rng(77)
    n = 1_200_000
    k = rng.integers(0, 50_000, n).astype(np.int32)
    q = rng.integers(1, 51, n).astype(np.int32)
    f = rng.standard_normal(n)
    fvalid = rng.random(n) > 0.05
    types = [L.INT32, L.INT32, L.FLOAT64]
    aggs = [AggSpec(L.SUM, 1), AggSpec(L.COUNT, -1), AggSpec(L.AVG, 2), AggSpec(L.MAX, 2)]
    fm = bitmap_from_mask(fvalid)
    cols = [Col(ctx.to_device(k), L.INT32), Col(ctx.to_device(q), L.INT32), Col(ctx.to_device(f), L.FLOAT64,
                                                                                 ctx.to_device(fm))]
    sel = np.sort(rng.choice(n, n // 2, replace=False)).astype(np.uint32)
    agg = HashAggregate(ctx, types, [0], aggs, nullable=[False, False, True], group_hint=50_000)
    agg.consume(cols, n, sel=ctx.to_device(sel), pred=[Pred(1, L.LE, 25), Pred(2, L.GT, 0, -0.5)])
    agg.finalize()
    hcols = [(k, L.INT32), (q, L.INT32), (f, L.FLOAT64, fm)]
    s1 = O.filter_rows(hcols[1], n, L.LE, 25, sel=sel)
    ref = oracle_groups(O.hash_agg(hcols, [0], [(a.func, a.col) for a in aggs], sel=s1, pred=(2, L.GT, 0, -0.5)))
    abs_f = O.hash_agg(hcols + [(np.abs(f), L.FLOAT64, fm)], [0], [(O.SUM, 3)], sel=s1, pred=(2, L.GT, 0, -0.5))
    compare(gpu_groups(agg), ref, [(a.func, a.col) for a in aggs], types,
            abs_sums=[None, None, oracle_groups(abs_f)["acc"][:, 0]["sum_f64"], None])
