"""The GPU aggregate's partial-state spill run (wg_agg_spill_run) against the
oracle's run of the same rows: byte-identical when every accumulator is an
integer or Int128, and equal up to float-sum order otherwise."""
import numpy as np
import pytest

import oracle as O
from helpers import FLOAT_RTOL
from wadjet_amd import _lib as L
from wadjet_amd.gpu import AggSpec, Col, HashAggregate, bitmap_from_mask

pytestmark = pytest.mark.gpu


def run_both(ctx, host_cols, types, keys, aggs, nullable=None, scales=None, names=None):
    dev = []
    for arr, t, *rest in host_cols:
        v = rest[0] if rest else None
        dev.append(Col(ctx.to_device(arr), t, None if v is None else ctx.to_device(v), rest[1] if len(rest) > 1 else 0))
    n = len(host_cols[0][0])
    agg = HashAggregate(ctx, types, keys, aggs, scales=scales, nullable=nullable)
    agg.consume(dev, n)
    agg.finalize()
    got = agg.spill_run(names)
    oaggs = [(a.func, a.col) for a in aggs]
    res = O.hash_agg(host_cols, keys, oaggs)
    specs = [O.partial_spec(a.func, None if a.col < 0 else types[a.col], (scales or [0] * len(types))[a.col])
             for a in aggs]
    want = O.partial_run(res, specs, [types[k] for k in keys], names=names)
    agg.close()
    return got, want


def test_integer_and_decimal_run_is_byte_identical(ctx):
    """int32 composite key; SUM int64, COUNT(*), COUNT(col) with NULLs,
    MIN/MAX int, SUM DECIMAL(18,2) whose group sums pass 2^64."""
    print("[spill] integer/decimal run", flush=True)
    rng = np.random.default_rng(11)
    n = 400_000
    a = rng.integers(-3000, 3000, n).astype(np.int32)
    b = rng.integers(0, 7, n).astype(np.int32)
    v = rng.integers(-10**12, 10**12, n).astype(np.int64)
    vvalid = rng.random(n) > 0.1
    d = rng.integers(9 * 10**17, 9 * 10**18, n).astype(np.int64)
    types = [L.INT32, L.INT32, L.INT64, L.DECIMAL]
    hcols = [(a, L.INT32), (b, L.INT32), (v, L.INT64, bitmap_from_mask(vvalid)), (d, L.DECIMAL, None, 2)]
    aggs = [AggSpec(L.SUM, 2), AggSpec(L.COUNT, -1), AggSpec(L.COUNT, 2), AggSpec(L.MIN, 2), AggSpec(L.MAX, 2),
            AggSpec(L.SUM, 3)]
    got, want = run_both(ctx, hcols, types, [0, 1], aggs, nullable=[False, False, True, False],
                         scales=[0, 0, 0, 2], names=["a", "b"])
    assert O.read_partial_run(got)["n"] == 6000 * 7
    assert got == want


def test_null_and_sentinel_groups_int64_key(ctx):
    """A nullable int64 key: the NULL group ("<null>", tag NULL) and the
    sentinel-valued key sort and encode like every other group."""
    print("[spill] NULL and sentinel groups", flush=True)
    rng = np.random.default_rng(12)
    n = 200_000
    k = rng.integers(-500, 500, n).astype(np.int64)
    k[::997] = -0x7FFFFFFFFFFFFFFF
    kvalid = rng.random(n) > 0.05
    v = rng.integers(0, 1000, n).astype(np.int32)
    types = [L.INT64, L.INT32]
    hcols = [(k, L.INT64, bitmap_from_mask(kvalid)), (v, L.INT32)]
    aggs = [AggSpec(L.SUM, 1), AggSpec(L.AVG, 1), AggSpec(L.MAX, 1)]
    got, want = run_both(ctx, hcols, types, [0], aggs, nullable=[True, False], names=["k"])
    r = O.read_partial_run(got)
    assert any(kv == [None] for _, kv, _ in r["groups"])
    assert got == want


def test_float_sums_equal_up_to_order(ctx):
    """SUM/AVG over float64 and AVG over int64 are float sums: the runs agree
    in every byte but those sums, which agree within the stated tolerance."""
    print("[spill] float sums", flush=True)
    rng = np.random.default_rng(13)
    n = 300_000
    k = rng.integers(0, 2000, n).astype(np.int32)
    f = rng.normal(0, 1, n)
    w = rng.integers(-10**9, 10**9, n).astype(np.int64)
    types = [L.INT32, L.FLOAT64, L.INT64]
    hcols = [(k, L.INT32), (f, L.FLOAT64), (w, L.INT64)]
    aggs = [AggSpec(L.SUM, 1), AggSpec(L.MIN, 1), AggSpec(L.MAX, 1), AggSpec(L.AVG, 2), AggSpec(L.COUNT, -1)]
    got, want = run_both(ctx, hcols, types, [0], aggs)
    rg, rw = O.read_partial_run(got), O.read_partial_run(want)
    assert rg["cols"] == rw["cols"] and rg["specs"] == rw["specs"] and rg["n"] == rw["n"] == 2000
    absf = np.zeros(2000)
    np.add.at(absf, k, np.abs(f))
    absw = np.zeros(2000)
    np.add.at(absw, k, np.abs(w.astype(np.float64)))
    for (skg, kvg, ag), (skw, kvw, aw) in zip(rg["groups"], rw["groups"]):
        assert skg == skw and kvg == kvw
        key = kvg[0]
        for j, (x, y) in enumerate(zip(ag, aw)):
            if j in (0, 3):
                assert x["count"] == y["count"]
                lim = FLOAT_RTOL * (absf[key] if j == 0 else absw[key])
                assert abs(x["sum"] - y["sum"]) <= lim, (key, j, x, y)
            else:
                assert x == y, (key, j)


def _mixed_rows(seed, n):
    rng = np.random.default_rng(seed)
    k = rng.integers(-400, 400, n).astype(np.int64)
    k[::1009] = -0x7FFFFFFFFFFFFFFF
    kvalid = rng.random(n) > 0.03
    v = rng.integers(-10**12, 10**12, n).astype(np.int64)
    d = rng.integers(9 * 10**17, 9 * 10**18, n).astype(np.int64)
    f = rng.normal(0, 1, n)
    return k, kvalid, v, d, f


MIX_TYPES = [L.INT64, L.INT64, L.DECIMAL, L.FLOAT64]
MIX_AGGS = [AggSpec(L.SUM, 1), AggSpec(L.COUNT, -1), AggSpec(L.MIN, 1), AggSpec(L.MAX, 3), AggSpec(L.SUM, 2),
            AggSpec(L.AVG, 1), AggSpec(L.SUM, 3)]


def _dev_cols(ctx, k, kvalid, v, d, f):
    return [Col(ctx.to_device(k), L.INT64, ctx.to_device(bitmap_from_mask(kvalid))), Col(ctx.to_device(v), L.INT64),
            Col(ctx.to_device(d), L.DECIMAL, None, 2), Col(ctx.to_device(f), L.FLOAT64)]


def _compare_all(ctx, agg, parts):
    from helpers import compare, gpu_groups, oracle_groups
    k, kvalid, v, d, f = (np.concatenate(x) for x in zip(*parts))
    hcols = [(k, L.INT64, bitmap_from_mask(kvalid)), (v, L.INT64), (d, L.DECIMAL, None, 2), (f, L.FLOAT64)]
    oaggs = [(a.func, a.col) for a in MIX_AGGS]
    ref = oracle_groups(O.hash_agg(hcols, [0], oaggs))
    keys = np.where(kvalid, k, 0)
    gk = np.unique(ref["keys"][:, 0][~ref["nulls"][:, 0]])
    absf = [None] * len(MIX_AGGS)
    for j, a in enumerate(MIX_AGGS):
        if a.func in (L.SUM, L.AVG) and MIX_TYPES[a.col] in (L.FLOAT64,) or (a.func == L.AVG and a.col == 1):
            vals = np.abs(f) if a.col == 3 else np.abs(v.astype(np.float64))
            s = np.zeros(len(ref["keys"]))
            # oracle group order: key-sorted with the NULL group where _order puts it
            order = {(bool(nl), int(kk)): g for g, (kk, nl) in enumerate(zip(ref["keys"][:, 0], ref["nulls"][:, 0]))}
            for kk, ok, x in zip(keys, kvalid, vals):
                s[order[(not ok, int(kk) if ok else 0)]] += x
            absf[j] = s
    compare(gpu_groups(agg), ref, oaggs, MIX_TYPES, [0, 0, 2, 0], abs_sums=absf)
    assert len(gk) > 0


def test_spill_then_merge_round_trip(ctx):
    """Aggregate A spills its run; aggregate B consumes other rows, merges the
    run, consumes more and finalizes: equal to the oracle over all rows
    (NULL and sentinel-valued groups, Int128 sums, float sums in tolerance)."""
    print("[spill] spill -> merge round trip", flush=True)
    parts = [_mixed_rows(21 + i, 60_000) for i in range(3)]
    a = HashAggregate(ctx, MIX_TYPES, [0], MIX_AGGS, scales=[0, 0, 2, 0], nullable=[True, False, False, False])
    a.consume(_dev_cols(ctx, *parts[0]), 60_000)
    a.finalize()
    run = a.spill_run(["k"])
    a.reset()  # the spill drains the state, as drainStateToRuns does
    b = HashAggregate(ctx, MIX_TYPES, [0], MIX_AGGS, scales=[0, 0, 2, 0], nullable=[True, False, False, False])
    b.consume(_dev_cols(ctx, *parts[1]), 60_000)
    b.merge_run(run)
    b.consume(_dev_cols(ctx, *parts[2]), 60_000)
    b.finalize()
    _compare_all(ctx, b, parts)
    a.close()
    b.close()


def test_merge_a_run_written_by_the_cpu_operator(ctx):
    """A partial computed on the CPU (the oracle's aggregate written as a run)
    merges into a GPU aggregate: the result equals the oracle over all rows."""
    print("[spill] merge a CPU-written run", flush=True)
    parts = [_mixed_rows(31 + i, 50_000) for i in range(2)]
    k, kvalid, v, d, f = parts[0]
    hcols = [(k, L.INT64, bitmap_from_mask(kvalid)), (v, L.INT64), (d, L.DECIMAL, None, 2), (f, L.FLOAT64)]
    res = O.hash_agg(hcols, [0], [(x.func, x.col) for x in MIX_AGGS])
    specs = [O.partial_spec(x.func, None if x.col < 0 else MIX_TYPES[x.col], 2 if x.col == 2 else 0)
             for x in MIX_AGGS]
    run = O.partial_run(res, specs, [L.INT64], names=["k"])
    g = HashAggregate(ctx, MIX_TYPES, [0], MIX_AGGS, scales=[0, 0, 2, 0], nullable=[True, False, False, False])
    g.merge_run(run)
    g.consume(_dev_cols(ctx, *parts[1]), 50_000)
    g.finalize()
    _compare_all(ctx, g, parts)
    g.close()


def test_merge_run_rejects_another_query(ctx):
    """A run of different aggregates is refused before any state changes."""
    from wadjet_amd._lib import WgError
    k = np.arange(1000, dtype=np.int64) % 10
    a = HashAggregate(ctx, [L.INT64, L.INT64], [0], [AggSpec(L.SUM, 1)])
    a.consume([Col(ctx.to_device(k), L.INT64), Col(ctx.to_device(k), L.INT64)], 1000)
    a.finalize()
    run = a.spill_run()
    b = HashAggregate(ctx, [L.INT64, L.INT64], [0], [AggSpec(L.MAX, 1)])
    with pytest.raises(WgError):
        b.merge_run(run)
    with pytest.raises(WgError):
        b.merge_run(run[:-3])
    a.close()
    b.close()


def test_gpu_run_bytes_known_answer(ctx):
    """The GPU run of a hand-checked input equals bytes assembled field by
    field from the reference writer (writeHeader / Write / emitAcc / Close,
    aggregate_partial_spill.go:171-351): GROUP BY k int64 SUM(v), COUNT(*),
    MIN(f), SUM(d DECIMAL(10,2)), MAX(d). MAX over a DECIMAL column is NULL on
    the flat accumulators (agg_scatter.go:555-576), so its HasMax byte is 0."""
    print("[spill] byte-level known answer", flush=True)
    import struct

    def le(n, x, signed=False):
        return int(x).to_bytes(n, "little", signed=signed)

    k = np.array([10, 9, 10, 10], np.int64)
    v = np.array([-10, 40, 1, 2], np.int64)
    f = np.array([0.5, 7.25, -1.5, 2.0], np.float64)
    d = np.array([2**62, 5, 2**62, -(2**60)], np.int64)
    dev = [Col(ctx.to_device(k), L.INT64), Col(ctx.to_device(v), L.INT64), Col(ctx.to_device(f), L.FLOAT64),
           Col(ctx.to_device(d), L.DECIMAL, None, 2)]
    aggs = [AggSpec(L.SUM, 1), AggSpec(L.COUNT, -1), AggSpec(L.MIN, 2), AggSpec(L.SUM, 3), AggSpec(L.MAX, 3)]
    agg = HashAggregate(ctx, [L.INT64, L.INT64, L.FLOAT64, L.DECIMAL], [0], aggs, scales=[0, 0, 0, 2])
    agg.consume(dev, 4)
    agg.finalize()
    got = agg.spill_run(["k"])
    agg.close()
    want = bytearray(b"WAGS\x01") + le(4, 1) + le(2, 1) + b"k" + le(2, O.PARQUET_INT64) + le(4, 5)
    for func, is_f, is_d, sc in [(0, 0, 0, 0), (1, 0, 0, 0), (2, 1, 0, 0), (0, 0, 1, 2), (3, 0, 1, 0)]:
        want += le(4, func) + bytes([is_f, is_d]) + le(4, sc, True)
    want += le(4, 2)
    dsum = 2**62 + 2**62 - 2**60  # group 10's DECIMAL sum (fits 64 bits, not 63)
    want += b"\x01" + le(4, 2) + b"10" + b"\x03" + le(8, 10, True)
    want += le(8, 3, True) + le(8, -7, True) + le(8, 3, True) + b"\x01" + struct.pack("<d", -1.5)
    want += le(8, 3, True) + le(8, 0, True) + le(8, dsum) + b"\x00"
    want += b"\x01" + le(4, 1) + b"9" + b"\x03" + le(8, 9, True)
    want += le(8, 1, True) + le(8, 40, True) + le(8, 1, True) + b"\x01" + struct.pack("<d", 7.25)
    want += le(8, 1, True) + le(8, 0, True) + le(8, 5) + b"\x00"
    want += b"\x00"
    assert got == bytes(want)
