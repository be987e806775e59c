"""GPU parity of the shuffle routing (wg_shuffle_partition) against the oracle's
restatement of hashRowsIntoPartitions + perPartRows
(internal/worker/partitioned_shuffle_sink.go:199-290, :1039-1171): the
partition of every active row, the rows grouped by partition in input order,
and the offsets — all bit-exact."""
import numpy as np
import pytest

import oracle as O
from wadjet_amd import _lib as L
from wadjet_amd._lib import WgError
from wadjet_amd.gpu import Col, bitmap_from_mask, shuffle_partition

pytestmark = pytest.mark.gpu

TYPES = {"int64": L.INT64, "int32": L.INT32, "uint16": L.UINT16, "uint32": L.UINT32, "ipv4": L.IPV4,
         "float64": L.FLOAT64, "decimal": L.DECIMAL, "decimal128": L.DECIMAL128}


def make_col(kind, n, rng, null_p=0.0):
    if kind == "int64":
        v = rng.integers(-2**63, 2**63 - 1, n, dtype=np.int64)
    elif kind == "int32":
        v = rng.integers(-2**31, 2**31 - 1, n, dtype=np.int64).astype(np.int32)
    elif kind == "uint16":
        v = rng.integers(0, 65536, n).astype(np.uint16)
    elif kind == "uint32":
        v = rng.integers(0, 2**32, n, dtype=np.int64).astype(np.uint32)
    elif kind == "ipv4":
        v = rng.integers(0, 2**32, n, dtype=np.int64)
    elif kind == "float64":
        v = rng.normal(size=n)
        if n >= 4:
            v[:4] = [0.0, -0.0, np.nan, -np.inf]
    elif kind == "decimal":
        v = rng.integers(-10**15, 10**15, n, dtype=np.int64)
    else:
        v = rng.integers(-2**62, 2**62, 2 * n, dtype=np.int64)
    valid = bitmap_from_mask(rng.random(n) >= null_p) if null_p > 0 else None
    return v, TYPES[kind], valid


def run_both(ctx, host_cols, key_cols, num_parts, n, sel=None):
    dev = [Col(ctx.to_device(v), t, ctx.to_device(m) if m is not None else None) for v, t, m in host_cols]
    dsel = ctx.to_device(sel) if sel is not None else None
    parts, rows, offs = shuffle_partition(ctx, dev, key_cols, num_parts, n, dsel, len(sel) if sel is not None else 0)
    m = len(sel) if sel is not None else n
    ocols = [(v if t != L.DECIMAL128 else v[::2], t, mk) for v, t, mk in host_cols]
    eparts, erows, eoffs = O.shuffle_partition(ocols, key_cols, num_parts, sel, n)
    assert np.array_equal(offs, eoffs), "offsets"
    assert np.array_equal(parts.to_host(np.uint32, m), eparts), "partition ids"
    assert np.array_equal(rows.to_host(np.uint32, m), erows), "grouped rows"
    return eoffs


@pytest.mark.parametrize("kind", list(TYPES))
@pytest.mark.parametrize("null_p", [0.0, 0.2])
def test_each_key_type(ctx, kind, null_p):
    print(f"[shuffle] type {kind} nulls {null_p}", flush=True)
    rng = np.random.default_rng(hash((kind, null_p)) % 2**32)
    n = 10_007
    run_both(ctx, [make_col(kind, n, rng, null_p)], [0], 24, n)


@pytest.mark.parametrize("num_parts", [0, 1, 2, 3, 4, 7, 24, 64, 100, 1000, 4096, 5000, 8192, 12_345, 16_384])
def test_partition_counts(ctx, num_parts):
    """Powers of two (mask) and others (32-bit remainder form), the
    workgroup shapes for up to 4096, 8192 and 16384 partitions, and 0."""
    print(f"[shuffle] {num_parts} partitions", flush=True)
    rng = np.random.default_rng(num_parts)
    n = 300_001
    cols = [make_col("int64", n, rng, 0.01), make_col("int32", n, rng)]
    offs = run_both(ctx, cols, [0, 1], num_parts, n)
    assert offs[-1] == n


def test_multi_column_keys_with_selection(ctx):
    """Five key columns of mixed widths with NULLs, through an ascending Sel:
    rows are Sel values, parts indexed by active position."""
    print("[shuffle] multi-column + sel", flush=True)
    rng = np.random.default_rng(5)
    n = 200_000
    cols = [make_col(k, n, rng, 0.05) for k in ("ipv4", "uint16", "int32", "float64", "decimal")]
    sel = np.sort(rng.choice(n, 123_457, replace=False)).astype(np.uint32)
    run_both(ctx, cols, [0, 1, 2, 3, 4], 37, n, sel)
    run_both(ctx, cols, [3, 0], 64, n, sel)


@pytest.mark.parametrize("n", [0, 1, 5, 63, 64, 65, 1023, 1025, 4097])
def test_ragged_and_tiny_batches(ctx, n):
    print(f"[shuffle] n={n}", flush=True)
    rng = np.random.default_rng(n + 11)
    cols = [make_col("int64", max(n, 1), rng, 0.1)]
    offs = run_both(ctx, cols, [0], 5, n)
    assert offs[-1] == n


def test_no_key_columns(ctx):
    rng = np.random.default_rng(3)
    n = 70_000
    sel = np.sort(rng.choice(n, 5000, replace=False)).astype(np.uint32)
    run_both(ctx, [make_col("int64", n, rng)], [], 8, n, sel)


def test_skewed_keys(ctx):
    """One key value carries most rows (long runs of one partition per
    wavefront step) and a handful of others."""
    print("[shuffle] skew", flush=True)
    rng = np.random.default_rng(9)
    n = 1_000_000
    k = np.where(rng.random(n) < 0.9, 42, rng.integers(0, 10, n)).astype(np.int64)
    run_both(ctx, [(k, L.INT64, None)], [0], 16, n)


def test_full_size_properties(ctx):
    """100M rows of (a int32 Zipf-like, b int32): every row routed once, each
    partition's rows ascending and hashing to it (checked on a sample
    against the oracle), counts equal the oracle's bincount."""
    print("[shuffle] 100M rows", flush=True)
    from wadjet_amd import workloads as W
    n = 100_000_000
    cols = W.gen_zipf_groupby(ctx, n)
    parts, rows, offs = shuffle_partition(ctx, cols[:2], [0, 1], 256, n)
    assert offs[-1] == n and np.all(np.diff(offs) >= 0)
    r = rows.to_host(np.uint32, n)
    p = parts.to_host(np.uint32, n)
    assert np.array_equal(np.bincount(p, minlength=256), np.diff(offs))
    seg = np.repeat(np.arange(256, dtype=np.uint32), np.diff(offs).astype(np.int64))
    assert np.array_equal(p[r], seg), "row grouped under a partition it does not hash to"
    within = np.diff(r.astype(np.int64))
    starts = offs[1:-1]
    ok = within > 0
    ok[starts[(starts > 0) & (starts < n)] - 1] = True  # partition boundaries may step down
    assert ok.all(), "rows out of input order within a partition"
    del r
    a = cols[0].data.to_host(np.int32, n)
    b = cols[1].data.to_host(np.int32, n)
    idx = np.random.default_rng(0).choice(n, 200_000, replace=False)
    h = O.shuffle_hash([(a[idx], O.INT32), (b[idx], O.INT32)], [0, 1])
    assert np.array_equal((h % np.uint64(256)).astype(np.uint32), p[idx])


def test_errors(ctx):
    rng = np.random.default_rng(1)
    cols = [Col(ctx.to_device(make_col("int64", 100, rng)[0]), L.INT64)]
    with pytest.raises(WgError):
        shuffle_partition(ctx, cols, [0], 16_385, 100)
    with pytest.raises(WgError):
        shuffle_partition(ctx, cols, [3], 4, 100)


def test_state_bytes_tracks_groups(ctx):
    """wg_agg_state_bytes (HashAggregate.StateBytes, the bound the capped
    partial aggregate flushes on): 0 before any row, groups x per-group state
    after host and device batches, the same after finalize, 0 after reset."""
    from wadjet_amd.gpu import AggSpec, HashAggregate
    agg = HashAggregate(ctx, [L.INT64, L.INT64], [0], [AggSpec(L.SUM, 1), AggSpec(L.COUNT, -1)], group_hint=16)
    assert agg.state_bytes() == 0
    k = np.arange(1000, dtype=np.int64) % 300
    agg.consume([Col(k, L.INT64), Col(k, L.INT64)], 1000, device=False)  # staged host batch
    b1 = agg.state_bytes()
    assert b1 > 0 and b1 % 300 == 0
    k2 = (np.arange(5000, dtype=np.int64) % 700) + 10_000
    agg.consume([Col(ctx.to_device(k2), L.INT64), Col(ctx.to_device(k2), L.INT64)], 5000)
    assert agg.state_bytes() == b1 // 300 * 1000
    agg.finalize()
    assert agg.num_groups() == 1000 and agg.state_bytes() == b1 // 300 * 1000
    agg.reset()
    assert agg.state_bytes() == 0
    agg.close()
