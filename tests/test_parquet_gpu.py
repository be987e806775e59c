"""Parquet → device ingestion feeding the GPU aggregate: the 1M-row file of the
plain method (read_parquet, SUM/COUNT/MIN/MAX GROUP BY k) and a file with
NULLs and DECIMAL columns, each streamed as host batches through pinned
staging, against the oracle over the same decoded rows."""
import numpy as np
import pytest

import oracle as O
from helpers import compare, gpu_groups, oracle_groups
from wadjet_amd import _lib as L
from wadjet_amd import parquet as Q
from wadjet_amd.gpu import AggSpec, HashAggregate

pytestmark = pytest.mark.gpu
pa = pytest.importorskip("pyarrow")
pq = pytest.importorskip("pyarrow.parquet")


def decoded(path, names):
    """(values, type, validity words, scale) per column over the whole file."""
    t = pq.read_table(path, columns=names).combine_chunks()
    out = []
    for i in range(len(names)):
        v, ty, sc, valid = Q._column(t.column(i).chunk(0))
        out.append((v, ty, valid, sc))
    return out


def test_kv1k_read_parquet_aggregate(ctx, tmp_path):
    path = str(tmp_path / "kv1k.parquet")
    Q.write_kv1k(path)
    print("[parquet] k-over-1000 file, 1M rows, 256K-row batches", flush=True)
    aggs = [AggSpec(L.SUM, 1), AggSpec(L.COUNT, -1), AggSpec(L.MIN, 1), AggSpec(L.MAX, 1)]
    agg = HashAggregate(ctx, [L.INT64, L.INT64], [0], aggs)
    assert Q.consume_parquet(agg, path, ["k", "v"], batch_rows=1 << 18) == 1_000_000
    agg.finalize()
    k, v = Q.gen_kv1k_host()
    ref = oracle_groups(O.hash_agg([(k, O.INT64), (v, O.INT64)], [0], [(a.func, a.col) for a in aggs]))
    compare(gpu_groups(agg), ref, [(a.func, a.col) for a in aggs], [L.INT64, L.INT64])
    assert agg.num_groups() == 1000


def test_parquet_nulls_and_decimals(ctx, tmp_path):
    rng = np.random.default_rng(12)
    n = 200_000
    g = rng.integers(0, 50, n).astype(np.int32)
    g_null = rng.random(n) < 0.05
    price = rng.integers(-10**12, 10**12, n)
    price_null = rng.random(n) < 0.1
    big = [int(x) * 10**10 for x in rng.integers(-10**9, 10**9, n)]
    from decimal import Decimal
    tbl = pa.table({
        "g": pa.array(g, pa.int32(), mask=g_null),
        "p": pa.array([Decimal(int(x)).scaleb(-2) for x in price], pa.decimal128(15, 2), mask=price_null),
        "b": pa.array([Decimal(x).scaleb(-2) for x in big], pa.decimal128(30, 2)),
    })
    path = str(tmp_path / "nd.parquet")
    pq.write_table(tbl, path, row_group_size=50_000)
    types, scales = Q.schema_of(path, ["g", "p", "b"])
    assert types == [L.INT32, L.DECIMAL, L.DECIMAL128] and scales == [0, 2, 2]
    print(f"[parquet] nulls+decimals {n} rows", flush=True)
    aggs = [AggSpec(L.SUM, 1), AggSpec(L.COUNT, 1), AggSpec(L.SUM, 2), AggSpec(L.AVG, 1), AggSpec(L.COUNT, -1)]
    agg = HashAggregate(ctx, types, [0], aggs, scales=scales, nullable=[True, True, False])
    Q.consume_parquet(agg, path, ["g", "p", "b"], batch_rows=30_000)
    agg.finalize()
    ref = oracle_groups(O.hash_agg(decoded(path, ["g", "p", "b"]), [0], [(a.func, a.col) for a in aggs]))
    compare(gpu_groups(agg), ref, [(a.func, a.col) for a in aggs], types, scales)
