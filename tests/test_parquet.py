"""Parquet ingestion (host side): the seeded generator of the 1M-row file the
plain CPU method is quoted on, the write/read round trip, the type mapping to
the library's column layouts, and the oracle on that file against an
independent implementation (pyarrow's own group-by)."""
import numpy as np
import pytest

import oracle as O
from wadjet_amd import _lib as L
from wadjet_amd import parquet as Q

pa = pytest.importorskip("pyarrow")
pq = pytest.importorskip("pyarrow.parquet")

# SHA-256 of the generated (k, v) columns, little-endian: pins the generator
KV1K_DIGEST = "18c4e25125938c227192433c92a64a3de69fe8cb9fa84418dc9bae8aeb40fc3e"


def test_kv1k_generator_is_pinned():
    k, v = Q.gen_kv1k_host()
    assert Q.columns_digest([k, v]) == KV1K_DIGEST
    assert len(k) == 1_000_000 and k.dtype == np.int64 and v.dtype == np.int64
    assert len(np.unique(k)) == 1000
    assert v.min() >= 0 and v.max() < 1_000_000


def test_round_trip_in_batches(tmp_path):
    path = str(tmp_path / "kv1k.parquet")
    Q.write_kv1k(path, n=300_000, row_group_rows=65_536)
    k, v = Q.gen_kv1k_host(300_000)
    got_k, got_v, rows = [], [], 0
    for n, batch in Q.iter_parquet(path, ["k", "v"], batch_rows=50_000):
        (kv, kt, ks, kval), (vv, vt, vs, vval) = batch
        assert kt == L.INT64 and vt == L.INT64 and kval is None and vval is None
        got_k.append(kv)
        got_v.append(vv)
        rows += n
    assert rows == 300_000
    np.testing.assert_array_equal(np.concatenate(got_k), k)
    np.testing.assert_array_equal(np.concatenate(got_v), v)


def test_oracle_on_kv1k_matches_pyarrow_group_by(tmp_path):
    """SELECT k, SUM(v), COUNT(*), MIN(v), MAX(v) GROUP BY k: the oracle over
    the decoded file equals pyarrow's hash aggregation of the same file."""
    path = str(tmp_path / "kv1k.parquet")
    Q.write_kv1k(path)
    cols = []
    for n, batch in Q.iter_parquet(path, ["k", "v"], batch_rows=1 << 20):
        cols.append(batch)
    k = np.concatenate([b[0][0] for b in cols])
    v = np.concatenate([b[1][0] for b in cols])
    r = O.hash_agg([(k, O.INT64), (v, O.INT64)], [0], [(O.SUM, 1), (O.COUNT, -1), (O.MIN, 1), (O.MAX, 1)])
    o = np.argsort(r["keys"][:, 0])
    acc = r["acc"][o]
    t = pq.read_table(path).group_by("k").aggregate([("v", "sum"), ("v", "count"), ("v", "min"), ("v", "max")])
    t = t.sort_by("k")
    assert r["n"] == t.num_rows == 1000
    np.testing.assert_array_equal(r["keys"][o, 0], t["k"].to_numpy())
    np.testing.assert_array_equal(acc[:, 0]["sum_i64"], t["v_sum"].to_numpy())
    np.testing.assert_array_equal(acc[:, 1]["count"], t["v_count"].to_numpy())
    np.testing.assert_array_equal(acc[:, 2]["min_i64"], t["v_min"].to_numpy())
    np.testing.assert_array_equal(acc[:, 3]["max_i64"], t["v_max"].to_numpy())


def test_type_mapping_decimals_and_nulls(tmp_path):
    from decimal import Decimal
    path = str(tmp_path / "types.parquet")
    d15 = [Decimal("123.45"), None, Decimal("-0.01"), Decimal("9999999999999.99")]
    d30 = [Decimal("1" + "0" * 25 + ".50"), Decimal("-2.00"), None, Decimal("0.00")]
    tbl = pa.table({
        "i32": pa.array([1, None, -3, 4], pa.int32()),
        "u16": pa.array([1, 2, 65535, 0], pa.uint16()),
        "f": pa.array([0.5, None, -1.25, float("inf")], pa.float64()),
        "d15": pa.array(d15, pa.decimal128(15, 2)),
        "d30": pa.array(d30, pa.decimal128(30, 2)),
    })
    pq.write_table(tbl, path)
    (n, batch), = list(Q.iter_parquet(path, ["i32", "u16", "f", "d15", "d30"]))
    assert n == 4
    (i32, ti, _, vi), (u16, tu, _, vu), (f, tf, _, vf), (d15v, td, sd, vd), (d30v, t128, s128, v128) = batch
    assert (ti, tu, tf, td, t128) == (L.INT32, L.UINT16, L.FLOAT64, L.DECIMAL, L.DECIMAL128)
    assert (sd, s128) == (2, 2)
    valid_bits = lambda w: [(int(w[0]) >> i) & 1 for i in range(4)]
    assert valid_bits(vi) == [1, 0, 1, 1] and vu is None and valid_bits(vf) == [1, 0, 1, 1]
    np.testing.assert_array_equal(i32, [1, 0, -3, 4])
    assert d15v.tolist() == [12345, 0, -1, 999999999999999] and valid_bits(vd) == [1, 0, 1, 1]
    # DECIMAL128 rows are {Hi, Lo} pairs
    want = [10**27 + 50, -200, 0, 0]
    got = [(int(np.int64(d30v[2 * i].view(np.int64))) << 64) | int(d30v[2 * i + 1]) for i in range(4)]
    assert got == want and valid_bits(v128) == [1, 1, 0, 1]
