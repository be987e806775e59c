"""Parquet pages decoded on the device (wg_parquet_decode_pages /
wg_parquet_decode_page, driven by wadjet_amd.parquet) against pyarrow's decode of the same
files: values and NULLs bit-exact for every page layout pyarrow writes —
PLAIN and dictionary encodings (and the fallback from one to the other within
a chunk), data page v1 and v2, each codec, nullable and required columns,
many pages and row groups."""
import numpy as np
import pytest

from wadjet_amd import _lib as L
from wadjet_amd import parquet as PQ

pytestmark = pytest.mark.gpu
pa = pytest.importorskip("pyarrow")
pq = pytest.importorskip("pyarrow.parquet")

NP = {L.INT32: np.int32, L.UINT32: np.uint32, L.INT64: np.int64, L.DECIMAL: np.int64, L.FLOAT64: np.float64}


def table(n, seed, null_p):
    rng = np.random.default_rng(seed)
    mask = lambda: (rng.random(n) < null_p) if null_p else None
    f = rng.normal(size=n)
    f[:3] = [0.0, -0.0, np.inf]
    cols = {
        "k64": pa.array(rng.integers(0, 1000, n), pa.int64(), mask=mask()),
        "u64": pa.array(rng.integers(-2**62, 2**62, n), pa.int64(), mask=mask()),
        "k32": pa.array(rng.integers(-50, 50, n).astype(np.int32), pa.int32(), mask=mask()),
        "u32": pa.array(rng.integers(0, 2**32, n, dtype=np.int64).astype(np.uint32), pa.uint32(), mask=mask()),
        "f": pa.array(f, pa.float64(), mask=mask()),
    }
    return pa.table(cols)


def check(ctx, path, names):
    n, cols = PQ.read_parquet_gpu(ctx, path, names)
    t = pq.read_table(path, columns=names).combine_chunks()
    assert n == t.num_rows
    for name, col in zip(names, cols):
        arr = t.column(name).chunk(0)
        want_valid = ~arr.is_null().to_numpy(zero_copy_only=False)
        got = col.data.to_host(NP[col.type], n)
        if col.validity is not None:
            words = col.validity.to_host(np.uint64, (n + 63) // 64)
            got_valid = ((words[np.arange(n) >> 6] >> (np.arange(n) & 63).astype(np.uint64)) & np.uint64(1)) == 1
        else:
            got_valid = np.ones(n, bool)
        np.testing.assert_array_equal(got_valid, want_valid, err_msg=f"{name}: validity")
        if col.type == L.DECIMAL:
            want = np.frombuffer(arr.buffers()[1], np.int64)[2 * arr.offset:2 * (arr.offset + n):2]
        else:
            want = arr.fill_null(0).to_numpy(zero_copy_only=False).astype(NP[col.type])
        if col.type == L.FLOAT64:
            np.testing.assert_array_equal(got.view(np.uint64)[want_valid], want.view(np.uint64)[want_valid],
                                          err_msg=name)
        else:
            np.testing.assert_array_equal(got[want_valid], want[want_valid], err_msg=name)
        assert not np.any(got[~want_valid]), f"{name}: NULL rows must decode to 0"


@pytest.mark.parametrize("compression", ["none", "snappy", "zstd", "gzip"])
@pytest.mark.parametrize("version", ["1.0", "2.0"])
def test_codecs_and_page_versions(ctx, tmp_path, compression, version):
    print(f"[pq] {compression} v{version}", flush=True)
    path = str(tmp_path / "a.parquet")
    t = table(120_000, 1, 0.1)
    pq.write_table(t, path, compression=compression, data_page_version=version, data_page_size=32_768,
                   row_group_size=50_000)
    check(ctx, path, ["k64", "u64", "k32", "u32", "f"])


@pytest.mark.parametrize("dictionary", [True, False])
def test_required_columns_plain_and_dictionary(ctx, tmp_path, dictionary):
    print(f"[pq] required dict={dictionary}", flush=True)
    path = str(tmp_path / "b.parquet")
    t = table(300_001, 2, 0.0)
    schema = pa.schema([pa.field(f.name, f.type, nullable=False) for f in t.schema])
    pq.write_table(t.cast(schema), path, use_dictionary=dictionary, row_group_size=100_000)
    check(ctx, path, ["k64", "u64", "k32", "u32", "f"])


def test_dictionary_fallback_to_plain_within_a_chunk(ctx, tmp_path):
    """A small dictionary page limit makes the writer switch from dictionary
    to PLAIN pages inside one column chunk (TestDictFallbackMixedPages,
    dict_fallback_test.go:44)."""
    print("[pq] dictionary fallback", flush=True)
    path = str(tmp_path / "c.parquet")
    t = table(200_000, 3, 0.05)
    pq.write_table(t, path, dictionary_pagesize_limit=4096, data_page_size=8192)
    md = pq.ParquetFile(path).metadata
    assert "PLAIN" in md.row_group(0).column(1).encodings
    check(ctx, path, ["k64", "u64", "k32", "f"])


def test_decimal_column_and_all_null_pages(ctx, tmp_path):
    """DECIMAL(15,2) stored as INT64 (the scaled integers, the device layout
    of DECIMAL) and a column whose pages hold only NULLs."""
    print("[pq] decimal + all-null", flush=True)
    rng = np.random.default_rng(4)
    n = 70_000
    vals = [None if i % 5 == 0 else int(x) for i, x in enumerate(rng.integers(-10**12, 10**12, n))]
    import decimal
    dec = pa.array([None if v is None else decimal.Decimal(v).scaleb(-2) for v in vals], pa.decimal128(15, 2))
    nulls = pa.array([None] * n, pa.int64())
    path = str(tmp_path / "d.parquet")
    pq.write_table(pa.table({"dec": dec, "nulls": nulls}), path, data_page_size=4096, store_decimal_as_integer=True)
    assert pq.ParquetFile(path).schema.column(0).physical_type == "INT64"
    check(ctx, path, ["dec", "nulls"])


def test_gpu_decoded_columns_feed_the_aggregate(ctx, tmp_path):
    """The kv1k file (k int64 over 1000 values, v int64), decoded on the
    device, aggregated on the device, equals the host-decoded path."""
    from wadjet_amd.gpu import AggSpec, HashAggregate
    path = str(tmp_path / "kv1k.parquet")
    PQ.write_kv1k(path, n=300_000)
    n, cols = PQ.read_parquet_gpu(ctx, path, ["k", "v"])
    aggs = [AggSpec(L.SUM, 1), AggSpec(L.COUNT, -1), AggSpec(L.MIN, 1), AggSpec(L.MAX, 1)]
    # the writer's columns are OPTIONAL, so the decoded columns carry validity
    a = HashAggregate(ctx, [L.INT64, L.INT64], [0], aggs,
                      nullable=[c.validity is not None for c in cols])
    a.consume(cols, n)
    a.finalize()
    b = HashAggregate(ctx, [L.INT64, L.INT64], [0], aggs)
    PQ.consume_parquet(b, path, ["k", "v"])
    b.finalize()
    ra, rb = a.fetch(), b.fetch()
    oa, ob = np.argsort(ra["keys"][0]), np.argsort(rb["keys"][0])
    np.testing.assert_array_equal(ra["keys"][0][oa], rb["keys"][0][ob])
    for x, y in zip(ra["values"], rb["values"]):
        np.testing.assert_array_equal(x[oa], y[ob])


def test_chunk_lands_at_an_unaligned_row_offset(ctx, tmp_path):
    """A chunk decoded into a larger column at row 37: its values and
    validity bits land there, and the rows around it are left alone."""
    path = str(tmp_path / "off.parquet")
    pq.write_table(table(30_000, 11, 0.3), path, row_group_size=10_000, data_page_size=8192)
    t = pq.read_table(path, columns=["u64"]).combine_chunks()
    arr = t.column("u64").chunk(0).slice(10_000, 10_000)
    total, row0 = 10_100, 37
    out = ctx.to_device(np.full(total, -1, np.int64))
    words = (total + 63) // 64
    sentinel = np.zeros(words, np.uint64)
    sentinel[0] = np.uint64((1 << 37) - 1)  # rows before row0 already valid
    valid = ctx.to_device(sentinel)
    n, col = PQ.decode_column_chunk(ctx, path, 1, "u64", out=out, out_valid=valid, row0=row0)
    assert n == 10_000
    got = out.to_host(np.int64, total)
    bits = valid.to_host(np.uint64, words)
    v = ((bits[np.arange(total) >> 6] >> (np.arange(total) & 63).astype(np.uint64)) & np.uint64(1)) == 1
    want_valid = ~arr.is_null().to_numpy(zero_copy_only=False)
    np.testing.assert_array_equal(v[:row0], True)
    np.testing.assert_array_equal(v[row0:row0 + n], want_valid)
    np.testing.assert_array_equal(v[row0 + n:], False)
    np.testing.assert_array_equal(got[:row0], -1)
    np.testing.assert_array_equal(got[row0 + n:], -1)
    want = arr.fill_null(0).to_numpy(zero_copy_only=False)
    np.testing.assert_array_equal(got[row0:row0 + n], want)


def test_out_of_range_dictionary_index_is_an_error(ctx, tmp_path):
    """A page whose indices name entries past its dictionary is refused
    (ValidateDictIndices): the staged page's dictionary is cut to one entry."""
    path = str(tmp_path / "dict.parquet")
    pq.write_table(pa.table({"k": pa.array(np.arange(5000) % 7, pa.int64())}), path)
    st = PQ.StagedColumn(ctx, path, "k")
    try:
        assert st.pages[0].encoding in (2, 8) and st.pages[0].dict_n == 7
        st.decode()  # intact: decodes
        st.pages[0].dict_n = 1
        with pytest.raises(L.WgError, match="out of range"):
            st.decode()
    finally:
        st.free()


def test_all_row_groups_decode_in_one_call(ctx, tmp_path):
    """Many chunks, each with its own dictionary, staged together and decoded
    by one call; chunks of 4097 rows straddle the 4096-row tiles."""
    path = str(tmp_path / "many.parquet")
    n = 450_000
    rng = np.random.default_rng(5)
    k = rng.integers(0, 3000, n)
    t = pa.table({"k": pa.array(k, pa.int64(), mask=rng.random(n) < 0.05)})
    pq.write_table(t, path, row_group_size=4097, data_page_size=1 << 20)
    st = PQ.StagedColumn(ctx, path, "k")
    try:
        assert st.n_pages >= 100
        col = st.decode()
    finally:
        st.free()
    got = col.data.to_host(np.int64, n)
    arr = pq.read_table(path).combine_chunks().column("k").chunk(0)
    nul = arr.is_null().to_numpy(zero_copy_only=False)
    np.testing.assert_array_equal(got[~nul], arr.fill_null(0).to_numpy()[~nul])
    assert not np.any(got[nul])
