"""Host half of the device Parquet decode: the library's page walk,
decompressors and hybrid run scanner (wg_pq_stage with no context stages into
host memory; wg_pq_decompress; wg_pq_scan_runs), checked against
pyarrow-written files, pyarrow's codecs and hand-built streams. The staged
pages are expanded here on the CPU by the device's rules (runs, bit offsets,
dictionary lookups) and must give back pyarrow's column."""
import ctypes as C

import numpy as np
import pyarrow as pa
import pyarrow.parquet as pq
import pytest

from wadjet_amd import _lib as L
from wadjet_amd import parquet as PQ


def scan(buf: bytes, bw: int, count: int):
    """Run table (uint32 x 4 per run), runs, bytes used, values equal to 1."""
    arr = np.frombuffer(buf, np.uint8)
    ptr = arr.ctypes.data_as(C.c_void_p) if len(arr) else None
    nr, used, ones = C.c_uint64(), C.c_uint64(), C.c_uint64()
    L.call("wg_pq_scan_runs", ptr, len(arr), bw, count, None, 0, C.byref(nr), C.byref(used), C.byref(ones))
    runs = np.zeros(max(nr.value, 1) * 4, np.uint32)
    L.call("wg_pq_scan_runs", ptr, len(arr), bw, count, runs.ctypes.data_as(C.c_void_p), nr.value,
           C.byref(nr), C.byref(used), C.byref(ones))
    return runs, nr.value, used.value, ones.value


def expand_runs(runs: np.ndarray, n_runs: int, data: bytes, bw: int, count: int) -> np.ndarray:
    """Values of a hybrid stream from its run table (the device's rule, on
    the host): RLE value, or bw bits LSB-first from the run's bit offset."""
    out = np.zeros(count, np.int64)
    bits = np.unpackbits(np.frombuffer(data, np.uint8), bitorder="little")
    for r in range(n_runs):
        start, cnt, payload, packed = (int(x) for x in runs[4 * r:4 * r + 4])
        if packed:
            pos = payload + np.arange(cnt, dtype=np.int64)[:, None] * bw + np.arange(bw)[None, :]
            out[start:start + cnt] = (bits[pos].astype(np.int64) << np.arange(bw)).sum(axis=1)
        else:
            out[start:start + cnt] = payload
    return out


def test_hand_built_streams():
    """RLE run of 5 x 3 (header 5<<1, value byte), then a bit-packed group of
    8 values at width 3 (header 1<<1|1, 3 bytes): the Encodings.md example
    values 0..7 packed LSB-first are 0x88 0xC6 0xFA."""
    s = bytes([10, 3, 3, 0x88, 0xC6, 0xFA])
    runs, nr, used, _ = scan(s, 3, 13)
    assert expand_runs(runs, nr, s + b"\0" * 8, 3, 13).tolist() == [3] * 5 + list(range(8))
    assert nr == 2 and used == len(s)
    # width 1: the count of ones (present rows of a definition-level stream)
    lv = bytes([3, 0b10110101])  # one bit-packed group: 1,0,1,0,1,1,0,1
    assert scan(lv, 1, 8)[3] == 5
    assert scan(lv, 1, 6)[3] == 4  # only the first `count` values count
    with pytest.raises(L.WgError):
        scan(bytes([10]), 3, 5)  # RLE value missing


def _mem(addr: int, n: int) -> bytes:
    return C.string_at(addr, n) if n else b""


def reconstruct(st) -> tuple:
    """(values, valid) of a host-staged column, page by page."""
    dt = {PQ.PT_INT32: np.int32, PQ.PT_INT64: np.int64, PQ.PT_DOUBLE: np.float64}[st.phys]
    vals, valid = [], []
    for i in range(st.n_pages):
        p = st.pages[i]
        n = p.num_rows
        if p.def_.n_runs:
            runs = np.frombuffer(_mem(p.def_.runs, 16 * p.def_.n_runs), np.uint32)
            need = int(max(runs[2::4][runs[3::4] == 1] + runs[1::4][runs[3::4] == 1], default=0)) // 8 + 16
            present = expand_runs(runs, p.def_.n_runs, _mem(p.def_.data, need), 1, n) == 1
        else:
            present = np.ones(n, bool)
        k = int(present.sum())
        if p.encoding == PQ.ENC_PLAIN:
            v = np.frombuffer(_mem(p.plain, k * st.width), dt)
        else:
            bw = p.idx.bit_width
            runs = np.frombuffer(_mem(p.idx.runs, 16 * p.idx.n_runs), np.uint32) if p.idx.n_runs else np.zeros(0, np.uint32)
            packed = runs[3::4] == 1
            need = int(max(runs[2::4][packed] + runs[1::4][packed] * bw, default=0)) // 8 + 16
            idx = expand_runs(runs, p.idx.n_runs, _mem(p.idx.data, need), bw, k)
            d = np.frombuffer(_mem(p.dict, p.dict_n * st.width), dt)
            assert (idx < p.dict_n).all()
            v = d[idx]
        full = np.zeros(n, dt)
        full[present] = v
        vals.append(full)
        valid.append(present)
    return np.concatenate(vals) if vals else np.zeros(0, dt), np.concatenate(valid) if valid else np.zeros(0, bool)


def _check(path, column, row_groups=None, threads=0):
    st = PQ.StagedColumn(None, path, column, row_groups=row_groups, threads=threads)
    try:
        got, ok = reconstruct(st)
        rows = st.rows
    finally:
        st.free()
    tab = pq.ParquetFile(path)
    if row_groups is None:
        arr = tab.read(columns=[column]).column(0).combine_chunks()
    else:
        arr = pa.concat_arrays([tab.read_row_group(g, columns=[column]).column(0).combine_chunks() for g in row_groups])
    if pa.types.is_decimal(arr.type):
        raw = np.frombuffer(arr.buffers()[1], np.int64)[2 * arr.offset: 2 * (arr.offset + len(arr))][0::2]
        want = raw
    else:
        want = arr.fill_null(0).to_numpy(zero_copy_only=False)
    want_ok = ~arr.is_null().to_numpy(zero_copy_only=False)
    assert rows == len(arr)
    np.testing.assert_array_equal(ok, want_ok)
    np.testing.assert_array_equal(got[want_ok], np.asarray(want)[want_ok].astype(got.dtype))
    assert not np.any(got[~want_ok])
    return rows


def _table(n, seed=1):
    rng = np.random.default_rng(seed)
    return pa.table({
        "k": pa.array(rng.integers(0, 1000, n), pa.int64(), mask=rng.random(n) < 0.1),
        "g": pa.array(rng.integers(0, 16, n).astype(np.int32), pa.int32()),
        "v": pa.array(rng.normal(size=n), pa.float64(), mask=rng.random(n) < 0.3),
        "u": pa.array(rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32), pa.uint32()),
    })


@pytest.mark.parametrize("compression", ["none", "snappy", "gzip", "zstd", "lz4", "lz4_raw"])
@pytest.mark.parametrize("version", ["1.0", "2.0"])
def test_staged_pages_give_back_the_column(tmp_path, compression, version):
    path = str(tmp_path / "t.parquet")
    pq.write_table(_table(120_000), path, compression=compression, data_page_version=version,
                   data_page_size=8192, row_group_size=50_000)
    for c in ("k", "g", "v", "u"):
        assert _check(path, c) == 120_000


def test_dictionary_fallback_and_row_group_subsets(tmp_path):
    """A dictionary that outgrows its page falls back to PLAIN inside the
    chunk; a subset of row groups stages alone; one thread or many."""
    path = str(tmp_path / "fb.parquet")
    n = 200_000
    rng = np.random.default_rng(3)
    t = pa.table({"k": pa.array(rng.integers(0, 1 << 40, n), pa.int64(), mask=rng.random(n) < 0.02)})
    pq.write_table(t, path, dictionary_pagesize_limit=16_384, data_page_size=16_384, row_group_size=60_000)
    for threads in (1, 3, 0):
        _check(path, "k", threads=threads)
    _check(path, "k", row_groups=[1, 3])


def test_decimals_all_null_pages_and_required_columns(tmp_path):
    from decimal import Decimal
    path = str(tmp_path / "d.parquet")
    n = 30_000
    dec = [None if (i // 5000) % 2 else Decimal(i * 7 - 100_000).scaleb(-2) for i in range(n)]
    t = pa.table({"d": pa.array(dec, pa.decimal128(18, 2)),
                  "z": pa.array([None] * n, pa.int64()),
                  "r": pa.array(np.arange(n, dtype=np.int64) % 97)})
    schema = pa.schema([pa.field("d", pa.decimal128(18, 2)), pa.field("z", pa.int64()),
                        pa.field("r", pa.int64(), nullable=False)])
    pq.write_table(t.cast(schema), path, data_page_size=4096, store_decimal_as_integer=True)
    for c in ("d", "z", "r"):
        _check(path, c)


def test_errors_name_the_problem(tmp_path):
    path = str(tmp_path / "e.parquet")
    pq.write_table(_table(20_000), path, compression="brotli")
    with pytest.raises(L.UnsupportedError, match="codec 4"):
        PQ.StagedColumn(None, path, "k")
    path2 = str(tmp_path / "e2.parquet")
    pq.write_table(_table(20_000), path2)
    cs, chunks, n = PQ.column_chunks(path2, "k")
    chunks[0].num_rows += 1
    h = C.c_void_p()
    with pytest.raises(L.WgError, match="pages hold 20000 rows, the row group 20001"):
        L.call("wg_pq_stage", None, path2.encode(), chunks, n, PQ.PT_INT64, 1, 0, C.byref(h))
    chunks[0].num_rows -= 1
    chunks[0].offset += 3  # not at a page header
    with pytest.raises(L.WgError):
        L.call("wg_pq_stage", None, path2.encode(), chunks, n, PQ.PT_INT64, 1, 0, C.byref(h))
    with pytest.raises(L.WgError, match="cannot open"):
        L.call("wg_pq_stage", None, str(tmp_path / "missing.parquet").encode(), chunks, n, PQ.PT_INT64, 1, 0,
               C.byref(h))


def _decompress(codec: int, src: bytes, size: int) -> bytes:
    s = np.frombuffer(src, np.uint8)
    dst = np.zeros(max(size, 1), np.uint8)
    L.call("wg_pq_decompress", codec, s.ctypes.data_as(C.c_void_p) if len(s) else None, len(s),
           dst.ctypes.data_as(C.c_void_p), size)
    return dst[:size].tobytes()


@pytest.mark.parametrize("name,codec", [("snappy", 1), ("gzip", 2), ("zstd", 6), ("lz4", 5), ("lz4_raw", 7)])
def test_decompressors_against_pyarrow_codecs(name, codec):
    """Incompressible bytes (long literals), repeats at short distances
    (overlapping back references), a long zero run and an empty page."""
    rng = np.random.default_rng(11)
    cases = [rng.integers(0, 256, 300_000, dtype=np.uint8).tobytes(),
             (b"abcab" * 40_000) + bytes(rng.integers(0, 4, 70_000, dtype=np.uint8)),
             bytes(1 << 20), b"x"]
    if name != "gzip":
        cases.append(b"")
    cod = pa.Codec(name)
    for raw in cases:
        comp = cod.compress(raw, asbytes=True)
        assert _decompress(codec, comp, len(raw)) == raw
    # a size that disagrees with the stream, and a cut stream, are errors
    comp = cod.compress(cases[1], asbytes=True)
    with pytest.raises(L.WgError):
        _decompress(codec, comp, len(cases[1]) + 1)
    with pytest.raises(L.WgError):
        _decompress(codec, comp[:len(comp) // 2], len(cases[1]))
