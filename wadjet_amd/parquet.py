"""Parquet → device ingestion for the aggregate: host decode, pinned upload.

The reference's `read_parquet()` table function (planner/physical/
table_func.go:135-185) decodes a file with its own reader
(storage/parquet/reader.go) into RecordBatches that feed HashAggregate. Here
the file is decoded on the host by pyarrow, batch by batch, and each batch is
handed to `wg_agg_consume` as host memory (WG_MEM_HOST): the library gathers
it into pinned staging buffers and ships large batches to HBM, overlapping the
copy with the previous batch's kernels.

Column types map as the reference's vectors do (vector.go:302-346): int32,
int64, uint32 (compact IPv4), uint16 (compact Port), float64, decimal128 with
precision <= 18 as scaled int64 (DECIMAL), wider decimals as Int128
{Hi, Lo} (DECIMAL128). Validity becomes the reference's Bitmap layout (bit
i % 64 of word i / 64, 1 = valid).

Also the seeded generator and writer of the 1M-row file the plain CPU method
is quoted on: k int64 uniform over 1000 distinct values, v int64 uniform in
[0, 1e6).
"""
from __future__ import annotations

import hashlib
import os
from typing import Iterator, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib as L

KV1K_ROWS = 1_000_000
KV1K_SEED = 1


def _mix64(x: np.ndarray) -> np.ndarray:
    x = x.astype(np.uint64)
    with np.errstate(over="ignore"):
        x ^= x >> np.uint64(33)
        x *= np.uint64(0xFF51AFD7ED558CCD)
        x ^= x >> np.uint64(33)
        x *= np.uint64(0xC4CEB9FE1A85EC53)
        x ^= x >> np.uint64(33)
    return x


# the 1000 distinct group keys: spread over the whole int64 range
KV1K_DISTINCT = _mix64(np.arange(1000, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15)).view(np.int64)


def gen_kv1k_host(n: int = KV1K_ROWS, seed: int = KV1K_SEED) -> Tuple[np.ndarray, np.ndarray]:
    """Seeded (k, v): k uniform over KV1K_DISTINCT, v uniform in [0, 1e6)."""
    i = np.arange(n, dtype=np.uint64)
    s = np.uint64(seed) * np.uint64(0x632BE59BD9B4E019)
    k = KV1K_DISTINCT[(_mix64(i + s) % np.uint64(1000)).astype(np.int64)]
    v = (_mix64(i + s + np.uint64(0x3C6EF372FE94F82B)) % np.uint64(1_000_000)).astype(np.int64)
    return k, v


def columns_digest(arrays: Sequence[np.ndarray]) -> str:
    """SHA-256 over the columns' little-endian bytes (pins the generator)."""
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).astype(a.dtype.newbyteorder("<")).tobytes())
    return h.hexdigest()


def write_kv1k(path: str, n: int = KV1K_ROWS, seed: int = KV1K_SEED, row_group_rows: int = 131_072) -> None:
    import pyarrow as pa
    import pyarrow.parquet as pq
    k, v = gen_kv1k_host(n, seed)
    pq.write_table(pa.table({"k": k, "v": v}), path, row_group_size=row_group_rows)


# ------------------------------------------------------------------ reading

def _validity_words(arr) -> Optional[np.ndarray]:
    if arr.null_count == 0:
        return None
    valid = ~arr.is_null().to_numpy(zero_copy_only=False)
    bits = np.packbits(valid, bitorder="little")
    pad = (-len(bits)) % 8
    if pad:
        bits = np.concatenate([bits, np.zeros(pad, np.uint8)])
    return bits.view(np.uint64).copy()


def _column(arr) -> Tuple[np.ndarray, int, int, Optional[np.ndarray]]:
    """(values, wg type, scale, validity words) of one Arrow array."""
    import pyarrow as pa
    t = arr.type
    valid = _validity_words(arr)
    if pa.types.is_decimal(t):
        if not pa.types.is_decimal128(t):
            raise L.UnsupportedError("parquet", L.WG_ERR_UNSUPPORTED, f"decimal type {t} not on the GPU path")
        raw = np.frombuffer(arr.buffers()[1], dtype=np.uint64)[2 * arr.offset: 2 * (arr.offset + len(arr))]
        lo, hi = raw[0::2], raw[1::2]
        if t.precision <= 18:
            return lo.view(np.int64).copy(), L.DECIMAL, t.scale, valid
        out = np.empty(2 * len(arr), np.uint64)
        out[0::2], out[1::2] = hi, lo  # {Hi, Lo} per row
        return out, L.DECIMAL128, t.scale, valid
    m = {pa.int32(): (np.int32, L.INT32), pa.int64(): (np.int64, L.INT64), pa.uint32(): (np.uint32, L.UINT32),
         pa.uint16(): (np.uint16, L.UINT16), pa.float64(): (np.float64, L.FLOAT64)}
    if t not in m:
        raise L.UnsupportedError("parquet", L.WG_ERR_UNSUPPORTED, f"column type {t} not on the GPU path")
    dt, wt = m[t]
    vals = arr.to_numpy(zero_copy_only=False)
    if valid is not None:  # nulls decode as NaN / object; store zeros under them
        vals = np.where(arr.is_null().to_numpy(zero_copy_only=False), 0, arr.fill_null(0).to_numpy()).astype(dt)
    return np.ascontiguousarray(vals, dtype=dt), wt, 0, valid


def iter_parquet(path: str, columns: Sequence[str],
                 batch_rows: int = 1 << 20) -> Iterator[Tuple[int, List[tuple]]]:
    """Decoded batches: (rows, per column (values, wg type, scale, validity words))."""
    import pyarrow.parquet as pq
    f = pq.ParquetFile(path)
    for b in f.iter_batches(batch_size=batch_rows, columns=list(columns)):
        yield b.num_rows, [_column(b.column(i)) for i in range(b.num_columns)]


def schema_of(path: str, columns: Sequence[str]) -> Tuple[List[int], List[int]]:
    """wg types and scales of the named columns (from the first batch)."""
    for _, batch in iter_parquet(path, columns, batch_rows=1):
        return [c[1] for c in batch], [c[2] for c in batch]
    raise ValueError(f"{path} has no rows")


def consume_parquet(agg, path: str, columns: Sequence[str], pred=None, batch_rows: int = 1 << 20) -> int:
    """Stream a Parquet file's columns into a HashAggregate (host batches).
    Returns the rows consumed."""
    from .gpu import Col
    rows = 0
    for n, batch in iter_parquet(path, columns, batch_rows):
        cols = [Col(v, t, valid, s) for (v, t, s, valid) in batch]
        agg.consume(cols, n, pred=pred, device=False)
        rows += n
    return rows


# ------------------------------------------------------ device page decode
#
# The reference's own Parquet reader (internal/storage/parquet: footer.go,
# thrift.go, page_reader.go, rle.go) decodes every page on the CPU. Here the
# footer comes from pyarrow's metadata (chunk offsets, sizes, codecs, types);
# the library stages each column's chunks on a pool of host threads
# (wg_pq_stage: page headers, decompression, run headers of each RLE /
# bit-packing hybrid stream, pinned uploads) and the device decodes the values
# (wg_parquet_decode_pages): definition levels, PLAIN values, dictionary
# indices resolved against the dictionary page.

PAGE_DATA, PAGE_INDEX, PAGE_DICTIONARY, PAGE_DATA_V2 = 0, 1, 2, 3
ENC_PLAIN, ENC_PLAIN_DICTIONARY, ENC_RLE_DICTIONARY = 0, 2, 8
PT_INT32, PT_INT64, PT_DOUBLE = 1, 2, 5
# CompressionCodec (parquet.thrift) by pyarrow's metadata name
CODECS = {"UNCOMPRESSED": 0, "SNAPPY": 1, "GZIP": 2, "LZO": 3, "BROTLI": 4, "LZ4": 5, "ZSTD": 6, "LZ4_RAW": 7}


def _wg_type_of(col_schema) -> Tuple[int, int, int]:
    """(physical type id, wg type, scale) of a flat column."""
    phys = {"INT32": PT_INT32, "INT64": PT_INT64, "DOUBLE": PT_DOUBLE}.get(col_schema.physical_type)
    if phys is None:
        raise L.UnsupportedError("parquet", L.WG_ERR_UNSUPPORTED,
                                 f"physical type {col_schema.physical_type} not on the GPU decode path")
    lt = col_schema.logical_type
    ltype = getattr(lt, "type", "NONE")
    if ltype == "DECIMAL":
        if phys != PT_INT64:
            raise L.UnsupportedError("parquet", L.WG_ERR_UNSUPPORTED,
                                     "decimals stored as INT32 are not on the GPU decode path")
        return phys, L.DECIMAL, col_schema.scale
    if phys == PT_INT32:
        lts = str(lt).replace(" ", "")
        if "isSigned=false" in lts and "bitWidth=32" in lts:
            return phys, L.UINT32, 0
        return phys, L.INT32, 0
    return phys, (L.INT64 if phys == PT_INT64 else L.FLOAT64), 0


def column_chunks(path: str, column: str, row_groups: Optional[Sequence[int]] = None):
    """(column schema, [wg_pq_chunk]) of one flat column, from the footer."""
    import pyarrow.parquet as pq
    f = pq.ParquetFile(path)
    md = f.metadata
    ci = f.schema_arrow.get_field_index(column)
    if ci < 0:
        raise KeyError(column)
    cs = f.schema.column(ci)
    if cs.max_repetition_level != 0:
        raise L.UnsupportedError("parquet", L.WG_ERR_UNSUPPORTED, "repeated columns are not on the GPU decode path")
    rgs = list(range(md.num_row_groups)) if row_groups is None else list(row_groups)
    chunks = (L.PqChunk * max(len(rgs), 1))()
    for ch, rg in zip(chunks, rgs):
        cc = md.row_group(rg).column(ci)
        ch.offset = (cc.dictionary_page_offset if cc.has_dictionary_page and cc.dictionary_page_offset
                     else cc.data_page_offset)
        ch.bytes = cc.total_compressed_size
        ch.num_rows = md.row_group(rg).num_rows
        ch.codec = CODECS.get(cc.compression, -1)
    return cs, chunks, len(rgs)


class StagedColumn:
    """One column's chunks (all row groups, or those named) read, parsed,
    decompressed and scanned by the library on host threads and staged in
    HBM (wg_pq_stage): the input of the device decode, which `decode()` runs
    in one wg_parquet_decode_pages call. ctx=None stages into host memory
    (the host half alone; `decode()` then refuses). `page_bytes` = the stream
    bytes the decode reads."""

    def __init__(self, ctx, path: str, column: str, row_groups: Optional[Sequence[int]] = None, threads: int = 0):
        import ctypes as C
        cs, chunks, n_chunks = column_chunks(path, column, row_groups)
        self.ctx = ctx
        self.max_def = cs.max_definition_level
        self.phys, self.type, self.scale = _wg_type_of(cs)
        self.width = 4 if self.phys == PT_INT32 else 8
        h = C.c_void_p()
        L.call("wg_pq_stage", ctx.h if ctx is not None else None, os.fsencode(path), chunks, n_chunks, self.phys,
               self.max_def, threads, C.byref(h))
        self.h = h
        rows, n_pages, stream_bytes, staged = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_uint64()
        L.call("wg_pq_staged_info", h, C.byref(rows), C.byref(n_pages), C.byref(stream_bytes), C.byref(staged))
        self.rows, self.n_pages = rows.value, n_pages.value
        self.page_bytes, self.staged_bytes = stream_bytes.value, staged.value
        self.pages = (L.PqPage * max(self.n_pages, 1))()
        L.call("wg_pq_staged_pages", h, self.pages, self.n_pages)

    def decode(self, out=None, out_valid=None, row0: int = 0):
        """Decode every page on the device. With `out` given (a device buffer
        of the column's width, and `out_valid` zeroed bitmap words for a
        nullable column) the rows land at row0; else new buffers. -> Col."""
        from .gpu import Col
        if self.ctx is None:
            raise L.WgError("wg_parquet_decode_pages", L.WG_ERR_INVALID, "staged in host memory")
        if out is None:
            out = self.ctx.alloc(max(self.rows, 1) * self.width)
            row0 = 0
            if self.max_def > 0:
                out_valid = self.ctx.alloc(max((self.rows + 63) // 64, 1) * 8)
                self.ctx.memset(out_valid, 0)
        L.call("wg_parquet_decode_pages", self.ctx.h, self.pages, self.n_pages, out.ptr,
               out_valid.ptr if (out_valid is not None and self.max_def > 0) else None, row0)
        return Col(out, self.type, out_valid if self.max_def > 0 else None, self.scale)

    def free(self):
        if getattr(self, "h", None):
            L.call("wg_pq_staged_free", self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def decode_column_chunk(ctx, path: str, row_group: int, column: str, out=None, out_valid=None, row0: int = 0):
    """Decode one column chunk on the device. Returns (rows, Col); with `out`
    given (a device buffer of the column's width), pages land at row0."""
    st = StagedColumn(ctx, path, column, [row_group])
    try:
        return st.rows, st.decode(out, out_valid, row0)
    finally:
        st.free()


def read_parquet_gpu(ctx, path: str, columns: Sequence[str]):
    """Every row group of the named columns decoded on the device into one
    device column each: (rows, [Col])."""
    n, cols = 0, []
    for name in columns:
        st = StagedColumn(ctx, path, name)
        try:
            cols.append(st.decode())
            n = st.rows
        finally:
            st.free()
    return n, cols
