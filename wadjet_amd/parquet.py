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
            raise L.UnsupportedError(f"decimal type {t} not on the GPU path")
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
        raise L.UnsupportedError(f"column type {t} not on the GPU path")
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
