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


# ------------------------------------------------------ device page decode
#
# The reference's own Parquet reader (internal/storage/parquet: footer.go,
# thrift.go, page_reader.go, rle.go) decodes every page on the CPU. Here the
# host reads a column chunk's bytes, parses the Thrift compact page headers,
# decompresses with pyarrow's codecs and scans the run headers of each
# RLE / bit-packing hybrid stream (wg_pq_scan_runs); the device decodes the
# values (wg_parquet_decode_page): definition levels, PLAIN values,
# dictionary indices resolved against the dictionary page.

_CT_STOP, _CT_TRUE, _CT_FALSE, _CT_BYTE, _CT_I16, _CT_I32, _CT_I64, _CT_DOUBLE, _CT_BINARY, _CT_LIST, _CT_SET, \
    _CT_MAP, _CT_STRUCT = range(13)

PAGE_DATA, PAGE_INDEX, PAGE_DICTIONARY, PAGE_DATA_V2 = 0, 1, 2, 3
ENC_PLAIN, ENC_PLAIN_DICTIONARY, ENC_RLE_DICTIONARY = 0, 2, 8
PT_INT32, PT_INT64, PT_DOUBLE = 1, 2, 5


class _Compact:
    """Thrift compact protocol reader for page headers (thrift.go restates the
    same protocol): field headers with deltas, zigzag varints, nested structs;
    unknown fields are skipped."""

    def __init__(self, buf: bytes, pos: int = 0):
        self.b = buf
        self.p = pos

    def varint(self) -> int:
        v, shift = 0, 0
        while True:
            c = self.b[self.p]
            self.p += 1
            v |= (c & 0x7F) << shift
            shift += 7
            if not c & 0x80:
                return v

    def zigzag(self) -> int:
        v = self.varint()
        return (v >> 1) ^ -(v & 1)

    def value(self, t: int):
        if t in (_CT_TRUE, _CT_FALSE):
            return t == _CT_TRUE
        if t == _CT_BYTE:
            self.p += 1
            return self.b[self.p - 1]
        if t in (_CT_I16, _CT_I32, _CT_I64):
            return self.zigzag()
        if t == _CT_DOUBLE:
            self.p += 8
            return None
        if t == _CT_BINARY:
            n = self.varint()
            self.p += n
            return self.b[self.p - n:self.p]
        if t in (_CT_LIST, _CT_SET):
            h = self.b[self.p]
            self.p += 1
            size, et = h >> 4, h & 0x0F
            if size == 15:
                size = self.varint()
            return [self._bool_elem() if et in (_CT_TRUE, _CT_FALSE) else self.value(et) for _ in range(size)]
        if t == _CT_MAP:
            size = self.varint()
            if size:
                kv = self.b[self.p]
                self.p += 1
                for _ in range(size):
                    self.value(kv >> 4)
                    self.value(kv & 0x0F)
            return None
        if t == _CT_STRUCT:
            return self.struct()
        raise ValueError(f"thrift compact type {t}")

    def _bool_elem(self):
        self.p += 1
        return self.b[self.p - 1] == 1

    def struct(self) -> dict:
        out, fid = {}, 0
        while True:
            h = self.b[self.p]
            self.p += 1
            if h == _CT_STOP:
                return out
            delta, t = h >> 4, h & 0x0F
            fid = fid + delta if delta else self.zigzag()
            out[fid] = self.value(t)


def _read_page_header(buf: bytes, pos: int):
    r = _Compact(buf, pos)
    h = r.struct()
    return h, r.p


def _align(n: int, a: int = 8) -> int:
    return (n + a - 1) // a * a


class _Arena:
    """Host staging of one column chunk's streams, 8-byte aligned, each
    followed by 16 readable padding bytes (the device reads whole words)."""

    def __init__(self):
        self.parts, self.size = [], 0

    def add(self, data) -> int:
        off = self.size
        b = np.frombuffer(data, np.uint8) if not isinstance(data, np.ndarray) else data.view(np.uint8).ravel()
        self.parts.append((off, b))
        self.size = _align(off + len(b) + 16)
        return off

    def pack(self) -> np.ndarray:
        out = np.zeros(max(self.size, 8), np.uint8)
        for off, b in self.parts:
            out[off:off + len(b)] = b
        return out


def _scan_runs(buf, bit_width: int, count: int):
    """Run table (uint32 x 4 per run), bytes used, and values equal to 1."""
    import ctypes as C
    arr = np.frombuffer(buf, np.uint8)
    ptr = arr.ctypes.data_as(C.c_void_p) if len(arr) else None
    nr, used, ones = C.c_uint64(), C.c_uint64(), C.c_uint64()
    L.call("wg_pq_scan_runs", ptr, len(arr), bit_width, count, None, 0, C.byref(nr), C.byref(used), C.byref(ones))
    runs = np.zeros(max(nr.value, 1) * 4, np.uint32)
    L.call("wg_pq_scan_runs", ptr, len(arr), bit_width, count, runs.ctypes.data_as(C.c_void_p), nr.value,
           C.byref(nr), C.byref(used), C.byref(ones))
    return runs, nr.value, used.value, ones.value


def _wg_type_of(col_schema) -> Tuple[int, int, int]:
    """(physical type id, wg type, scale) of a flat column."""
    phys = {"INT32": PT_INT32, "INT64": PT_INT64, "DOUBLE": PT_DOUBLE}.get(col_schema.physical_type)
    if phys is None:
        raise L.UnsupportedError(f"parquet physical type {col_schema.physical_type} not on the GPU decode path")
    lt = col_schema.logical_type
    ltype = getattr(lt, "type", "NONE")
    if ltype == "DECIMAL":
        if phys != PT_INT64:
            raise L.UnsupportedError("decimals stored as INT32 are not on the GPU decode path")
        return phys, L.DECIMAL, col_schema.scale
    if phys == PT_INT32:
        lts = str(lt)
        if "isSigned=false" in lts.replace(" ", "") and "bitWidth=32" in lts.replace(" ", ""):
            return phys, L.UINT32, 0
        return phys, L.INT32, 0
    return phys, (L.INT64 if phys == PT_INT64 else L.FLOAT64), 0


def _chunk_pages(raw: bytes, compression: str, n_rows: int, max_def: int, width: int, arena: _Arena) -> list:
    """The data pages of one column chunk's bytes, decompressed into `arena`:
    per page its rows, encoding and the arena offsets of its streams
    (ColumnPageReader.ReadPage, storage/parquet/page_reader.go:434-520)."""
    import pyarrow as pa
    codec = None if compression == "UNCOMPRESSED" else pa.Codec(compression.lower())

    def inflate(b, size):
        # an Arrow buffer read in place (no copy into a Python bytes object)
        return b if codec is None else memoryview(codec.decompress(b, decompressed_size=size)).cast("B")

    pages, dict_ref = [], None
    pos, rows_seen = 0, 0
    while pos < len(raw) and rows_seen < n_rows:
        h, pos = _read_page_header(raw, pos)
        ptype, usize, csize = h[1], h[2], h[3]
        body = raw[pos:pos + csize]
        pos += csize
        if ptype == PAGE_DICTIONARY:
            dict_n = h[7][1]
            vals = inflate(body, usize)
            if len(vals) < dict_n * width:
                raise L.WgError("dictionary page shorter than its values")
            dict_ref = (arena.add(vals[:dict_n * width]), dict_n)
            continue
        if ptype == PAGE_DATA:
            dph = h[5]
            nv, enc = dph[1], dph[2]
            data = inflate(body, usize)
            off = 0
            levels = None
            if max_def > 0:
                ll = int.from_bytes(data[0:4], "little")
                levels = data[4:4 + ll]
                off = 4 + ll
            values = data[off:]
        elif ptype == PAGE_DATA_V2:
            dph = h[8]
            nv, enc = dph[1], dph[4]
            dl, rl = dph.get(5, 0), dph.get(6, 0)
            levels = body[rl:rl + dl] if max_def > 0 else None
            rest = body[rl + dl:]
            values = inflate(rest, usize - rl - dl) if dph.get(7, True) else bytes(rest)
        else:
            continue  # index pages carry no values
        page = dict(rows=nv, enc=enc)
        present = nv
        if levels is not None:
            runs, nr, _, present = _scan_runs(levels, 1, nv)
            if present < nv:  # a page without NULLs decodes as a required one
                page["def"] = (arena.add(levels), arena.add(runs), nr)
        if enc == ENC_PLAIN:
            if len(values) < present * width:
                raise L.WgError("PLAIN page shorter than its values")
            page["plain"] = arena.add(values[:present * width])
        elif enc in (ENC_PLAIN_DICTIONARY, ENC_RLE_DICTIONARY):
            if dict_ref is None:
                raise L.WgError("dictionary-encoded page before a dictionary page")
            bw = values[0]
            runs, nr, _, _ = _scan_runs(values[1:], bw, present)
            page["idx"] = (arena.add(values[1:]), arena.add(runs), nr, bw)
            page["dict"] = dict_ref
        else:
            raise L.UnsupportedError(f"parquet encoding {enc} not on the GPU decode path")
        pages.append(page)
        rows_seen += nv
    got = sum(pg["rows"] for pg in pages)
    if got != n_rows:
        raise L.WgError(f"column chunk pages hold {got} rows, the row group {n_rows}")
    return pages


class StagedColumn:
    """One column's chunks (all row groups, or those named) parsed and
    decompressed on the host and uploaded to HBM in one buffer: the input of
    the device decode, which `decode()` runs in one wg_parquet_decode_pages
    call. `page_bytes` = the decompressed stream bytes the decode reads."""

    def __init__(self, ctx, path: str, column: str, row_groups: Optional[Sequence[int]] = None):
        import pyarrow.parquet as pq
        f = pq.ParquetFile(path)
        md = f.metadata
        ci = f.schema_arrow.get_field_index(column)
        if ci < 0:
            raise KeyError(column)
        cs = f.schema.column(ci)
        if cs.max_repetition_level != 0:
            raise L.UnsupportedError("repeated columns are not on the GPU decode path")
        self.ctx = ctx
        self.max_def = cs.max_definition_level
        self.phys, self.type, self.scale = _wg_type_of(cs)
        self.width = 4 if self.phys == PT_INT32 else 8
        rgs = list(range(md.num_row_groups)) if row_groups is None else list(row_groups)

        def stage(rg):
            # one chunk: read, decompress, scan its streams into its own arena
            # (file reads, decompression and run scans release the GIL)
            cc = md.row_group(rg).column(ci)
            n_rows = md.row_group(rg).num_rows
            start = (cc.dictionary_page_offset if cc.has_dictionary_page and cc.dictionary_page_offset
                     else cc.data_page_offset)
            with open(path, "rb") as fh:
                fh.seek(start)
                raw = fh.read(cc.total_compressed_size)
            arena = _Arena()
            return n_rows, arena, _chunk_pages(raw, cc.compression, n_rows, self.max_def, self.width, arena)

        import concurrent.futures as cf
        # the page walk holds the GIL: past two threads the contention costs more than the overlap gains
        workers = max(1, min(len(rgs), 2, os.cpu_count() or 1))
        with cf.ThreadPoolExecutor(max_workers=workers) as ex:
            staged = list(ex.map(stage, rgs))
        # the chunks' arenas laid end to end in one host buffer, one upload
        offs, total = [], 0
        for _, arena, _ in staged:
            offs.append(total)
            total += arena.size
        host = np.zeros(max(total, 8), np.uint8)

        def place(i):
            for off, b in staged[i][1].parts:
                host[offs[i] + off:offs[i] + off + len(b)] = b

        with cf.ThreadPoolExecutor(max_workers=workers) as ex:
            list(ex.map(place, range(len(staged))))
        self.rows = sum(n for n, _, _ in staged)
        self.page_bytes = sum(len(b) for _, arena, _ in staged for _, b in arena.parts)
        self.dev = ctx.to_device(host)
        del host
        base = self.dev.ptr.value if hasattr(self.dev.ptr, "value") else int(self.dev.ptr)
        pages = []
        for (_, _, chunk_pages), off in zip(staged, offs):
            for pg in chunk_pages:
                pg = dict(pg)
                for key in ("def", "plain", "idx"):
                    if key in pg:
                        v = pg[key]
                        pg[key] = (v + off) if key == "plain" else (v[0] + off, v[1] + off) + tuple(v[2:])
                if "dict" in pg:
                    pg["dict"] = (pg["dict"][0] + off, pg["dict"][1])
                pages.append(pg)
        self.pages = (L.PqPage * max(len(pages), 1))()
        for P, pg in zip(self.pages, pages):
            P.num_rows = pg["rows"]
            P.physical_type = self.phys
            P.encoding = pg["enc"]
            if "def" in pg:
                d_off, r_off, nr = pg["def"]
                P.def_ = L.PqRuns(base + r_off, nr, 1, base + d_off)
            if "plain" in pg:
                P.plain = base + pg["plain"]
            if "idx" in pg:
                d_off, r_off, nr, bw = pg["idx"]
                P.idx = L.PqRuns(base + r_off, nr, bw, base + d_off)
                P.dict = base + pg["dict"][0]
                P.dict_n = pg["dict"][1]
        self.n_pages = len(pages)

    def decode(self, out=None, out_valid=None, row0: int = 0):
        """Decode every page on the device. With `out` given (a device buffer
        of the column's width, and `out_valid` zeroed bitmap words for a
        nullable column) the rows land at row0; else new buffers. -> Col."""
        from .gpu import Col
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
        if self.dev is not None:
            self.dev.free()
            self.dev = None


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
