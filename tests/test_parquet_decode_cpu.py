"""Host half of the device Parquet decode: the Thrift compact page-header
reader and the hybrid run scanner (wg_pq_scan_runs, host code in the
library) against pyarrow-written files and hand-built streams."""
import ctypes as C

import numpy as np
import pytest

from wadjet_amd import _lib as L
from wadjet_amd import parquet as PQ


def scan(buf: bytes, bw: int, count: int):
    return PQ._scan_runs(buf, bw, count)


def expand(buf: bytes, bw: int, count: int) -> np.ndarray:
    """Values of a hybrid stream from its run table (the device's rule, on
    the host): RLE value, or bits LSB-first from the bit offset."""
    runs, nr, _, _ = scan(buf, bw, count)
    bits = np.unpackbits(np.frombuffer(buf + b"\0" * 8, np.uint8), bitorder="little")
    out = np.zeros(count, np.int64)
    for r in range(nr):
        start, cnt, payload, packed = runs[4 * r:4 * r + 4]
        for j in range(cnt):
            if packed:
                b = payload + j * bw
                out[start + j] = int(sum(int(bits[b + k]) << k for k in range(bw)))
            else:
                out[start + j] = payload
    return out


def test_hand_built_streams():
    """RLE run of 5 x 3 (header 5<<1, value byte), then a bit-packed group of
    8 values at width 3 (header 1<<1|1, 3 bytes): the Encodings.md example
    values 0..7 packed LSB-first are 0x88 0xC6 0xFA."""
    s = bytes([10, 3, 3, 0x88, 0xC6, 0xFA])
    assert expand(s, 3, 13).tolist() == [3] * 5 + list(range(8))
    runs, nr, used, _ = scan(s, 3, 13)
    assert nr == 2 and used == len(s)
    # width 1: the count of ones (present rows of a definition-level stream)
    lv = bytes([3, 0b10110101])  # one bit-packed group: 1,0,1,0,1,1,0,1
    assert scan(lv, 1, 8)[3] == 5
    assert scan(lv, 1, 6)[3] == 4  # only the first `count` values count
    with pytest.raises(L.WgError):
        scan(bytes([10]), 3, 5)  # RLE value missing


def test_page_headers_of_a_pyarrow_file(tmp_path):
    """Every page header of a pyarrow file parses; data pages add up to the
    row group's rows; the dictionary page comes first."""
    import pyarrow as pa
    import pyarrow.parquet as pq
    path = str(tmp_path / "h.parquet")
    n = 50_000
    pq.write_table(pa.table({"k": np.arange(n, dtype=np.int64) % 777}), path, data_page_size=4096,
                   compression="snappy")
    md = pq.ParquetFile(path).metadata
    cc = md.row_group(0).column(0)
    with open(path, "rb") as fh:
        fh.seek(cc.dictionary_page_offset)
        raw = fh.read(cc.total_compressed_size)
    pos, kinds, rows = 0, [], 0
    while pos < len(raw):
        h, pos = PQ._read_page_header(raw, pos)
        kinds.append(h[1])
        pos += h[3]
        if h[1] == PQ.PAGE_DATA:
            rows += h[5][1]
    assert kinds[0] == PQ.PAGE_DICTIONARY and rows == n and len(kinds) > 3
