"""Pins of the oracle's shuffle routing (oracle.shuffle_hash / shuffle_partition,
restating hashRowsIntoPartitions, internal/worker/partitioned_shuffle_sink.go:
1039-1171, and the perPartRows grouping, :272-290)."""
import numpy as np
import pytest

import oracle as O


def test_fnv1a64_published_vectors():
    """FNV-1a 64 known answers (the published FNV test suite): these separate
    FNV-1a from FNV-1 (multiply before xor: "a" -> 0xaf63bd4c8601b7be) and pin
    the offset basis and prime."""
    assert O.fnv1a64(b"") == 0xCBF29CE484222325
    assert O.fnv1a64(b"a") == 0xAF63DC4C8601EC8C
    assert O.fnv1a64(b"foobar") == 0x85944171F73967E8
    assert O.fnv1a64(b"a") != 0xAF63BD4C8601B7BE


def _rows_bytes_hash(cols, key_cols, rows):
    out = []
    for r in map(int, rows):
        vals, types, nulls = [], [], []
        for c in key_cols:
            data, t = cols[c][0], cols[c][1]
            valid = cols[c][2] if len(cols[c]) > 2 else None
            vals.append(data[r])
            types.append(t)
            nulls.append(valid is not None and not ((int(valid[r >> 6]) >> (r & 63)) & 1))
        out.append(O.fnv1a64(O.shuffle_row_bytes(vals, types, nulls)))
    return np.array(out, np.uint64)


def _mixed_cols(n, seed, nulls=True):
    rng = np.random.default_rng(seed)
    def bm(p):
        if not nulls:
            return None
        m = rng.random(n) > p
        return np.packbits(np.concatenate([m, np.zeros((-n) % 64, bool)]), bitorder="little").view(np.uint64)
    f = rng.normal(size=n)
    f[:3] = [0.0, -0.0, np.inf]
    return [
        (rng.integers(-2**63, 2**63 - 1, n, dtype=np.int64), O.INT64, bm(0.1)),
        (rng.integers(-2**31, 2**31 - 1, n, dtype=np.int64).astype(np.int32), O.INT32, bm(0.1)),
        (rng.integers(0, 65536, n).astype(np.uint16), O.UINT16, bm(0.1)),
        (rng.integers(0, 2**32, n, dtype=np.int64).astype(np.uint32), O.UINT32, bm(0.1)),
        (rng.integers(0, 2**32, n, dtype=np.int64), O.IPV4, bm(0.1)),
        (f, O.FLOAT64, bm(0.1)),
    ]


@pytest.mark.parametrize("key_cols", [[0], [1], [2], [3], [4], [5], [0, 1], [1, 2, 5], [4, 3, 2, 1, 0]])
def test_vectorised_hash_equals_the_per_row_byte_stream(key_cols):
    """The numpy restatement against a byte-at-a-time FNV-1a over each row's
    key bytes (the hashVectorValue form the reference's routing test uses):
    little-endian storage bytes, storage widths of the reference's vectors, 0xff
    for NULL. A wrong width, byte order or NULL marker fails here."""
    cols = _mixed_cols(500, 7)
    sel = np.sort(np.random.default_rng(1).choice(500, 300, replace=False)).astype(np.uint32)
    got = O.shuffle_hash(cols, key_cols, sel)
    assert np.array_equal(got, _rows_bytes_hash(cols, key_cols, sel.astype(np.int64)))


def test_storage_widths_of_the_reference_vectors():
    """A compact uint32 IPv4 hashes as the 8-byte Int64Data it is in the
    reference; a compact uint16 Port as the 4-byte Int32Data; int32 as its
    4 two's-complement bytes."""
    ip = np.array([0xC0A80001], np.uint32)
    assert O.shuffle_hash([(ip, O.UINT32)], [0])[0] == O.fnv1a64((0xC0A80001).to_bytes(8, "little"))
    assert O.shuffle_hash([(ip.astype(np.int64), O.IPV4)], [0])[0] == O.fnv1a64((0xC0A80001).to_bytes(8, "little"))
    port = np.array([443], np.uint16)
    assert O.shuffle_hash([(port, O.UINT16)], [0])[0] == O.fnv1a64((443).to_bytes(4, "little"))
    neg = np.array([-2], np.int32)
    assert O.shuffle_hash([(neg, O.INT32)], [0])[0] == O.fnv1a64(b"\xfe\xff\xff\xff")


def test_decimal_keys_take_the_default_case():
    """A decimal key column mixes one 0x00 byte per row, NULL or not
    (partitioned_shuffle_sink.go:1158-1162): all its rows route alike."""
    d = np.array([5, -7, 1 << 40], np.int64)
    valid = np.array([0b101], np.uint64)  # row 1 NULL
    h = O.shuffle_hash([(d, O.DECIMAL, valid)], [0])
    assert np.all(h == np.uint64(O.fnv1a64(b"\x00")))


def test_no_key_columns_route_everything_to_partition_zero():
    parts, rows, offs = O.shuffle_partition([(np.arange(10), O.INT64)], [], 4)
    assert np.all(parts == 0) and list(rows) == list(range(10)) and list(offs) == [0, 10, 10, 10, 10]


def test_reference_round_trip_case():
    """TestPartitionedShuffleSink_RoundTrip (partitioned_shuffle_sink_test.go:
    19-78): 1000 sequential int64 keys over 4 partitions; every row sits in
    the partition its own FNV-1a of the 8 little-endian key bytes names, and
    no row is lost."""
    k = np.arange(1000, dtype=np.int64)
    parts, rows, offs = O.shuffle_partition([(k, O.INT64)], [0], 4)
    assert offs[-1] == 1000 and sorted(rows.tolist()) == list(range(1000))
    for p in range(4):
        for r in rows[offs[p]:offs[p + 1]]:
            assert O.fnv1a64(int(k[r]).to_bytes(8, "little", signed=True)) % 4 == p


@pytest.mark.parametrize("num_parts", [1, 3, 24, 1000])
def test_grouping_is_stable_and_complete(num_parts):
    """perPartRows appends in input order: within a partition rows ascend;
    offsets are the running counts; with a Sel the rows are Sel values."""
    cols = _mixed_cols(5000, num_parts, nulls=True)
    sel = np.sort(np.random.default_rng(num_parts).choice(5000, 3001, replace=False)).astype(np.uint32)
    parts, rows, offs = O.shuffle_partition(cols, [0, 1], num_parts, sel)
    assert len(rows) == len(sel) and offs[-1] == len(sel)
    assert np.array_equal(np.sort(rows), sel)
    pos = {int(s): i for i, s in enumerate(sel)}
    for p in range(num_parts):
        seg = rows[offs[p]:offs[p + 1]]
        assert np.all(np.diff(seg.astype(np.int64)) > 0)
        assert all(parts[pos[int(r)]] == p for r in seg)


def test_mod_zero_counts_as_one():
    """numParts 0: `mod == 0 -> 1` (partitioned_shuffle_sink.go:1163-1166)."""
    parts, rows, offs = O.shuffle_partition([(np.arange(5), O.INT64)], [0], 0)
    assert np.all(parts == 0) and list(offs) == [0, 5]
