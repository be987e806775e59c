"""Pins of the oracle's partial-state spill run writer / reader
(aggregate_partial_spill.go): the reference's own round-trip case, the header
offset its Close() patches, and the bytes.Compare order of decimal-text keys."""
import struct

import numpy as np

import oracle as O


def acc_rows(rows):
    a = np.zeros((len(rows), len(rows[0])), dtype=O.ACC_DTYPE)
    for g, r in enumerate(rows):
        for j, fields in enumerate(r):
            for f, v in fields.items():
                a[g, j][f] = v
    return a


def test_reference_round_trip_case():
    """TestExternalMergeSpill_PartialFileRoundTrip (aggregate_partial_spill_test.go:289-392):
    k int64 groups 1 and 2 with SUM (int) {Count 5, SumI64 100} / {7, 200},
    COUNT {5} / {7}, MIN (float) {3.14} / {2.71}."""
    res = dict(n=2, keys=np.array([[1], [2]], np.int64), key_null=np.zeros((2, 1), bool),
               acc=acc_rows([[{"count": 5, "sum_i64": 100}, {"count": 5}, {"count": 5, "min_f64": 3.14, "has_min": 1}],
                             [{"count": 7, "sum_i64": 200}, {"count": 7}, {"count": 7, "min_f64": 2.71, "has_min": 1}]]))
    specs = [(O.SUM, False, False, 0), (O.COUNT, False, False, 0), (O.MIN, True, False, 0)]
    run = O.partial_run(res, specs, [O.INT64], names=["k"], type_ids=[O.PARQUET_INT64])
    r = O.read_partial_run(run)
    assert r["cols"] == [("k", 2)] and r["specs"] == specs and r["n"] == 2
    (sk0, kv0, a0), (sk1, kv1, a1) = r["groups"]
    assert kv0 == [1] and kv1 == [2]
    assert a0 == [{"count": 5, "sum": 100}, {"count": 5}, {"has": True, "value": 3.14}]
    assert a1 == [{"count": 7, "sum": 200}, {"count": 7}, {"has": True, "value": 2.71}]
    # Close() patches numGroups at magic(5) + numGroupCols(4) + per col
    # (2 + len(name) + 2) + numAggs(4) + 10 per agg (aggregate_partial_spill.go:274-283)
    off = 5 + 4 + (2 + 1 + 2) + 4 + 10 * 3
    assert struct.unpack_from("<I", run, off)[0] == 2
    assert run[off + 4] == 0x01 and run[-1] == 0x00
    # per group: marker, key len, key, tag + int64, SUM (8 + 8), COUNT 8, MIN (1 + 8)
    assert len(run) == off + 4 + 2 * (1 + 4 + 1 + 1 + 8 + 16 + 8 + 9) + 1


def test_sort_key_order_is_bytes_compare_of_decimal_text():
    """int-mode sort keys are strconv.AppendInt text (aggregate_partial_drain_cursor.go:440-452)
    ordered by bytes.Compare: "-1" < "-2" < "0" < "10" < "100" < "9"; a NULL is "<null>"."""
    keys = np.array([[10], [9], [-1], [-2], [100], [0]], np.int64)
    nulls = np.zeros((6, 1), bool)
    res = dict(n=6, keys=keys, key_null=nulls, acc=acc_rows([[{"count": i}] for i in range(6)]))
    r = O.read_partial_run(O.partial_run(res, [(O.COUNT, False, False, 0)], [O.INT64]))
    assert [g[0] for g in r["groups"]] == [b"-1", b"-2", b"0", b"10", b"100", b"9"]
    assert [g[1][0] for g in r["groups"]] == [-1, -2, 0, 10, 100, 9]
    assert O.partial_sort_key([5, 0], [False, True]) == b"5\x00<null>"


def test_oracle_aggregate_round_trips_through_the_run():
    """A hash_agg result written as a run reads back with every accumulator
    intact: int and DECIMAL (Int128 past 2^64) sums, counts, MIN/MAX, an int32
    composite key and the flags initFlatAccums sets."""
    rng = np.random.default_rng(3)
    n = 5000
    a = rng.integers(-50, 50, n).astype(np.int32)
    b = rng.integers(0, 4, n).astype(np.int32)
    v = rng.integers(-10**6, 10**6, n).astype(np.int64)
    d = np.empty(2 * n, np.uint64)
    big = [int(x) * (1 << 70) + 12345 for x in rng.integers(-3, 4, n)]
    for i, x in enumerate(big):
        lo, hi = O.i128_split(x)
        d[2 * i], d[2 * i + 1] = np.uint64(hi & ((1 << 64) - 1)), np.uint64(lo)
    cols = [(a, O.INT32), (b, O.INT32), (v, O.INT64), (d, O.DECIMAL128, None, 2)]
    aggs = [(O.SUM, 2), (O.COUNT, -1), (O.MIN, 2), (O.MAX, 2), (O.SUM, 3), (O.AVG, 2)]
    res = O.hash_agg(cols, [0, 1], aggs)
    specs = [O.partial_spec(f, None if c < 0 else cols[c][1], 2 if c == 3 else 0) for f, c in aggs]
    assert specs[4] == (O.SUM, False, True, 2) and specs[5] == (O.AVG, True, False, 0)
    # the flags agree with the oracle's own accumulators (its initFlatAccums restatement)
    for j, (f, c) in enumerate(aggs):
        if f != O.COUNT:
            assert bool(res["acc"][0, j]["is_float"]) == specs[j][1], j
            assert bool(res["acc"][0, j]["is_decimal"]) == specs[j][2], j
    r = O.read_partial_run(O.partial_run(res, specs, [O.INT32, O.INT32]))
    got = {tuple(kv): accs for _, kv, accs in r["groups"]}
    assert len(got) == res["n"]
    for g in range(res["n"]):
        x = res["acc"][g]
        accs = got[(int(res["keys"][g, 0]), int(res["keys"][g, 1]))]
        assert accs[0] == {"count": int(x[0]["count"]), "sum": int(x[0]["sum_i64"])}
        assert accs[1] == {"count": int(x[1]["count"])}
        assert accs[2] == {"has": True, "value": int(x[2]["min_i64"])}
        assert accs[3] == {"has": True, "value": int(x[3]["max_i64"])}
        assert accs[4]["sum"] == O.i128_join(int(x[4]["sum_lo"]), int(x[4]["sum_hi"]))
        assert accs[5]["sum"] == float(x[5]["sum_f64"])
    # exact Int128 sums, checked against Python integers
    want = {}
    for i in range(n):
        want[(int(a[i]), int(b[i]))] = want.get((int(a[i]), int(b[i])), 0) + big[i]
    for k, accs in got.items():
        assert accs[4]["sum"] == want[k]


def le(n, v, signed=False):
    return int(v).to_bytes(n, "little", signed=signed)


def test_run_bytes_known_answer():
    """The run of two groups of GROUP BY k int64 with SUM(v int64), COUNT(*),
    MIN(f float64), SUM(d DECIMAL(10,2)), MAX(d) assembled byte by byte from
    the reference writer, each field at its place:
      writeHeader  aggregate_partial_spill.go:171-220 (magic, numGroupCols,
                   nameLen/name/typeID, numAggs, func/isFloat/isDecimal/decScale,
                   numGroups)
      Write        :226-257 (rowMarker, sortKeyLen, sortKey, tagged key values
                   via writeKeyValue :465-536: tag 3 + int64 LE)
      emitAcc      :302-351 (SUM/AVG: Count then SumI64 / SumF64 / SumDec as
                   Hi, Lo; COUNT: Count; MIN/MAX: HasMin byte, then the value;
                   a present decimal MIN/MAX is an Int128 Hi, Lo :702-707)
      Close        :259-290 (endMarker 0x00, numGroups patched)
    A swapped Count / Sum order, a missing has byte or an 8-byte decimal
    MAX changes the bytes."""
    specs = [(O.SUM, False, False, 0), (O.COUNT, False, False, 0), (O.MIN, True, False, 0),
             (O.SUM, False, True, 2), (O.MAX, False, True, 0)]
    rows = [[{"count": 3, "sum_i64": -7}, {"count": 3}, {"has_min": 1, "min_f64": -1.5},
             {"count": 3, "sum_lo": 2**64 - 5, "sum_hi": 1}, {"has_max": 1, "max_i64": -12}],
            [{"count": 1, "sum_i64": 40}, {"count": 1}, {"has_min": 0},
             {"count": 0, "sum_lo": 0, "sum_hi": 0}, {"has_max": 0}]]
    res = dict(n=2, keys=np.array([[10], [9]], np.int64), key_null=np.zeros((2, 1), bool), acc=acc_rows(rows))
    want = bytearray(b"WAGS\x01")
    want += le(4, 1)                                            # numGroupCols
    want += le(2, 1) + b"k" + le(2, O.PARQUET_INT64)            # nameLen, name, typeID
    want += le(4, 5)                                            # numAggs
    for func, is_f, is_d, sc in specs:                          # func u32, isFloat, isDecimal, decScale i32
        want += le(4, func) + bytes([is_f, is_d]) + le(4, sc, True)
    want += le(4, 2)                                            # numGroups
    # sort keys "10" < "9" (bytes.Compare of the decimal text)
    want += b"\x01" + le(4, 2) + b"10" + b"\x03" + le(8, 10, True)
    want += le(8, 3, True) + le(8, -7, True)                   # SUM: Count, SumI64
    want += le(8, 3, True)                                      # COUNT
    want += b"\x01" + struct.pack("<d", -1.5)                   # MIN: HasMin, MinF64
    want += le(8, 3, True) + le(8, 1, True) + le(8, 2**64 - 5)  # SUM dec: Count, Hi, Lo
    want += b"\x01" + le(8, -1, True) + le(8, 2**64 - 12)       # MAX dec: HasMax, Hi, Lo of -12
    want += b"\x01" + le(4, 1) + b"9" + b"\x03" + le(8, 9, True)
    want += le(8, 1, True) + le(8, 40, True)
    want += le(8, 1, True)
    want += b"\x00"                                             # HasMin false: no value
    want += le(8, 0, True) + le(8, 0, True) + le(8, 0)
    want += b"\x00"                                             # HasMax false
    want += b"\x00"                                             # endMarker
    run = O.partial_run(res, specs, [O.INT64], names=["k"], type_ids=[O.PARQUET_INT64])
    assert run == bytes(want)
    r = O.read_partial_run(bytes(want))
    assert r["groups"][0][2][4] == {"has": True, "value": -12}
    assert r["groups"][1][2][2] == {"has": False}
