"""Pins the CPU oracle (oracle/) before it is trusted as the checker:
the reference's own known-answer tests, brute force on small inputs in pure
Python, closed forms on constructed data, and planted mistakes that the
comparison must catch."""
import math

import numpy as np
import pytest

import oracle as O

U64 = (1 << 64) - 1


# ------------------------------------------------ Int128: decimal_test.go KATs

@pytest.mark.parametrize("v,hi,lo", [(0, 0, 0), (1, 0, 1), (-1, -1, U64), (12345, 0, 12345),
                                     (-12345, -1, (~12344) & U64), (2**63 - 1, 0, 2**63 - 1)])
def test_int128_from_kat(v, hi, lo):
    # decimal_test.go:8-30 (TestInt128From)
    assert O.i128_from_i64(v) == (lo, hi)


def test_int128_arithmetic_kat():
    # decimal_test.go:32-50
    assert O.i128_add(100, 50) == 150
    assert O.i128_add(100, O.i128_neg(50)) == 50
    assert O.i128_neg(100) == -100


def test_int128_comparison_kat():
    # decimal_test.go:52-78
    assert O.i128_less(100, 200) and not O.i128_less(200, 100)
    assert O.i128_less(-50, 100)


def test_int128_from_float_kat():
    # decimal_test.go:80-97
    assert O.i128_from_f64(123.45, 2) == 12345
    assert O.i128_from_f64(-99.99, 2) == -9999
    assert O.i128_from_f64(42.7, 0) == 43


def test_int128_to_float_kat():
    # decimal_test.go:99-111
    assert abs(O.i128_to_f64(12345, 2) - 123.45) <= 0.001
    assert abs(O.i128_to_f64(-9999, 2) - (-99.99)) <= 0.001


@pytest.mark.parametrize("v,scale,want", [(12345, 2, "123.45"), (100, 2, "1.0"), (0, 2, "0.0"), (12345, 0, "12345"),
                                          (-9999, 2, "-99.99"), (1000000, 4, "100.0"), (123456789, 6, "123.456789")])
def test_format_decimal_kat(v, scale, want):
    # decimal_test.go:113-133
    assert O.i128_format(v, scale) == want


@pytest.mark.parametrize("s,scale,want", [("123.45", 2, 12345), ("-99.99", 2, -9999), ("100", 2, 10000),
                                          ("0.5", 2, 50), ("0.005", 3, 5), ("  42.0  ", 1, 420)])
def test_parse_decimal_kat(s, scale, want):
    # decimal_test.go:135-155
    assert O.parse_decimal(s, scale) == want


def test_int128_random_vs_python_ints():
    """Add/neg/less agree with exact Python integers mod 2^128; ToFloat64
    follows decimal.go:39-50 operation by operation."""
    rng = np.random.default_rng(0)
    M = 1 << 128
    def wrap(x):
        x %= M
        return x - M if x >= M // 2 else x
    for _ in range(2000):
        a = int(rng.integers(-2**62, 2**62)) * int(rng.integers(1, 2**62)) * int(rng.choice([1, -1]))
        b = int(rng.integers(-2**63, 2**63 - 1))
        a, b = wrap(a), wrap(b)
        assert O.i128_add(a, b) == wrap(a + b)
        assert O.i128_neg(a) == wrap(-a)
        assert O.i128_less(a, b) == (a < b)
        s = int(rng.integers(0, 6))
        lo, hi = O.i128_split(a)
        if -2**63 <= a < 2**63:
            want = float(a) / float(10**s)
        else:
            want = (float(hi) * 2.0**64 + float(lo)) / float(10**s)
        got = O.i128_to_f64(a, s)
        assert got == want or (math.isnan(got) and math.isnan(want))


# --------------------------------------------- aggregate: brute force / KATs

def brute_agg(cols, keys, aggs, sel=None, pred=None):
    """Pure-Python restatement of agg_scatter.go over dicts, row order."""
    n = len(cols[0][0])
    rows = range(n) if sel is None else [int(x) for x in sel]
    def valid(c, r):
        v = c[2] if len(c) > 2 else None
        return v is None or bool((int(v[r >> 6]) >> (r & 63)) & 1)
    groups = {}
    order = []
    for r in rows:
        if pred is not None:
            pc, op, ci, cf = pred
            c = cols[pc]
            if not valid(c, r):
                continue
            x = c[0][r]
            k = float(x) if c[1] == O.FLOAT64 else int(x)
            cst = cf if c[1] == O.FLOAT64 else ci
            ok = [k == cst, k != cst, k < cst, k <= cst, k > cst, k >= cst][op]
            if not ok:
                continue
        key = tuple((None if not valid(cols[k], r) else int(cols[k][0][r])) for k in keys)
        if key not in groups:
            groups[key] = [dict(count=0, si=0, sf=0.0, sd=0, mn=None, mx=None) for _ in aggs]
            order.append(key)
        for a, (f, c) in enumerate(aggs):
            st = groups[key][a]
            if c < 0:
                st["count"] += 1
                continue
            col = cols[c]
            if not valid(col, r):
                continue
            t = col[1]
            v = col[0][r]
            if f == O.COUNT:
                st["count"] += 1
            elif f in (O.SUM, O.AVG):
                if t == O.FLOAT64:
                    st["sf"] += float(v); st["count"] += 1
                elif t == O.DECIMAL:
                    st["sd"] += int(v); st["count"] += 1
                elif t == O.INT64 and f == O.AVG:
                    st["sf"] += float(int(v)); st["count"] += 1
                elif t in (O.INT64, O.INT32, O.UINT16):
                    st["si"] = ((st["si"] + int(v) + 2**63) % 2**64) - 2**63; st["count"] += 1
            elif f == O.MIN and t != O.DECIMAL:
                x = float(v) if t == O.FLOAT64 else int(v)
                if st["mn"] is None or x < st["mn"]:
                    st["mn"] = x
            elif f == O.MAX and t != O.DECIMAL:
                x = float(v) if t == O.FLOAT64 else int(v)
                if st["mx"] is None or x > st["mx"]:
                    st["mx"] = x
    return order, groups


def check_against_brute(cols, keys, aggs, sel=None, pred=None):
    r = O.hash_agg(cols, keys, aggs, sel=sel, pred=pred)
    order, groups = brute_agg(cols, keys, aggs, sel, pred)
    assert r["n"] == len(order)
    for g, key in enumerate(order):  # same first-seen order
        got_key = tuple(None if r["key_null"][g, k] else int(r["keys"][g, k]) for k in range(len(keys)))
        assert got_key == key
        for a, (f, c) in enumerate(aggs):
            acc = r["acc"][g, a]
            st = groups[key][a]
            assert acc["count"] == st["count"] or f in (O.MIN, O.MAX)
            if f in (O.SUM, O.AVG) and c >= 0:
                t = cols[c][1]
                if t == O.DECIMAL:
                    assert O.i128_join(int(acc["sum_lo"]), int(acc["sum_hi"])) == st["sd"]
                elif t == O.FLOAT64 or (t == O.INT64 and f == O.AVG):
                    assert acc["sum_f64"] == st["sf"]  # same order, bit-identical
                elif t in (O.INT64, O.INT32, O.UINT16):
                    assert int(acc["sum_i64"]) == st["si"]
            if f == O.MIN and st["mn"] is not None:
                assert acc["has_min"] == 1
                assert (acc["min_f64"] if cols[c][1] == O.FLOAT64 else int(acc["min_i64"])) == st["mn"]
            if f == O.MAX and st["mx"] is not None:
                assert (acc["max_f64"] if cols[c][1] == O.FLOAT64 else int(acc["max_i64"])) == st["mx"]
            if f == O.MIN and st["mn"] is None and c >= 0:
                assert acc["has_min"] == 0


def test_hash_agg_brute_force_all_types():
    rng = np.random.default_rng(1)
    n = 3000
    k1 = rng.integers(-5, 5, n).astype(np.int64)
    k2 = rng.integers(0, 3, n).astype(np.int32)
    i32 = rng.integers(-2**31, 2**31 - 1, n).astype(np.int32)
    i64 = rng.integers(-2**63, 2**63 - 1, n, dtype=np.int64)
    f64 = rng.standard_normal(n)
    dec = rng.integers(-10**17, 10**17, n).astype(np.int64)
    u16 = rng.integers(0, 2**16, n).astype(np.uint16)
    from wadjet_amd.gpu import bitmap_from_mask
    m = [bitmap_from_mask(rng.random(n) > 0.1) for _ in range(7)]
    cols = [(k1, O.INT64, m[0]), (k2, O.INT32), (i32, O.INT32, m[1]), (i64, O.INT64, m[2]), (f64, O.FLOAT64, m[3]),
            (dec, O.DECIMAL, m[4], 2), (u16, O.UINT16, m[5])]
    aggs = [(O.COUNT, -1)] + [(f, c) for c in range(2, 7) for f in (O.SUM, O.COUNT, O.MIN, O.MAX, O.AVG)]
    check_against_brute(cols, [0, 1], aggs)
    sel = np.sort(rng.choice(n, n // 3, replace=False)).astype(np.uint32)
    check_against_brute(cols, [0], aggs, sel=sel, pred=(4, O.GT, 0, -0.5))
    check_against_brute(cols, [1], aggs, pred=(2, O.LE, 12345, 0.0))


def test_hash_aggregate_kat_exec_test():
    # exec_test.go:39-96 (string keys "a"/"b" coded as ints): SUM(amount) of
    # group a = 40.0, COUNT(amount) = 2.
    g = np.array([0, 1, 0, 1], np.int64)
    amount = np.array([10.0, 20.0, 30.0, 40.0])
    r = O.hash_agg([(g, O.INT64), (amount, O.FLOAT64)], [0], [(O.SUM, 1), (O.COUNT, 1)])
    assert r["n"] == 2
    a = int(np.nonzero(r["keys"][:, 0] == 0)[0][0])
    v, ok = O.finalize(r["acc"][:, 0], O.SUM, O.FLOAT64)
    assert ok[a] and v[a] == 40.0
    c, _ = O.finalize(r["acc"][:, 1], O.COUNT, O.FLOAT64)
    assert c[a] == 2


def test_packed_near_unique_closed_form():
    # bench_test.go:430-509 dataset: row r -> wid = r*2654435761, refresh = r&1,
    # width = 1024 + r%512; every row is its own group: COUNT 1, SUM(refresh)
    # = r&1, AVG(width) = 1024 + r%512 exactly.
    rows = np.arange(16 * 2048, dtype=np.int64)
    wid = rows * 2654435761
    refresh = rows & 1
    width = 1024 + rows % 512
    r = O.hash_agg([(wid, O.INT64), (refresh, O.INT64), (width, O.INT64)], [0],
                   [(O.COUNT, -1), (O.SUM, 1), (O.AVG, 2)])
    assert r["n"] == len(rows)
    src = r["keys"][:, 0] // 2654435761
    assert (r["acc"][:, 0]["count"] == 1).all()
    assert (r["acc"][:, 1]["sum_i64"] == (src & 1)).all()
    avg, ok = O.finalize(r["acc"][:, 2], O.AVG, O.INT64)
    assert ok.all() and (avg == 1024 + src % 512).all()


def test_min_max_decimal_and_sum_ipv4_are_null():
    # agg_scatter.go:519-578 has no MIN/MAX case for DECIMAL and no SUM case
    # for IPv4: those accumulators stay empty and finalize to NULL.
    k = np.zeros(4, np.int64)
    d = np.array([5, -7, 9, 1], np.int64)
    ip = np.array([1, 2, 3, 4], np.int64)
    r = O.hash_agg([(k, O.INT64), (d, O.DECIMAL, None, 2), (ip, O.IPV4)], [0],
                   [(O.MIN, 1), (O.MAX, 1), (O.SUM, 2), (O.MIN, 2)])
    for a, (f, t) in enumerate([(O.MIN, O.DECIMAL), (O.MAX, O.DECIMAL), (O.SUM, O.IPV4)]):
        _, ok = O.finalize(r["acc"][:, a], f, t, 2)
        assert not ok[0]
    v, ok = O.finalize(r["acc"][:, 3], O.MIN, O.IPV4)
    assert ok[0] and v[0] == 1


def test_parallel_baseline_equals_serial_on_integers():
    rng = np.random.default_rng(2)
    n = 200_000
    k = rng.integers(0, 777, n).astype(np.int64)
    v = rng.integers(-10**12, 10**12, n).astype(np.int64)
    cols = [(k, O.INT64), (v, O.INT64)]
    aggs = [(O.SUM, 1), (O.COUNT, -1), (O.MIN, 1), (O.MAX, 1)]
    a = O.hash_agg(cols, [0], aggs)
    b = O.hash_agg(cols, [0], aggs, threads=4)
    oa, ob = np.argsort(a["keys"][:, 0]), np.argsort(b["keys"][:, 0])
    for f in ("sum_i64", "count", "min_i64", "max_i64"):
        np.testing.assert_array_equal(a["acc"][oa][f], b["acc"][ob][f])


# ----------------------------------------------------------------- filter

@pytest.mark.parametrize("op", range(6))
def test_filter_matches_numpy(op):
    rng = np.random.default_rng(op)
    n = 5000
    f = rng.standard_normal(n)
    f[::17] = np.nan
    i = rng.integers(-100, 100, n).astype(np.int32)
    ops = [np.equal, np.not_equal, np.less, np.less_equal, np.greater, np.greater_equal]
    got = O.filter_rows((f, O.FLOAT64), n, op, 0, -0.5)
    np.testing.assert_array_equal(got, np.nonzero(ops[op](f, -0.5))[0])
    got = O.filter_rows((i, O.INT32), n, op, 7, 0.0)
    np.testing.assert_array_equal(got, np.nonzero(ops[op](i, 7))[0])
    sel = np.arange(0, n, 3, dtype=np.uint32)
    got = O.filter_rows((i, O.INT32), n, op, 7, 0.0, sel=sel)
    np.testing.assert_array_equal(got, sel[ops[op](i[sel], 7)])


# ------------------------------------------------------------------- join

def test_join_brute_force_with_duplicates_and_nulls():
    rng = np.random.default_rng(4)
    nb, npr = 500, 2000
    bk = rng.integers(0, 300, nb).astype(np.int64)
    pk = rng.integers(0, 400, npr).astype(np.int32)
    from wadjet_amd.gpu import bitmap_from_mask
    bm = rng.random(nb) > 0.05
    pm = rng.random(npr) > 0.05
    p, b = O.hash_join((bk, O.INT64, bitmap_from_mask(bm)), (pk, O.INT32, bitmap_from_mask(pm)))
    want = sorted((i, j) for i in range(npr) if pm[i] for j in range(nb) if bm[j] and bk[j] == pk[i])
    assert sorted(zip(p.tolist(), b.tolist())) == want
    # chains are LIFO: for each probe row the build rows come newest first
    for i in np.unique(p):
        bs = b[p == i]
        assert (np.diff(bs.astype(np.int64)) < 0).all()


@pytest.mark.parametrize("threads", [1, 7])
def test_join_agg_equals_aggregated_pairs(threads):
    """or_join_agg_parallel == or_hash_join's pairs, each joined row's
    qty * p2 formed as an exact Python int and summed per p1 group; duplicate
    build keys (chains) and NULLs in every input column. A dropped carry or a
    truncated 64-bit product would change the Int128 sums."""
    from wadjet_amd.gpu import bitmap_from_mask
    rng = np.random.default_rng(20 + threads)
    nb, npr = 3000, 40_000
    bk = rng.integers(0, 2500, nb).astype(np.int64)
    g = rng.integers(-3, 9, nb).astype(np.int32)
    d = rng.integers(-2**62, 2**62, nb).astype(np.int64)
    pk = rng.integers(0, 3000, npr).astype(np.int64)
    q = rng.integers(-2**31, 2**31 - 1, npr).astype(np.int32)
    bm, pm, dm, qm = (rng.random(x) > 0.03 for x in (nb, npr, nb, npr))
    r = O.join_agg((bk, O.INT64, bitmap_from_mask(bm)), (g, O.INT32), (d, O.DECIMAL, bitmap_from_mask(dm), 2),
                   (pk, O.INT64, bitmap_from_mask(pm)), (q, O.INT32, bitmap_from_mask(qm)), threads=threads)
    p, b = O.hash_join((bk, O.INT64, bitmap_from_mask(bm)), (pk, O.INT64, bitmap_from_mask(pm)), cap=npr * 20)
    want = {}
    for i, j in zip(p.tolist(), b.tolist()):
        c, s = want.get(int(g[j]), (0, 0))
        if qm[i] and dm[j]:
            s += int(q[i]) * int(d[j])
        want[int(g[j])] = (c + 1, s)
    keys = sorted(want)
    assert r["n"] == len(keys) and r["keys"].tolist() == keys
    for x, k in enumerate(keys):
        c, s = want[k]
        assert r["count"][x] == c
        assert O.i128_join(int(r["sum_lo"][x]), int(r["sum_hi"][x])) == s % (1 << 128) - ((s % (1 << 128)) >> 127 << 128)
    assert sum(r["count"]) == len(p)


def test_join_kats():
    # join_test.go:10-56: probe alice, bob, alice, unknown vs build alice, bob,
    # carol (strings coded 1, 2, 1, 9 / 1, 2, 3) -> 3 rows
    p, b = O.hash_join((np.array([1, 2, 3], np.int64), O.INT64), (np.array([1, 2, 1, 9], np.int64), O.INT64))
    assert len(p) == 3
    # join_coverage_test.go:101-141: build ids 1, 2; probe 1, 3 -> 1 row
    p, b = O.hash_join((np.array([1, 2], np.int64), O.INT64), (np.array([1, 3], np.int64), O.INT64))
    assert list(zip(p, b)) == [(0, 0)]
    # dual_int_test.go:11-80 on the packed pair: 2 matches of 3 probe rows
    bk = np.array([1 | 10 << 32, 2 | 20 << 32, 3 | 30 << 32], np.int64)
    pk = np.array([1 | 10 << 32, 2 | 20 << 32, 99 | 99 << 32], np.int64)
    p, b = O.hash_join((bk, O.INT64), (pk, O.INT64))
    assert len(p) == 2 and (bk[b] == pk[p]).all()


def test_digest_is_order_free_and_restated():
    rng = np.random.default_rng(6)
    rows = rng.integers(-2**63, 2**63 - 1, (1000, 4), dtype=np.int64)
    d1 = O.digest_rows(rows)
    d2 = O.digest_rows(rows[rng.permutation(1000)])
    assert d1 == d2
    M = (1 << 64) - 1
    def mix(k):
        k ^= k >> 33; k = (k * 0xff51afd7ed558ccd) & M
        k ^= k >> 33; k = (k * 0xc4ceb9fe1a85ec53) & M
        return k ^ (k >> 33)
    a = b = 0
    for r in rows[:50]:
        x, y = 0x243F6A8885A308D3, 0x13198A2E03707344
        for v in r:
            v = int(v) & M
            x = mix(x ^ v)
            y = mix((y + v * 0x9E3779B97F4A7C15) & M)
        a, b = (a + x) & M, (b + y) & M
    assert O.digest_rows(rows[:50]) == (a, b)
    # a single changed value changes the digest
    r2 = rows.copy(); r2[7, 2] ^= 1
    assert O.digest_rows(r2) != d1


# ------------------------------------------------ planted mistakes are caught

def test_comparison_catches_planted_mistakes():
    """The GPU-vs-oracle comparison must fail on a dropped row, a wrong sign,
    a swapped key and an Int128 carry error planted in the oracle's output."""
    from helpers import compare, oracle_groups
    rng = np.random.default_rng(8)
    n = 2000
    k = rng.integers(0, 50, n).astype(np.int64)
    d = rng.integers(-10**18, 10**18, n).astype(np.int64)
    f = rng.standard_normal(n)
    cols = [(k, O.INT64), (d, O.DECIMAL, None, 2), (f, O.FLOAT64)]
    aggs = [(O.SUM, 1), (O.COUNT, -1), (O.MIN, 2), (O.AVG, 2)]
    types = [O.INT64, O.DECIMAL, O.FLOAT64]
    ref = oracle_groups(O.hash_agg(cols, [0], aggs))

    def as_gpu(r):
        vals, valid, state = [], [], []
        for a, (fn, c) in enumerate(aggs):
            v, ok = O.finalize(r["acc"][:, a], fn, types[c] if c >= 0 else O.INT64, 2 if c == 1 else 0)
            vals.append(v); valid.append(ok)
            state.append(dict(sum_lo=r["acc"][:, a]["sum_lo"].copy(), sum_hi=r["acc"][:, a]["sum_hi"].copy()))
        return dict(n=r["n"], keys=r["keys"].copy(), nulls=r["nulls"].copy(), values=vals, valid=valid, state=state)

    compare(as_gpu(ref), ref, aggs, types, [0, 2, 0])  # identical passes
    bad = as_gpu(ref); bad["values"][1][3] += 1          # COUNT off by one
    with pytest.raises(AssertionError):
        compare(bad, ref, aggs, types, [0, 2, 0])
    bad = as_gpu(ref); bad["state"][0]["sum_hi"][5] += 1  # Int128 carry lost
    with pytest.raises(AssertionError):
        compare(bad, ref, aggs, types, [0, 2, 0])
    bad = as_gpu(ref); bad["values"][3][2] = -bad["values"][3][2]  # float sign
    with pytest.raises(AssertionError):
        compare(bad, ref, aggs, types, [0, 2, 0])
    bad = as_gpu(ref); bad["keys"][[1, 2]] = bad["keys"][[2, 1]]  # swapped keys
    with pytest.raises(AssertionError):
        compare(bad, ref, aggs, types, [0, 2, 0])
