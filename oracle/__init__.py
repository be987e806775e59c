"""CPU oracle for the HIP hash aggregate / hash join (TEST INFRASTRUCTURE).

A C restatement of wadjet's operator semantics (oracle/oracle.c, built into
oracle/liboracle.so by oracle/Makefile) plus numpy helpers. Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this package;
the product library never does.

Pinning (how each function is checked, see tests/test_oracle.py):
  Int128 add/neg/less/to_f64/from_f64/format/parse   reference KATs (decimal_test.go:8-157)
                                                     + Python-int arithmetic on random values
  or_hash_agg integer/decimal/count/min/max          brute force in pure Python dicts on small
                                                     inputs, closed forms on constructed data,
                                                     reference KATs (exec_test.go:39-96, dual_int_test.go)
  or_hash_agg float64 SUM/AVG                        math.fsum-based bound + row-order Python loop
  or_filter                                          numpy boolean masks
  or_hash_join                                       brute-force nested loop, reference KATs
                                                     (join_test.go:10-56, join_coverage_test.go:101-141)
  or_join_agg_parallel                               or_hash_join's pairs aggregated by or_hash_agg over
                                                     the materialized joined rows (exact Python-int
                                                     products), at 1 and 7 threads; the matched-row count
                                                     of the fk generator's closed form
  or_semi_join                                       brute-force membership loop with NULLs, reference
                                                     KATs (join_coverage_test.go:272-351, :711-735)
  or_digest_rows                                     order invariance + a Python restatement
  partial_run / read_partial_run                     the reference's round-trip case
                                                     (aggregate_partial_spill_test.go:289-392), the
                                                     numGroups offset Close() patches (:259-290),
                                                     bytes.Compare order of decimal-text keys
  bloom_build / bloom_filter                         no false negatives, the size rule, bit positions
                                                     recomputed per key (tests/test_bloom_oracle.py)
  fnv1a64 / shuffle_hash / shuffle_partition         published FNV-1a 64 vectors, a byte-at-a-time FNV-1a
                                                     over each row's key bytes, the reference's routing
                                                     test (partitioned_shuffle_sink_test.go:19-78)
  view_flatten / view_compose                        the reference's view KATs (batch/view_test.go:53-93,
                                                     :155-218): permutation + duplication over NULL base
                                                     rows, own-NULL override, view-over-view composition,
                                                     NaN, empty and single; flatten(compose) equals
                                                     flatten(flatten) on random views
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

INT32, INT64, FLOAT64, DECIMAL, DECIMAL128, IPV4, UINT32, UINT16 = range(1, 9)
SUM, COUNT, MIN, MAX, AVG = range(5)
EQ, NE, LT, LE, GT, GE = range(6)


class OrCol(C.Structure):
    _fields_ = [("data", C.c_void_p), ("validity", C.c_void_p), ("type", C.c_int32), ("scale", C.c_int32)]


class OrAcc(C.Structure):
    _fields_ = [("count", C.c_int64), ("sum_i64", C.c_int64), ("sum_f64", C.c_double), ("sum_lo", C.c_uint64),
                ("sum_hi", C.c_int64), ("min_i64", C.c_int64), ("max_i64", C.c_int64), ("min_f64", C.c_double),
                ("max_f64", C.c_double), ("has_min", C.c_int32), ("has_max", C.c_int32), ("is_float", C.c_int32),
                ("is_decimal", C.c_int32)]


ACC_DTYPE = np.dtype([("count", np.int64), ("sum_i64", np.int64), ("sum_f64", np.float64), ("sum_lo", np.uint64),
                      ("sum_hi", np.int64), ("min_i64", np.int64), ("max_i64", np.int64), ("min_f64", np.float64),
                      ("max_f64", np.float64), ("has_min", np.int32), ("has_max", np.int32),
                      ("is_float", np.int32), ("is_decimal", np.int32)])
assert ACC_DTYPE.itemsize == C.sizeof(OrAcc)

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        P, U64, I64, I32, D = C.c_void_p, C.c_uint64, C.c_int64, C.c_int32, C.c_double
        L.or_i128_from_i64.argtypes = [I64, C.POINTER(U64), C.POINTER(I64)]
        L.or_i128_add.argtypes = [U64, I64, U64, I64, C.POINTER(U64), C.POINTER(I64)]
        L.or_i128_neg.argtypes = [U64, I64, C.POINTER(U64), C.POINTER(I64)]
        L.or_i128_less.argtypes = [U64, I64, U64, I64]
        L.or_i128_less.restype = C.c_int
        L.or_i128_to_f64.argtypes = [U64, I64, C.c_int]
        L.or_i128_to_f64.restype = D
        L.or_i128_from_f64.argtypes = [D, C.c_int, C.POINTER(U64), C.POINTER(I64)]
        L.or_i128_format.argtypes = [U64, I64, C.c_int, C.c_char_p, C.c_int]
        L.or_parse_decimal.argtypes = [C.c_char_p, C.c_int]
        L.or_parse_decimal.restype = I64
        L.or_hash_agg.argtypes = [P, I32, U64, P, U64, I32, I32, I64, D, P, I32, P, P, I32, P, P, P, I64]
        L.or_hash_agg.restype = I64
        L.or_hash_agg_parallel.argtypes = [P, I32, U64, I32, I32, I64, D, P, I32, P, P, I32, I32, P, P, I64]
        L.or_hash_agg_parallel.restype = I64
        L.or_finalize.argtypes = [P, I32, I32, I32, C.POINTER(D), C.POINTER(I64)]
        L.or_finalize.restype = C.c_int
        L.or_filter.argtypes = [P, U64, P, U64, I32, I64, D, P]
        L.or_filter.restype = U64
        L.or_hash_join.argtypes = [P, U64, P, U64, P, U64, P, P, I64]
        L.or_semi_join.argtypes = [P, U64, P, U64, P, U64, C.c_int, P]
        L.or_semi_join.restype = I64
        L.or_hash_join.restype = I64
        L.or_digest_rows.argtypes = [P, U64, I32, P]
        L.or_mix64.argtypes = [U64]
        L.or_mix64.restype = U64
        L.or_join_agg_parallel.argtypes = [P, P, P, U64, P, P, U64, I32, P, P, P, P, I64]
        L.or_join_agg_parallel.restype = I64
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def make_cols(cols):
    """cols: list of (np.ndarray, type[, validity words[, scale]])."""
    keep = []
    arr = (OrCol * len(cols))()
    for i, c in enumerate(cols):
        data, t = c[0], c[1]
        valid = c[2] if len(c) > 2 else None
        scale = c[3] if len(c) > 3 else 0
        data = np.ascontiguousarray(data)
        keep.append(data)
        if valid is not None:
            valid = np.ascontiguousarray(valid, dtype=np.uint64)
            keep.append(valid)
        arr[i] = OrCol(_ptr(data), _ptr(valid), t, scale)
    return arr, keep


# ------------------------------------------------------------------ Int128

def i128_split(v: int):
    v &= (1 << 128) - 1
    lo = v & ((1 << 64) - 1)
    hi = v >> 64
    if hi >= 1 << 63:
        hi -= 1 << 64
    return lo, hi


def i128_join(lo: int, hi: int) -> int:
    v = (int(hi) << 64) | (int(lo) & ((1 << 64) - 1))
    if v >= 1 << 127:
        v -= 1 << 128
    return v


def i128_from_i64(v: int):
    lo, hi = C.c_uint64(), C.c_int64()
    lib().or_i128_from_i64(v, C.byref(lo), C.byref(hi))
    return lo.value, hi.value


def i128_add(a: int, b: int) -> int:
    alo, ahi = i128_split(a)
    blo, bhi = i128_split(b)
    lo, hi = C.c_uint64(), C.c_int64()
    lib().or_i128_add(alo, ahi, blo, bhi, C.byref(lo), C.byref(hi))
    return i128_join(lo.value, hi.value)


def i128_neg(a: int) -> int:
    lo, hi = C.c_uint64(), C.c_int64()
    alo, ahi = i128_split(a)
    lib().or_i128_neg(alo, ahi, C.byref(lo), C.byref(hi))
    return i128_join(lo.value, hi.value)


def i128_less(a: int, b: int) -> bool:
    alo, ahi = i128_split(a)
    blo, bhi = i128_split(b)
    return bool(lib().or_i128_less(alo, ahi, blo, bhi))


def i128_to_f64(v: int, scale: int) -> float:
    lo, hi = i128_split(v)
    return lib().or_i128_to_f64(lo, hi, scale)


def i128_from_f64(f: float, scale: int) -> int:
    lo, hi = C.c_uint64(), C.c_int64()
    lib().or_i128_from_f64(f, scale, C.byref(lo), C.byref(hi))
    return i128_join(lo.value, hi.value)


def i128_format(v: int, scale: int) -> str:
    lo, hi = i128_split(v)
    buf = C.create_string_buffer(128)
    lib().or_i128_format(lo, hi, scale, buf, 128)
    return buf.value.decode()


def parse_decimal(s: str, scale: int) -> int:
    return lib().or_parse_decimal(s.encode(), scale)


# --------------------------------------------------------------- operators

def hash_agg(cols, keys, aggs, sel=None, pred=None, max_groups=None, threads: int = 0):
    """GROUP BY keys over cols with aggs [(func, col)]; pred = (col, op, i64, f64).

    Returns dict(n, keys[n, k] int64, key_null[n, k] bool, acc structured[n, a]).
    threads > 0 uses the partial-aggregate + merge baseline (no sel, no nulls).
    """
    arr, keep = make_cols(cols)
    n_rows = len(cols[0][0]) if cols[0][1] != DECIMAL128 else len(cols[0][0]) // 2
    kc = np.asarray(keys, dtype=np.int32)
    af = np.asarray([a[0] for a in aggs], dtype=np.int32)
    ac = np.asarray([a[1] for a in aggs], dtype=np.int32)
    sel_a = None if sel is None else np.ascontiguousarray(sel, dtype=np.uint32)
    n_sel = 0 if sel_a is None else len(sel_a)
    mg = max_groups if max_groups is not None else max(n_sel if sel_a is not None else n_rows, 1)
    keys_out = np.empty((mg, len(keys)), dtype=np.int64)
    null_out = np.empty((mg, len(keys)), dtype=np.uint8)
    acc_out = np.zeros((mg, max(len(aggs), 1)), dtype=ACC_DTYPE)
    pc, po, pi, pf = (-1, 0, 0, 0.0) if pred is None else pred
    if threads:
        n = lib().or_hash_agg_parallel(arr, len(cols), n_rows, pc, po, pi, pf, _ptr(kc), len(keys), _ptr(af), _ptr(ac),
                                       len(aggs), threads, _ptr(keys_out), _ptr(acc_out), mg)
        null_out[:] = 0
    else:
        n = lib().or_hash_agg(arr, len(cols), n_rows, _ptr(sel_a), n_sel, pc, po, pi, pf, _ptr(kc), len(keys),
                              _ptr(af), _ptr(ac), len(aggs), _ptr(keys_out), _ptr(null_out), _ptr(acc_out), mg)
    if n < 0:
        raise RuntimeError("oracle hash_agg: group limit exceeded")
    return dict(n=n, keys=keys_out[:n], key_null=null_out[:n].astype(bool), acc=acc_out[:n])


def finalize(acc_col, func: int, in_type: int, scale: int = 0):
    """Finalized values for one aggregate column: (values, valid)."""
    n = len(acc_col)
    is_f = func in (SUM, AVG) or (func in (MIN, MAX) and in_type == FLOAT64)
    vals = np.empty(n, dtype=np.float64 if is_f else np.int64)
    valid = np.empty(n, dtype=bool)
    f, i = C.c_double(), C.c_int64()
    a = np.ascontiguousarray(acc_col)
    base = a.ctypes.data
    for g in range(n):
        ok = lib().or_finalize(C.c_void_p(base + g * ACC_DTYPE.itemsize), func, in_type, scale, C.byref(f),
                               C.byref(i))
        valid[g] = bool(ok)
        vals[g] = f.value if is_f else i.value
    return vals, valid


def filter_rows(col, n_rows, op, i64=0, f64=0.0, sel=None):
    arr, keep = make_cols([col])
    sel_a = None if sel is None else np.ascontiguousarray(sel, dtype=np.uint32)
    out = np.empty(max(n_rows if sel_a is None else len(sel_a), 1), dtype=np.uint32)
    k = lib().or_filter(arr, n_rows, _ptr(sel_a), 0 if sel_a is None else len(sel_a), op, i64, f64, _ptr(out))
    return out[:k]


def hash_join(build_key, probe_key, probe_sel=None, cap=None):
    """build_key/probe_key: (array, type[, validity]). Returns (probe_rows, build_rows)."""
    ba, k1 = make_cols([build_key])
    pa, k2 = make_cols([probe_key])
    nb = len(build_key[0])
    npr = len(probe_key[0])
    sel_a = None if probe_sel is None else np.ascontiguousarray(probe_sel, dtype=np.uint32)
    cap = cap if cap is not None else max(npr * 4, 16)
    op = np.empty(cap, dtype=np.uint32)
    ob = np.empty(cap, dtype=np.uint32)
    n = lib().or_hash_join(ba, nb, pa, npr, _ptr(sel_a), 0 if sel_a is None else len(sel_a), _ptr(op), _ptr(ob), cap)
    if n < 0:
        raise RuntimeError("oracle hash_join: output capacity exceeded")
    return op[:n], ob[:n]


def join_agg(build_key, build_group, build_dec, probe_key, probe_mul, threads: int = 1, max_groups: int = 1 << 16):
    """JOIN probe_key = build_key, then GROUP BY build_group SUM(probe_mul *
    build_dec) as exact Int128, COUNT(*). Columns: (array, type[, validity[,
    scale]]). Returns dict(n, keys int64 ascending, count, sum_lo uint64,
    sum_hi int64)."""
    ba, k1 = make_cols([build_key, build_group, build_dec])
    pa, k2 = make_cols([probe_key, probe_mul])
    nb, npr = len(build_key[0]), len(probe_key[0])
    keys = np.zeros(max_groups, np.int64)
    cnt = np.zeros(max_groups, np.int64)
    lo = np.zeros(max_groups, np.uint64)
    hi = np.zeros(max_groups, np.int64)
    sz = C.sizeof(OrCol)
    base_b, base_p = C.addressof(ba), C.addressof(pa)
    n = lib().or_join_agg_parallel(base_b, base_b + sz, base_b + 2 * sz, nb, base_p, base_p + sz, npr, threads,
                                   _ptr(keys), _ptr(cnt), _ptr(lo), _ptr(hi), max_groups)
    if n < 0:
        raise RuntimeError("oracle join_agg: group limit exceeded")
    return dict(n=n, keys=keys[:n], count=cnt[:n], sum_lo=lo[:n], sum_hi=hi[:n])


def semi_join(build_key, probe_key, probe_sel=None, anti=False):
    """build_key/probe_key: (array, type[, validity]). Probe positions, in
    input order, whose key is (semi) / is not (anti) among the non-NULL build
    keys; NULL probe keys never match."""
    ba, k1 = make_cols([build_key])
    pa, k2 = make_cols([probe_key])
    nb = len(build_key[0])
    npr = len(probe_key[0])
    sel_a = None if probe_sel is None else np.ascontiguousarray(probe_sel, dtype=np.uint32)
    out = np.empty(max(npr if sel_a is None else len(sel_a), 1), dtype=np.uint32)
    n = lib().or_semi_join(ba, nb, pa, npr, _ptr(sel_a), 0 if sel_a is None else len(sel_a), 1 if anti else 0,
                           _ptr(out))
    return out[:n]


def digest_rows(rows: np.ndarray):
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    n, w = rows.shape
    d = np.zeros(2, dtype=np.uint64)
    lib().or_digest_rows(_ptr(rows), n, w, _ptr(d))
    return int(d[0]), int(d[1])


def mix64(k: int) -> int:
    return lib().or_mix64(k & ((1 << 64) - 1))


# ------------------------------------------------ partial-state spill runs

# The external-merge spill run of HashAggregate (aggregate_partial_spill.go:19-52):
# header, groups in bytes.Compare order of their sort keys, end marker.
PARTIAL_MAGIC = b"WAGS\x01"
TAG_NULL, TAG_INT64, TAG_INT32 = 0, 3, 6  # partialTag* (:64-75)
PARQUET_INT32, PARQUET_INT64, PARQUET_IPV4, PARQUET_PORT = 1, 2, 8, 12  # parquet.TypeID (schema.go:12-35)


def partial_spec(func: int, in_type, scale: int = 0):
    """(func, is_float, is_decimal, dec_scale) of one aggregate, as
    initFlatAccums sets the flat accumulator flags (aggregate.go:6030-6119)."""
    if func == COUNT or in_type is None:
        return (func, False, False, 0)
    if func in (SUM, AVG):
        if in_type == FLOAT64:
            return (func, True, False, 0)
        if in_type in (DECIMAL, DECIMAL128):
            return (func, False, True, scale)
        if in_type == INT64 and func == AVG:  # overflow-safe AVG: float64 sum
            return (func, True, False, 0)
        return (func, False, False, 0)
    if in_type == FLOAT64:
        return (func, True, False, 0)
    if in_type in (DECIMAL, DECIMAL128):
        return (func, False, True, 0)
    return (func, False, False, 0)


def partial_sort_key(vals, nulls) -> bytes:
    """appendIntModeSortKey (aggregate_partial_drain_cursor.go:421-452): decimal
    text per column, 0-separated; a NULL value is "<null>" (sort.go:992-995)."""
    return b"\x00".join(b"<null>" if nl else str(int(v)).encode() for v, nl in zip(vals, nulls))


def partial_run(res, specs, key_types, names=None, type_ids=None) -> bytes:
    """One spill run of hash_agg's groups: writeHeader / Write / emitAcc /
    Close (aggregate_partial_spill.go:171-351). specs: partial_spec per agg."""
    import struct
    k = len(key_types)
    names = names or [""] * k
    if type_ids is None:
        type_ids = [PARQUET_INT32 if t == INT32 else PARQUET_PORT if t == UINT16 else
                    PARQUET_IPV4 if t in (IPV4, UINT32) else PARQUET_INT64 for t in key_types]
    out = bytearray(PARTIAL_MAGIC)
    out += struct.pack("<I", k)
    for nm, tid in zip(names, type_ids):
        b = nm.encode()
        out += struct.pack("<H", len(b)) + b + struct.pack("<H", tid)
    out += struct.pack("<I", len(specs))
    for func, is_f, is_d, sc in specs:
        out += struct.pack("<IBBi", func, int(is_f), int(is_d), sc)
    n = res["n"]
    out += struct.pack("<I", n)
    keys, nulls, acc = res["keys"], res["key_null"], res["acc"]
    order = sorted(range(n), key=lambda g: partial_sort_key(keys[g], nulls[g]))
    for g in order:
        sk = partial_sort_key(keys[g], nulls[g])
        out += b"\x01" + struct.pack("<I", len(sk)) + sk
        for c in range(k):
            if nulls[g][c]:
                out += bytes([TAG_NULL])
            elif key_types[c] in (INT32, UINT16):
                out += bytes([TAG_INT32]) + struct.pack("<i", int(keys[g][c]))
            else:
                out += bytes([TAG_INT64]) + struct.pack("<q", int(keys[g][c]))
        for a, (func, is_f, is_d, sc) in enumerate(specs):
            x = acc[g, a]
            if func == COUNT:
                out += struct.pack("<q", int(x["count"]))
            elif func in (SUM, AVG):
                out += struct.pack("<q", int(x["count"]))
                if is_d:
                    out += struct.pack("<qQ", int(x["sum_hi"]), int(x["sum_lo"]))
                elif is_f:
                    out += struct.pack("<d", float(x["sum_f64"]))
                else:
                    out += struct.pack("<q", int(x["sum_i64"]))
            else:
                has = bool(x["has_min"] if func == MIN else x["has_max"])
                out += bytes([1 if has else 0])
                if has:
                    if is_d:  # writeInt128: Hi, then Lo (aggregate_partial_spill.go:322-345, :702-707)
                        v = int(x["min_i64"] if func == MIN else x["max_i64"])
                        out += struct.pack("<qQ", -1 if v < 0 else 0, v & ((1 << 64) - 1))
                    elif is_f:
                        out += struct.pack("<d", float(x["min_f64"] if func == MIN else x["max_f64"]))
                    else:
                        out += struct.pack("<q", int(x["min_i64"] if func == MIN else x["max_i64"]))
    out += b"\x00"
    return bytes(out)


def read_partial_run(buf: bytes) -> dict:
    """readHeader / Next / readAcc (aggregate_partial_spill.go:759-877, :353-463):
    dict(cols=[(name, type_id)], specs=[(func, is_float, is_decimal, scale)],
    n, groups=[(sort_key, [key values, None = NULL], [acc dicts])])."""
    import struct
    p = 0

    def take(fmt):
        nonlocal p
        v = struct.unpack_from(fmt, buf, p)
        p += struct.calcsize(fmt)
        return v

    assert buf[:5] == PARTIAL_MAGIC, "bad magic"
    p = 5
    (k,) = take("<I")
    cols = []
    for _ in range(k):
        (ln,) = take("<H")
        nm = buf[p:p + ln].decode()
        p += ln
        (tid,) = take("<H")
        cols.append((nm, tid))
    (na,) = take("<I")
    specs = [tuple(take("<IBBi")) for _ in range(na)]
    specs = [(f, bool(i), bool(d), s) for f, i, d, s in specs]
    (n,) = take("<I")
    groups = []
    while True:
        (marker,) = take("<B")
        if marker == 0:
            break
        assert marker == 1, f"bad row marker {marker}"
        (sl,) = take("<I")
        sk = bytes(buf[p:p + sl])
        p += sl
        kv = []
        for _ in range(k):
            (tag,) = take("<B")
            kv.append(None if tag == TAG_NULL else take("<i")[0] if tag == TAG_INT32 else take("<q")[0])
        accs = []
        for func, is_f, is_d, sc in specs:
            a = {}
            if func == COUNT:
                (a["count"],) = take("<q")
            elif func in (SUM, AVG):
                (a["count"],) = take("<q")
                if is_d:
                    hi, lo = take("<qQ")
                    a["sum"] = (hi << 64) | lo  # the signed Int128
                elif is_f:
                    (a["sum"],) = take("<d")
                else:
                    (a["sum"],) = take("<q")
            else:
                (has,) = take("<B")
                a["has"] = bool(has)
                if has:
                    if is_d:  # readInt128 (aggregate_partial_spill.go:400-440, :709-717)
                        hi, lo = take("<qQ")
                        a["value"] = (hi << 64) | lo
                    else:
                        (a["value"],) = take("<d") if is_f else take("<q")
            accs.append(a)
        groups.append((sk, kv, accs))
    assert p == len(buf), "trailing bytes after the end marker"
    return dict(cols=cols, specs=specs, n=n, groups=groups)


# ------------------------------------------------------------ Bloom filter

BLOOM_MUL = 0x9E3779B97F4A7C15


def bloom_hash(keys: np.ndarray) -> np.ndarray:
    """bloomHashInt (join.go:1650-1654) over int64 keys, wrapping uint64."""
    with np.errstate(over="ignore"):
        x = keys.astype(np.int64).view(np.uint64) * np.uint64(BLOOM_MUL)
    return x ^ (x >> np.uint64(32))


def bloom_words_for(n_keys: int) -> int:
    """buildBloom's size (join.go:1452-1459): the smallest power of two of
    64-bit words holding 10 bits per key, at least 8."""
    w = 1
    while w * 64 < n_keys * 10:
        w *= 2
    return max(w, 8)


def bloom_build(build_keys: np.ndarray, valid=None) -> np.ndarray:
    """buildBloom + bloomSet over the distinct non-NULL build keys."""
    k = build_keys if valid is None else build_keys[valid]
    k = np.unique(k.astype(np.int64))
    words = np.zeros(bloom_words_for(len(k)), np.uint64)
    mask = np.uint64(len(words) - 1)
    h = bloom_hash(k)
    one = np.uint64(1)
    np.bitwise_or.at(words, (h & mask).astype(np.int64), one << (h & np.uint64(63)))
    np.bitwise_or.at(words, ((h >> np.uint64(17)) & mask).astype(np.int64), one << ((h >> np.uint64(6)) & np.uint64(63)))
    return words


def bloom_filter(words: np.ndarray, probe_keys: np.ndarray, valid=None, sel=None) -> np.ndarray:
    """BloomFilterOp.Execute's int-key path (bloom_filter_op.go:78-137):
    ascending positions (of sel, or of every row) whose non-NULL key passes
    bloomContains (:268-274)."""
    pos = np.arange(len(probe_keys), dtype=np.uint32) if sel is None else np.asarray(sel, np.uint32)
    keys = probe_keys.astype(np.int64)[pos]
    mask = np.uint64(len(words) - 1)
    h = bloom_hash(keys)
    one = np.uint64(1)
    b1 = (words[(h & mask).astype(np.int64)] >> (h & np.uint64(63))) & one
    b2 = (words[((h >> np.uint64(17)) & mask).astype(np.int64)] >> ((h >> np.uint64(6)) & np.uint64(63))) & one
    ok = (b1 == one) & (b2 == one)
    if valid is not None:
        ok &= np.asarray(valid, bool)[pos]
    return pos[ok]


# -------------------------------------------------------- shuffle routing

FNV_OFFSET64 = 14695981039346656037  # partitioned_shuffle_sink.go:1023
FNV_PRIME64 = 1099511628211          # partitioned_shuffle_sink.go:1024
_M64 = (1 << 64) - 1


def fnv1a64(data: bytes, h: int = FNV_OFFSET64) -> int:
    """FNV-1a 64 over a byte string (Go's hash/fnv New64a), one byte at a time."""
    for c in data:
        h = ((h ^ c) * FNV_PRIME64) & _M64
    return h


def shuffle_key_width(t: int) -> int:
    """Bytes a key column contributes to the row hash: the width of its storage
    in the reference's vectors (hashRowsIntoPartitions' type switch,
    partitioned_shuffle_sink.go:1057-1162). Int32/Port/Date hash their
    Int32Data (4 bytes); Int64/Timestamp/IPv4 their Int64Data (8 bytes, so a
    compact uint32 IPv4 hashes zero-extended to 8); Float64 its bits (8).
    Anything else (decimals here) takes the default case: 0."""
    return {INT32: 4, UINT16: 4, INT64: 8, IPV4: 8, UINT32: 8, FLOAT64: 8}.get(t, 0)


def _key_words(data: np.ndarray, t: int) -> np.ndarray:
    """Each row's storage bits as uint64 (little-endian byte order = bit order)."""
    if t in (INT32, UINT16):
        return data.astype(np.int64).astype(np.uint32).astype(np.uint64)
    if t == UINT32:
        return data.astype(np.uint32).astype(np.uint64)
    if t == FLOAT64:
        return np.ascontiguousarray(data, np.float64).view(np.uint64)
    return data.astype(np.int64).view(np.uint64)


def shuffle_hash(cols, key_cols, sel=None, n_rows=None) -> np.ndarray:
    """hashRowsIntoPartitions' per-row hash (partitioned_shuffle_sink.go:1039-1162)
    for every active row: FNV-1a 64 over the key columns' little-endian storage
    bytes, column after column; a NULL value mixes one 0xff byte; a column of
    another type mixes one 0x00 byte per row whether or not it is NULL.
    cols: list of (values, type[, validity words]); returns uint64 per active row."""
    n = len(cols[0][0]) if n_rows is None else n_rows
    rows = np.arange(n, dtype=np.int64) if sel is None else np.asarray(sel, np.int64)
    h = np.full(len(rows), FNV_OFFSET64, np.uint64)
    prime = np.uint64(FNV_PRIME64)
    with np.errstate(over="ignore"):
        for c in key_cols:
            data, t = cols[c][0], cols[c][1]
            valid = cols[c][2] if len(cols[c]) > 2 else None
            w = shuffle_key_width(t)
            if w == 0:
                h = (h ^ np.uint64(0)) * prime
                continue
            v = _key_words(np.asarray(data)[rows], t)
            hv = h.copy()
            for b in range(w):
                hv = (hv ^ ((v >> np.uint64(8 * b)) & np.uint64(0xFF))) * prime
            if valid is not None:
                ok = ((np.asarray(valid, np.uint64)[rows >> 6] >> (rows & 63).astype(np.uint64)) & np.uint64(1)) == 1
                hv = np.where(ok, hv, (h ^ np.uint64(0xFF)) * prime)
            h = hv
    return h


def shuffle_partition(cols, key_cols, num_parts: int, sel=None, n_rows=None):
    """Routing of partitionedShuffleSink.Consume (partitioned_shuffle_sink.go:
    266-290): parts[i] = row hash % num_parts for the i-th active row (all 0
    with no key columns, :1045-1050), then the rows of each partition in input
    order (perPartRows: b.Sel[i] or i). Returns (parts uint32, rows uint32
    grouped by partition, offsets int64 of num_parts + 1)."""
    n = len(cols[0][0]) if n_rows is None else n_rows
    active = np.arange(n, dtype=np.uint32) if sel is None else np.asarray(sel, np.uint32)
    mod = max(int(num_parts), 1)
    if len(key_cols) == 0:
        parts = np.zeros(len(active), np.uint32)
    else:
        parts = (shuffle_hash(cols, key_cols, sel, n) % np.uint64(mod)).astype(np.uint32)
    order = np.argsort(parts, kind="stable")
    counts = np.bincount(parts, minlength=mod).astype(np.int64)
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    return parts, active[order], offsets


def shuffle_row_bytes(vals, types, nulls=None) -> bytes:
    """hashVectorValue's byte stream for one row (partitioned_shuffle_sink.go:
    1177-1195): each key column's little-endian storage bytes in order, 0xff
    for a NULL; the reference's own routing test hashes rows this way."""
    out = bytearray()
    for i, (v, t) in enumerate(zip(vals, types)):
        if nulls is not None and nulls[i]:
            out += b"\xff"
            continue
        w = shuffle_key_width(t)
        if w == 0:
            out += b"\x00"
        elif t == FLOAT64:
            out += np.float64(v).tobytes()
        else:
            out += (int(v) & ((1 << (8 * w)) - 1)).to_bytes(w, "little")
    return bytes(out)


# ------------------------------------------------ two-column join keys

def _valid_at(valid, n):
    if valid is None:
        return np.ones(n, bool)
    v = np.asarray(valid, np.uint64)
    r = np.arange(n, dtype=np.int64)
    return ((v[r >> 6] >> (r & 63).astype(np.uint64)) & np.uint64(1)) == 1


def hash_join2(build_keys, probe_keys, build_valid=(None, None), probe_valid=(None, None), probe_sel=None):
    """Inner join on two int-class keys (dualIntKeyFromVectors + the chain walk
    that verifies both components, join.go:1621-1648, :2848-3130): every
    (probe row, build row) with both keys equal as int64; a NULL in either key
    never matches. Returns (probe_rows, build_rows) as uint32 arrays, probe
    rows ascending, build rows ascending within a probe row."""
    ba, bb = (np.asarray(k).astype(np.int64) for k in build_keys)
    pa, pb = (np.asarray(k).astype(np.int64) for k in probe_keys)
    bok = _valid_at(build_valid[0], len(ba)) & _valid_at(build_valid[1], len(ba))
    pok = _valid_at(probe_valid[0], len(pa)) & _valid_at(probe_valid[1], len(pa))
    index = {}
    for r in np.nonzero(bok)[0]:
        index.setdefault((int(ba[r]), int(bb[r])), []).append(int(r))
    rows = np.arange(len(pa)) if probe_sel is None else np.asarray(probe_sel, np.int64)
    out_p, out_b = [], []
    for r in rows:
        if not pok[r]:
            continue
        for b in index.get((int(pa[r]), int(pb[r])), ()):
            out_p.append(int(r))
            out_b.append(b)
    return np.array(out_p, np.uint32), np.array(out_b, np.uint32)


def semi_join2(build_keys, probe_keys, build_valid=(None, None), probe_valid=(None, None), probe_sel=None,
               anti=False):
    """Semi / anti join on two int-class keys: probe positions (ascending)
    whose pair has a build row (anti: has none; a NULL in either probe key never
    matches, so only the anti join keeps it)."""
    ba, bb = (np.asarray(k).astype(np.int64) for k in build_keys)
    pa, pb = (np.asarray(k).astype(np.int64) for k in probe_keys)
    bok = _valid_at(build_valid[0], len(ba)) & _valid_at(build_valid[1], len(ba))
    pok = _valid_at(probe_valid[0], len(pa)) & _valid_at(probe_valid[1], len(pa))
    keys = {(int(x), int(y)) for x, y, ok in zip(ba, bb, bok) if ok}
    rows = np.arange(len(pa)) if probe_sel is None else np.asarray(probe_sel, np.int64)
    hit = np.array([bool(pok[r]) and (int(pa[r]), int(pb[r])) in keys for r in rows], bool)
    keep = ~hit if anti else hit
    return rows[keep].astype(np.uint32)


# ------------------------------------- late-materialized (view) columns

def bits_of(valid: np.ndarray) -> np.ndarray:
    """Validity words (bit i % 64 of word i / 64, 1 = valid) of a bool array."""
    valid = np.asarray(valid, bool)
    b = np.packbits(valid, bitorder="little")
    b = np.concatenate([b, np.zeros((-len(b)) % 8, np.uint8)])
    return b.view(np.uint64).copy() if len(b) else np.zeros(0, np.uint64)


def view_flatten(base_vals, base_valid, indices, own_valid=None):
    """Vector.Flatten of a view (internal/engine/batch/view.go:87-230): row i
    is NULL when it is NULL in the view's own bitmap (the base is then not
    read, view.go:106-112) or the base row indices[i] is NULL; otherwise it is
    base[indices[i]]. NULL rows hold 0 (newOwnedLike storage). Returns
    (values, valid bool array)."""
    vals = np.asarray(base_vals)
    idx = np.asarray(indices, np.int64)
    n = len(idx)
    own = _valid_at(own_valid, n)
    bok = _valid_at(base_valid, len(vals))
    valid = np.zeros(n, bool)
    out = np.zeros(n, vals.dtype) if vals.ndim == 1 else np.zeros((n,) + vals.shape[1:], vals.dtype)
    for i in range(n):
        if not own[i]:
            continue
        s = int(idx[i])
        if bok[s]:
            valid[i] = True
            out[i] = vals[s]
    return out, valid


def view_compose(inner_indices, inner_own_valid, outer_indices):
    """A view over a view (NewViewVectorReuse, view.go:43-80): the composed
    indices are inner[outer[i]], and row i is own-NULL when the inner view's
    row outer[i] is own-NULL; the base stays the inner view's base. Returns
    (indices uint32, own valid bool array)."""
    inner = np.asarray(inner_indices, np.uint32)
    outer = np.asarray(outer_indices, np.int64)
    own = _valid_at(inner_own_valid, len(inner))
    return inner[outer].astype(np.uint32), own[outer]
