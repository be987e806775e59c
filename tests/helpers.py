"""Shared comparison helpers: GPU results (wadjet_amd.gpu.HashAggregate) vs
the CPU oracle (oracle.hash_agg), sorted by key."""
from __future__ import annotations

import numpy as np

import oracle as O

# float64 SUM/AVG: the GPU adds in a different order than the reference's
# row-order loop. Tolerance: |gpu - ref| <= FLOAT_RTOL * sum(|v|) of the group
# (equal to the plain relative error when the values share a sign).
FLOAT_RTOL = 1e-12


def _order(keys_2d: np.ndarray, nulls_2d: np.ndarray) -> np.ndarray:
    cols = []
    for k in range(keys_2d.shape[1] - 1, -1, -1):
        cols.append(keys_2d[:, k])
        cols.append(nulls_2d[:, k].astype(np.int64))
    return np.lexsort(cols)


def gpu_groups(agg) -> dict:
    r = agg.fetch()
    n = r["n"]
    keys = np.stack([k.astype(np.int64) for k in r["keys"]], axis=1) if n else np.zeros((0, len(r["keys"])), np.int64)
    nulls = np.repeat((~r["key_valid"])[:, None], keys.shape[1], axis=1)
    o = _order(keys, nulls)
    states = [agg.fetch_state(a) for a in range(len(agg.aggs))]
    return dict(n=n, keys=keys[o], nulls=nulls[o], values=[v[o] for v in r["values"]],
                valid=[v[o] for v in r["valid"]], state=[{k: v[o] for k, v in s.items()} for s in states])


def oracle_groups(res: dict) -> dict:
    o = _order(res["keys"], res["key_null"])
    return dict(n=res["n"], keys=res["keys"][o], nulls=res["key_null"][o], acc=res["acc"][o])


def compare(gpu: dict, ref: dict, aggs, types, scales=None, abs_sums=None):
    """aggs: [(func, col)], types: column types. abs_sums[a] = per-group sum(|v|)
    (oracle order) for float SUM/AVG aggregates."""
    scales = scales or [0] * len(types)
    assert gpu["n"] == ref["n"], f"group count {gpu['n']} vs {ref['n']}"
    np.testing.assert_array_equal(gpu["keys"], ref["keys"])
    np.testing.assert_array_equal(gpu["nulls"], ref["nulls"])
    for a, (func, col) in enumerate(aggs):
        acc = ref["acc"][:, a]
        t = types[col] if col >= 0 else O.INT64
        sc = scales[col] if col >= 0 else 0
        want, want_ok = O.finalize(acc, func, t, sc)
        got, got_ok = gpu["values"][a], gpu["valid"][a]
        np.testing.assert_array_equal(got_ok, want_ok, err_msg=f"agg {a}: NULL pattern")
        st = gpu["state"][a]
        if func == O.COUNT:
            np.testing.assert_array_equal(got, want)
            continue
        if func in (O.MIN, O.MAX):
            m = want_ok
            if t == O.FLOAT64:
                np.testing.assert_array_equal(got[m].view(np.uint64), want[m].view(np.uint64),
                                              err_msg=f"agg {a}: float MIN/MAX")
            else:
                np.testing.assert_array_equal(got[m], want[m], err_msg=f"agg {a}: int MIN/MAX")
            continue
        # SUM / AVG
        m = want_ok
        if acc["is_decimal"].any():
            lo = acc["sum_lo"]
            hi = acc["sum_hi"]
            np.testing.assert_array_equal(st["sum_lo"][m], lo[m], err_msg=f"agg {a}: Int128 lo")
            np.testing.assert_array_equal(st["sum_hi"][m], hi[m], err_msg=f"agg {a}: Int128 hi")
            np.testing.assert_array_equal(got[m].view(np.uint64), want[m].view(np.uint64),
                                          err_msg=f"agg {a}: decimal finalization")
        elif acc["is_float"].any():
            scale = abs_sums[a] if abs_sums is not None and abs_sums[a] is not None else np.abs(acc["sum_f64"])
            if func == O.AVG:
                scale = scale / np.maximum(acc["count"], 1)
            err = np.abs(got[m] - want[m])
            lim = FLOAT_RTOL * np.maximum(scale[m], 1e-300)
            bad = np.nonzero(err > lim)[0]
            assert len(bad) == 0, f"agg {a}: {len(bad)} float sums out of tolerance, e.g. {got[m][bad[:3]]} vs {want[m][bad[:3]]}"
        else:
            np.testing.assert_array_equal(st["sum_lo"][m].view(np.int64), acc["sum_i64"][m], err_msg=f"agg {a}: int sum")
            np.testing.assert_array_equal(got[m].view(np.uint64), want[m].view(np.uint64), err_msg=f"agg {a}: int finalization")


def splitmix_at(seed: int, rows: np.ndarray) -> np.ndarray:
    """SplitMix64 output at positions `rows` (uint64 arithmetic mod 2^64): the
    stream every seeded device generator draws from (wadjet_gen.h)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (rows.astype(np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def fk_match_count(n: int, seed: int, permille: int, chunk: int = 1 << 24) -> int:
    """Probe rows of the foreign-key generator that draw an existing build key
    (closed form): row r matches iff splitmix(seed ^ 0x3C..3C, r) % 1000 <
    permille; every other row gets a key outside the build's domain."""
    s = (seed ^ 0x3C3C3C3C3C3C3C3C) & ((1 << 64) - 1)
    total = 0
    for lo in range(0, n, chunk):
        r = splitmix_at(s, np.arange(lo, min(n, lo + chunk), dtype=np.uint64))
        total += int(np.count_nonzero(r % np.uint64(1000) < np.uint64(permille)))
    return total
