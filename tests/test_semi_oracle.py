"""The oracle's semi / anti join probe (or_semi_join) pinned by the
reference's own test cases (restated with integer keys) and by a brute-force
membership loop with NULLs, duplicates and selection vectors."""
import numpy as np
import pytest

import oracle as O


def test_reference_cases():
    # join_coverage_test.go:272-312 TestSemiJoin: probe a, b, c; build a, b -> 2 rows
    build = np.array([10, 20], np.int64)
    probe = np.array([10, 20, 30], np.int64)
    assert O.semi_join((build, O.INT64), (probe, O.INT64)).tolist() == [0, 1]
    # join_coverage_test.go:314-351 TestAntiJoin: only c (position 2)
    assert O.semi_join((build, O.INT64), (probe, O.INT64), anti=True).tolist() == [2]
    # join_coverage_test.go:711-735 TestAntiJoinNoMatches: empty build, all rows pass
    empty = np.zeros(0, np.int64)
    assert O.semi_join((empty, O.INT64), (probe[:2], O.INT64), anti=True).tolist() == [0, 1]
    assert O.semi_join((empty, O.INT64), (probe[:2], O.INT64)).tolist() == []


def brute(bk, bvalid, pk, pvalid, sel, anti):
    keys = {int(k) for k, v in zip(bk, bvalid) if v}
    rows = range(len(pk)) if sel is None else sel
    out = []
    for r in rows:
        found = bool(pvalid[r]) and int(pk[r]) in keys
        if found != anti:
            out.append(int(r))
    return out


def words(mask):
    bits = np.packbits(mask, bitorder="little")
    bits = np.concatenate([bits, np.zeros((-len(bits)) % 8, np.uint8)])
    return bits.view(np.uint64)


@pytest.mark.parametrize("anti", [False, True])
@pytest.mark.parametrize("typ", ["i64", "i32"])
def test_brute_force_with_nulls_and_sel(anti, typ):
    rng = np.random.default_rng(7 if anti else 8)
    dt, t = (np.int64, O.INT64) if typ == "i64" else (np.int32, O.INT32)
    bk = rng.integers(-50, 50, 400).astype(dt)
    pk = rng.integers(-80, 80, 3000).astype(dt)
    bvalid = rng.random(400) > 0.1
    pvalid = rng.random(3000) > 0.1
    sel = np.sort(rng.choice(3000, 1700, replace=False)).astype(np.uint32)
    for s in (None, sel):
        got = O.semi_join((bk, t, words(bvalid)), (pk, t, words(pvalid)), probe_sel=s, anti=anti)
        assert got.tolist() == brute(bk, bvalid, pk, pvalid, None if s is None else s.tolist(), anti)


def test_planted_mistake_is_caught():
    """An anti join that drops NULL probe keys must differ from the oracle."""
    rng = np.random.default_rng(3)
    bk = rng.integers(0, 10, 50).astype(np.int64)
    pk = rng.integers(0, 20, 500).astype(np.int64)
    pvalid = rng.random(500) > 0.2
    got = O.semi_join((bk, O.INT64), (pk, O.INT64, words(pvalid)), anti=True).tolist()
    wrong = [r for r in brute(bk, np.ones(50, bool), pk, pvalid, None, True) if pvalid[r]]
    assert got != wrong
