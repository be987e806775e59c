"""Pins of the oracle's two-column-key joins (oracle.hash_join2 / semi_join2,
restating dualIntKeyFromVectors and the verifying chain walk,
internal/engine/exec/join.go:1621-1648, :2848-3130)."""
import numpy as np

import oracle as O


def test_reference_inner_case():
    """TestHashJoin_DualIntKey (dual_int_test.go:11-75): 3 build pairs, probes
    (1,10), (2,20), (99,99): two matches with equal keys."""
    p, b = O.hash_join2(([1, 2, 3], [10, 20, 30]), ([1, 2, 99], [10, 20, 99]))
    assert list(zip(p, b)) == [(0, 0), (1, 1)]


def test_reference_semi_anti_case():
    """TestHashJoin_DualIntKey_SemiAnti (dual_int_test.go:77-140): semi keeps
    2 of 3 probe rows, anti the other 1."""
    build, probe = ([1, 2], [10, 20]), ([1, 3, 2], [10, 30, 20])
    assert list(O.semi_join2(build, probe)) == [0, 2]
    assert list(O.semi_join2(build, probe, anti=True)) == [1]


def test_against_a_nested_loop_with_duplicates_and_nulls():
    """Brute force over every (probe, build) pair on small inputs: duplicate
    build pairs (multi-match), one key equal but not the other, NULLs in either
    key on both sides, int32 against int64 (equal as int64)."""
    rng = np.random.default_rng(3)
    nb, npr = 300, 400
    ba = rng.integers(-3, 4, nb).astype(np.int64)
    bb = rng.integers(0, 5, nb).astype(np.int32)
    pa = rng.integers(-3, 4, npr).astype(np.int32)
    pb = rng.integers(0, 5, npr).astype(np.int64)
    bm = (rng.random(nb) > 0.1, rng.random(nb) > 0.1)
    pm = (rng.random(npr) > 0.1, rng.random(npr) > 0.1)
    words = lambda m: np.packbits(np.concatenate([m, np.zeros((-len(m)) % 64, bool)]), bitorder="little").view(np.uint64)
    p, b = O.hash_join2((ba, bb), (pa, pb), (words(bm[0]), words(bm[1])), (words(pm[0]), words(pm[1])))
    want = [(i, j) for i in range(npr) for j in range(nb)
            if pm[0][i] and pm[1][i] and bm[0][j] and bm[1][j] and pa[i] == ba[j] and pb[i] == bb[j]]
    assert sorted(zip(p.tolist(), b.tolist())) == want
    semi = O.semi_join2((ba, bb), (pa, pb), (words(bm[0]), words(bm[1])), (words(pm[0]), words(pm[1])))
    assert semi.tolist() == sorted({i for i, _ in want})
    anti = O.semi_join2((ba, bb), (pa, pb), (words(bm[0]), words(bm[1])), (words(pm[0]), words(pm[1])), anti=True)
    assert anti.tolist() == [i for i in range(npr) if i not in {i for i, _ in want}]
