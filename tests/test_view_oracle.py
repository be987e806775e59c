"""The oracle's view flatten / compose (late materialization) pinned by the
reference's own view tests (internal/engine/batch/view_test.go) and by
flatten(compose) == flatten(flatten) on random views."""
import numpy as np

import oracle as O

IDX = [3, 0, 0, 2, 1, 3]  # permutation, duplication and repeats of a NULL base row (view_test.go:83-85)


def flat(vals, valid_bool, idx, own=None):
    bv = O.bits_of(np.asarray(valid_bool, bool))
    ov = None if own is None else O.bits_of(np.asarray(own, bool))
    return O.view_flatten(np.asarray(vals), bv, idx, ov)


def as_boxed(vals, valid):
    return [v.item() if ok else None for v, ok in zip(vals, valid)]


def test_flatten_flat_types_kat():
    """TestViewFlattenAllFlatTypes (view_test.go:53-93), the int/float/decimal cases."""
    cases = [
        (np.int32, [1, -7, 0, 2048], [1, 1, 0, 1], [2048, 1, 1, None, -7, 2048]),
        (np.int64, [1, 0, -9 * 10**15, 42], [1, 0, 1, 1], [42, 1, 1, -9 * 10**15, None, 42]),
        (np.float64, [1.5, 0.0, -2.75, 0.0], [1, 0, 1, 1], [0.0, 1.5, 1.5, -2.75, None, 0.0]),
        (np.int32, [443, 0, 0, 65535], [1, 0, 1, 1], [65535, 443, 443, 0, None, 65535]),  # port
        (np.int64, [12345, 0, -1, 0], [1, 0, 1, 1], [0, 12345, 12345, -1, None, 0]),  # DECIMAL(.,2)
    ]
    for dt, vals, ok, want in cases:
        out, valid = flat(np.array(vals, dt), ok, IDX)
        assert as_boxed(out, valid) == want
        assert not np.any(out[~valid]), "NULL rows hold 0"


def test_own_null_override_kat():
    """TestViewOwnNullOverride (view_test.go:155-169): own-NULL rows are NULL
    whatever their index."""
    out, valid = flat(np.array([10, 20], np.int64), [1, 1], [1, 0], own=[0, 1])
    assert as_boxed(out, valid) == [None, 10]


def test_composition_kat():
    """TestViewComposition (view_test.go:171-183): base [0,10,20,30], v1 =
    [3,2,1,0] with row 2 own-NULL, v2 over v1 = [2,0,1,2] -> [nil,30,20,nil]."""
    base = np.array([0, 10, 20, 30], np.int64)
    own1 = np.array([1, 1, 0, 1], bool)
    idx2, own2 = O.view_compose([3, 2, 1, 0], O.bits_of(own1), [2, 0, 1, 2])
    assert idx2.tolist() == [1, 3, 2, 1] and own2.tolist() == [False, True, True, False]
    out, valid = flat(base, [1, 1, 1, 1], idx2, own=own2)
    assert as_boxed(out, valid) == [None, 30, 20, None]


def test_nan_empty_single_kat():
    """TestViewFloat64NaN, TestViewEmptyAndSingle (view_test.go:185-219)."""
    out, valid = flat(np.array([np.nan, 1.0]), [1, 1], [0, 1, 0])
    assert np.isnan(out[0]) and out[1] == 1.0 and np.isnan(out[2]) and valid.all()
    out, valid = flat(np.array([5], np.int64), [1], [])
    assert len(out) == 0 and len(valid) == 0
    out, valid = flat(np.array([5], np.int64), [1], [0])
    assert as_boxed(out, valid) == [5]


def test_compose_equals_flatten_twice():
    """A view over a view flattens to what flattening the inner view first and
    then the outer one gives, NULLs included; both agree with a per-row
    restatement of the rule."""
    rng = np.random.default_rng(3)
    for _ in range(20):
        nb, n1, n2 = rng.integers(1, 60), rng.integers(1, 80), rng.integers(0, 90)
        base = rng.integers(-10**12, 10**12, nb)
        bok = rng.random(nb) < 0.8
        i1 = rng.integers(0, nb, n1).astype(np.uint32)
        own1 = rng.random(n1) < 0.85
        i2 = rng.integers(0, n1, n2).astype(np.uint32)
        own2 = rng.random(n2) < 0.9
        # compose, then flatten once
        ci, cown = O.view_compose(i1, O.bits_of(own1), i2)
        cown &= own2
        a, av = flat(base, bok, ci, own=cown)
        # flatten the inner view, then the outer one over it
        f1, f1v = flat(base, bok, i1, own=own1)
        b, bv = flat(f1, f1v, i2, own=own2)
        assert av.tolist() == bv.tolist() and a.tolist() == b.tolist()
        want = [int(base[i1[j]]) if own2[k] and own1[j] and bok[i1[j]] else None
                for k, j in enumerate(i2.tolist())]
        assert as_boxed(a, av) == want
