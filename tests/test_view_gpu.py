"""Late-materialized (view) columns on the device: wg_view_flatten against
the oracle's Vector.Flatten restatement for every storage width, with own and
base NULLs, a view over a view composed through the same call, and the join
probe's pair arrays used as views over the probe batch and the retained build
columns, equal to the oracle's eager join output."""
import numpy as np
import pytest

import oracle as O
from wadjet_amd import _lib as L
from wadjet_amd.gpu import Col, HashJoin, view_flatten

pytestmark = pytest.mark.gpu


def valid_bits(words, n):
    r = np.arange(n)
    return ((np.asarray(words, np.uint64)[r >> 6] >> (r & 63).astype(np.uint64)) & np.uint64(1)) == 1


CASES = [(L.INT32, np.int32), (L.UINT16, np.uint16), (L.UINT32, np.uint32), (L.INT64, np.int64),
         (L.FLOAT64, np.float64), (L.DECIMAL, np.int64)]


@pytest.mark.parametrize("wt,dt", CASES)
@pytest.mark.parametrize("own_nulls", [False, True])
def test_flatten_matches_oracle(ctx, wt, dt, own_nulls):
    rng = np.random.default_rng(wt * 7 + own_nulls)
    nb, n = 5000, 70_001
    base = rng.integers(0, np.iinfo(np.int32).max, nb).astype(dt) if dt != np.float64 else rng.normal(size=nb)
    bok = rng.random(nb) < 0.8
    idx = rng.integers(0, nb, n).astype(np.uint32)
    own = (rng.random(n) < 0.9) if own_nulls else None
    bw = O.bits_of(bok)
    ow = None if own is None else O.bits_of(own)
    got = view_flatten(ctx, Col(ctx.to_device(base), wt, ctx.to_device(bw)), nb, ctx.to_device(idx), n,
                       own_valid=None if ow is None else ctx.to_device(ow))
    want, wv = O.view_flatten(base, bw, idx, ow)
    gv = valid_bits(got.validity.to_host(np.uint64, (n + 63) // 64), n)
    np.testing.assert_array_equal(gv, wv)
    np.testing.assert_array_equal(got.data.to_host(dt, n).view(f"u{base.dtype.itemsize}"),
                                  want.view(f"u{base.dtype.itemsize}"))
    # bits past the last row are clear
    words = got.validity.to_host(np.uint64, (n + 63) // 64)
    assert int(words[-1]) >> (n % 64) == 0


def test_decimal128_and_no_validity(ctx):
    rng = np.random.default_rng(9)
    nb, n = 300, 1000
    pairs = rng.integers(-2**62, 2**62, (nb, 2)).astype(np.int64)
    idx = rng.integers(0, nb, n).astype(np.uint32)
    got = view_flatten(ctx, Col(ctx.to_device(pairs.ravel()), L.DECIMAL128), nb, ctx.to_device(idx), n)
    np.testing.assert_array_equal(got.data.to_host(np.int64, 2 * n).reshape(n, 2), pairs[idx])
    assert valid_bits(got.validity.to_host(np.uint64, (n + 63) // 64), n).all()


def test_view_over_view_composes(ctx):
    """TestViewComposition (view_test.go:171-183) on the device: compose by
    flattening the inner indices (a UINT32 column carrying the inner view's
    own NULLs), then flatten the base through the composed view."""
    base = np.array([0, 10, 20, 30], np.int64)
    inner = np.array([3, 2, 1, 0], np.uint32)
    inner_own = O.bits_of(np.array([1, 1, 0, 1], bool))
    outer = np.array([2, 0, 1, 2], np.uint32)
    comp = view_flatten(ctx, Col(ctx.to_device(inner), L.UINT32, ctx.to_device(inner_own)), 4, ctx.to_device(outer), 4)
    flat = view_flatten(ctx, Col(ctx.to_device(base), L.INT64), 4, comp.data, 4, own_valid=comp.validity)
    v = valid_bits(flat.validity.to_host(np.uint64, 1), 4)
    vals = flat.data.to_host(np.int64, 4)
    assert [int(x) if ok else None for x, ok in zip(vals, v)] == [None, 30, 20, None]


def test_index_past_the_base_is_an_error(ctx):
    base = np.arange(10, dtype=np.int64)
    idx = np.array([0, 9, 10, 3], np.uint32)  # 10 is past the base; the base is not read there
    with pytest.raises(L.WgError, match="past the base"):
        view_flatten(ctx, Col(ctx.to_device(base), L.INT64), 10, ctx.to_device(idx), 4)


def test_join_pairs_as_views(ctx):
    """The probe's pair arrays as views: the probe columns and the retained
    build columns flattened through them equal the oracle's eager join output
    (emitViewOutput vs the gathered output, join.go:2436-2600)."""
    rng = np.random.default_rng(4)
    nb, npr = 20_000, 150_000
    bkey = rng.permutation(40_000)[:nb].astype(np.int64)
    bpay = rng.integers(-1000, 1000, nb).astype(np.int32)
    bpay_ok = rng.random(nb) < 0.9
    pkey = np.concatenate([bkey[rng.integers(0, nb, npr - 5000)], rng.integers(50_000, 60_000, 5000)]).astype(np.int64)
    pval = rng.normal(size=npr)
    pval_ok = rng.random(npr) < 0.95
    j = HashJoin(ctx, [L.INT64, L.INT32], 0)
    j.build([Col(ctx.to_device(bkey), L.INT64), Col(ctx.to_device(bpay), L.INT32, ctx.to_device(O.bits_of(bpay_ok)))],
            nb)
    j.finish()
    cap = npr + 16
    outp, outb = ctx.alloc(cap * 4), ctx.alloc(cap * 4)
    got, cur = j.probe(Col(ctx.to_device(pkey), L.INT64), npr, outp, outb, cap)
    assert cur.done == 1
    pv = view_flatten(ctx, Col(ctx.to_device(pval), L.FLOAT64, ctx.to_device(O.bits_of(pval_ok))), npr, outp, got)
    bv = view_flatten(ctx, j.build_column(1), nb, outb, got)
    pr, br = outp.to_host(np.uint32, got), outb.to_host(np.uint32, got)
    want_p, want_b = O.hash_join((bkey, O.INT64), (pkey, O.INT64))
    assert got == len(want_p)
    order = np.lexsort((br, pr))
    worder = np.lexsort((want_b, want_p))
    np.testing.assert_array_equal(pr[order], want_p[worder])
    np.testing.assert_array_equal(br[order], want_b[worder])
    # the flattened columns, per pair, equal the eager gather of the oracle's pairs
    ev, evok = O.view_flatten(pval, O.bits_of(pval_ok), want_p[worder])
    eb, ebok = O.view_flatten(bpay, O.bits_of(bpay_ok), want_b[worder])
    gvv = valid_bits(pv.validity.to_host(np.uint64, (got + 63) // 64), got)[order]
    gbv = valid_bits(bv.validity.to_host(np.uint64, (got + 63) // 64), got)[order]
    np.testing.assert_array_equal(gvv, evok)
    np.testing.assert_array_equal(gbv, ebok)
    np.testing.assert_array_equal(pv.data.to_host(np.float64, got)[order].view(np.uint64), ev.view(np.uint64))
    np.testing.assert_array_equal(bv.data.to_host(np.int32, got)[order], eb)
