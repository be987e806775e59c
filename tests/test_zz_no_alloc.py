"""wg_filter_noalloc (the filter entry point that allocates nothing, its
scratch a static device array) against wg_filter through the C-ABI: two
calls in a row, the second on a smaller input, plus a Sel input, NULLs and
the size limit."""
import ctypes as C

import numpy as np
import pytest

from wadjet_amd import _lib as L
from wadjet_amd.gpu import Col, bitmap_from_mask, call, filter_sel

pytestmark = pytest.mark.gpu


def noalloc(ctx, col, n_rows, op, i64=0, f64=0.0, sel=None, n_sel=0, out=None):
    nout = C.c_uint64()
    cc = col.c()
    call("wg_filter_noalloc", ctx.h, C.byref(cc), n_rows, sel.ptr if sel is not None else None, n_sel, op, i64, f64,
         out.ptr, C.byref(nout))
    return nout.value


def test_noalloc_filter_equals_filter_two_calls(ctx):
    rng = np.random.default_rng(5)
    for n, t, op, ci, cf, use_sel in ((3_000_017, L.FLOAT64, L.GT, 0, -0.5, False),
                                      (70_001, L.INT32, L.NE, 7, 0.0, True)):
        print(f"[noalloc] n={n} type={t} op={op} sel={use_sel}", flush=True)
        vals = rng.standard_normal(n) if t == L.FLOAT64 else rng.integers(0, 20, n).astype(np.int32)
        valid = rng.random(n) > 0.05
        col = Col(ctx.to_device(vals), t, ctx.to_device(bitmap_from_mask(valid)))
        sel = np.sort(rng.choice(n, n // 2, replace=False)).astype(np.uint32) if use_sel else None
        dsel = ctx.to_device(sel) if use_sel else None
        k = len(sel) if use_sel else n
        want_buf, want_n = filter_sel(ctx, col, n, op, ci, cf, sel=dsel, n_sel=k if use_sel else 0)
        want = want_buf.to_host(np.uint32, want_n)
        out = ctx.alloc(k * 4)
        got_n = noalloc(ctx, col, n, op, ci, cf, sel=dsel, n_sel=k if use_sel else 0, out=out)
        assert got_n == want_n
        np.testing.assert_array_equal(out.to_host(np.uint32, got_n), want)


def test_noalloc_filter_rejects_oversized_input(ctx):
    col = Col(ctx.to_device(np.zeros(16, np.int64)), L.INT64)
    out = ctx.alloc(64)
    with pytest.raises(Exception) as e:
        noalloc(ctx, col, (1 << 26) + 1, L.EQ, 0, out=out)
    assert "at most" in str(e.value)
