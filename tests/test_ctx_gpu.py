"""Context-level entry points: the device-to-device copy kernel (also the
known-size kernel that calibrates HBM traffic counters) and its argument checks."""
import numpy as np
import pytest

from wadjet_amd import _lib as L

pytestmark = pytest.mark.gpu


def test_copy_d2d_exact(ctx):
    rng = np.random.default_rng(3)
    for n in (2, 1000, 1 << 20):
        x = rng.integers(-2**63, 2**63 - 1, n, dtype=np.int64)
        src = ctx.to_device(x)
        dst = ctx.alloc(n * 8)
        print(f"[ctx] copy {n} words", flush=True)
        ctx.copy(dst, src, n * 8)
        np.testing.assert_array_equal(dst.to_host(np.int64, n), x)


def test_copy_d2d_rejects_unaligned(ctx):
    src = ctx.alloc(64)
    dst = ctx.alloc(64)
    with pytest.raises(L.WgError):
        ctx.copy(dst, src, 24)
