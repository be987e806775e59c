"""Pins of the oracle's Bloom pre-filter (join.go:1441-1483, :1650-1654;
bloom_filter_op.go:39-274): the size rule, no false negatives, the false
positive rate two hash bits at 10-20 bits per key allow, order and NULLs."""
import numpy as np

import oracle as O


def test_size_rule():
    """nSlots = smallest power of two with nSlots * 64 >= 10 * keys, at least
    8: 1 key -> 8 words, 51 -> 8, 52 -> 16, 10M -> 2^21 (16 MiB)."""
    assert [O.bloom_words_for(n) for n in (0, 1, 51, 52, 103, 10_000_000)] == [8, 8, 8, 16, 32, 1 << 21]


def test_hash_is_the_golden_ratio_multiply_then_xorshift():
    """bloomHashInt(1) is the multiplier folded by its own high half; 0 maps to 0."""
    m = O.BLOOM_MUL
    assert int(O.bloom_hash(np.array([1]))[0]) == m ^ (m >> 32)
    assert int(O.bloom_hash(np.array([0]))[0]) == 0
    assert int(O.bloom_hash(np.array([-1]))[0]) == ((-m) % 2**64) ^ (((-m) % 2**64) >> 32)


def test_no_false_negatives_and_bounded_false_positives():
    rng = np.random.default_rng(5)
    build = rng.choice(10**9, 50_000, replace=False).astype(np.int64)
    words = O.bloom_build(build)
    assert len(words) == O.bloom_words_for(50_000)
    # every build key passes
    assert len(O.bloom_filter(words, build)) == len(build)
    # keys drawn outside the build pass at a few to ~10 %: not the ~1 % of two
    # independent bits, because the first bit's position is the low 6 bits of
    # its own word index (bloomSet takes both from the same hash bits), so
    # the first probe alone almost always hits. The reference's filter has
    # the same rate; what is pinned here is that it filters at all and never
    # drops a build key.
    others = rng.integers(10**9, 2 * 10**9, 200_000).astype(np.int64)
    rate = len(O.bloom_filter(words, others)) / len(others)
    assert 0.001 < rate < 0.25, rate
    # first-probe structure: for every word w the first probe can only set bit w % 64
    h = O.bloom_hash(build)
    mask = np.uint64(len(words) - 1)
    assert np.all((h & np.uint64(63)) == ((h & mask) & np.uint64(63)))
    # at most two bits per distinct key are set
    assert sum(bin(int(w)).count("1") for w in words) <= 2 * 50_000


def test_order_sel_and_nulls():
    """Positions come back ascending; a selection vector restricts them; a NULL
    probe key never passes even when its value would."""
    build = np.arange(0, 1000, 3, dtype=np.int64)
    words = O.bloom_build(build)
    probe = np.arange(0, 2000, dtype=np.int64)
    valid = np.ones(2000, bool)
    valid[::7] = False
    got = O.bloom_filter(words, probe, valid)
    assert np.all(np.diff(got.astype(np.int64)) > 0)
    assert set(np.arange(0, 1000, 3)[(np.arange(0, 1000, 3) % 7) != 0]).issubset(set(got.tolist()))
    assert not any(p % 7 == 0 for p in got)
    sel = np.arange(1, 2000, 2, dtype=np.uint32)
    got2 = O.bloom_filter(words, probe, valid, sel)
    assert set(got2.tolist()) == set(got.tolist()) & set(sel.tolist())
