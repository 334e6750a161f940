"""The indexed subset kernel (two stream sets per wave, 128-bit LDS rows),
forced on for every launch, against the oracle and against the bit-walking
kernel."""
import numpy as np
import pytest

import oracle_lib as orc
import synchrowords_amd as sw
from conftest import golden
from helpers import random_sets, words
from synchrowords_amd import gen, replay

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ixctx():
    c = sw.Context(0)
    c.set_subset_kernel(True, 1)  # every launch takes the indexed kernel where it fits
    yield c
    c.close()


@pytest.mark.parametrize("n", [1, 5, 63, 64, 65, 100, 128, 200, 256, 257, 300, 318, 319, 448])
def test_indexed_counts_rows_and_any(ixctx, n):
    rng = np.random.default_rng(500 + n)
    W = words(n)
    small = random_sets(rng, 901, n, 0.08 if n > 20 else 0.3)  # odd: a wave's last pair is half empty
    large = random_sets(rng, 5000, n, 0.75)
    for i in range(0, 901, 37):
        small[i] = small[i] & large[int(rng.integers(0, 5000))]
    small[3] = large[4999]
    small[4] = 0  # the empty set is in every set
    total, cnt, _ = orc.subset_count(small, large, W)
    res, c, anyf = ixctx.subset_check(n, small, large, thresh=1, per_row=True)
    assert c == total and anyf == (total > 0)
    np.testing.assert_array_equal(res, cnt >= 1)
    res2, _, _ = ixctx.subset_check(n, small, large, thresh=3, per_row=True)
    np.testing.assert_array_equal(res2, cnt >= 3)
    _, c3, _ = ixctx.subset_check(n, small, large, count=True)
    assert c3 == total
    _, _, any4 = ixctx.subset_check(n, small, large, count=False)
    assert any4 == (total > 0)


@pytest.mark.parametrize("shape", [(1, 1), (2, 4097), (3, 4096), (4097, 3), (7000, 9000)])
def test_indexed_shapes(ixctx, shape):
    n = 150
    rng = np.random.default_rng(sum(shape))
    small = random_sets(rng, shape[0], n, 0.06)
    large = random_sets(rng, shape[1], n, 0.8)
    total, _, _ = orc.subset_count(small, large, words(n))
    _, c, anyf = ixctx.subset_check(n, small, large, count=True)
    assert c == total and anyf == (total > 0)


def test_indexed_long_stream_sets(ixctx):
    """Stream sets of more than the 32 states kept in registers."""
    n = 256
    rng = np.random.default_rng(9)
    small = random_sets(rng, 600, n, 0.3)   # about 77 states each
    large = random_sets(rng, 3000, n, 0.97)
    for i in range(0, 600, 10):
        small[i] &= large[i]
    total, cnt, _ = orc.subset_count(small, large, words(n))
    res, c, _ = ixctx.subset_check(n, small, large, per_row=True)
    assert c == total
    np.testing.assert_array_equal(res, cnt >= 1)


@pytest.mark.parametrize("case", [c for c in golden("c1_random48.json")[:20]] +
                         [c for c in golden("c3_random200.json")[:4]], ids=lambda c: c["name"])
def test_exact_search_with_indexed_kernel(ixctx, case):
    """The whole search with every goal check and dominance pass indexed."""
    n, k = case["n"], case["k"]
    if case["non_synchro"]:
        pytest.skip("non-synchronizing")
    r = ixctx.exact(n, k, case["delta"])
    assert r.found and r.length == case["upper"]
    assert replay(n, k, case["delta"], r.word) == 1


@pytest.mark.parametrize("n,k", [(90, 3), (300, 2)])
def test_indexed_and_bitwalk_searches_agree(ixctx, n, k):
    """Same lengths and the same pruned sets (dominance flags are exact)."""
    bw = sw.Context(0)
    bw.set_subset_kernel(False)
    try:
        cfg = sw.ExactConfig(dominance_pairs=10**11)  # every pass up to 1e11 pairs
        for seed in range(2):
            d = gen.random_dfa(n, k, 8100 + seed)
            a = ixctx.exact(n, k, d, cfg=cfg)
            b = bw.exact(n, k, d, cfg=cfg)
            assert a.found and b.found and a.length == b.length
            assert a.stats["sets_forward"] == b.stats["sets_forward"]
            assert a.stats["sets_backward"] == b.stats["sets_backward"]
    finally:
        bw.close()


@pytest.mark.parametrize("walk", [1, 2, 3, 4])
@pytest.mark.parametrize("n", [7, 64, 130, 256, 300])
def test_walk_orders_give_the_same_results(ixctx, walk, n):
    """Set-order walk and rarest-first walk with early exit, on lists where
    most stream sets lie in no tile set (early exit taken) and on lists with
    many pairs (early exit not taken): counts, per-row counts and the flag
    equal the oracle's."""
    rng = np.random.default_rng(700 + n + walk)

    def mixed(count, ps):
        """Sets of widely different sizes, shuffled (sorting by size matters)."""
        parts = [random_sets(rng, count // len(ps) + 1, n, q) for q in ps]
        out = np.concatenate(parts)[:count]
        return out[rng.permutation(count)]

    ixctx.set_subset_walk(walk)
    try:
        for p_small, p_large in ((0.3, 0.3), (0.08, 0.75), (0.02, 0.95), ((0.01, 0.1, 0.4), (0.05, 0.3, 0.9))):
            if isinstance(p_small, tuple):
                small, large = mixed(1201, p_small), mixed(4500, p_large)
            else:
                small = random_sets(rng, 1201, n, p_small)
                large = random_sets(rng, 4500, n, p_large)
            for i in range(0, 1201, 53):
                small[i] &= large[int(rng.integers(0, 4500))]
            total, cnt, _ = orc.subset_count(small, large, words(n))
            res, c, anyf = ixctx.subset_check(n, small, large, per_row=True)
            assert c == total and anyf == (total > 0)
            np.testing.assert_array_equal(res, cnt >= 1)
            _, c2, _ = ixctx.subset_check(n, small, large, count=True)
            assert c2 == total
            _, _, any3 = ixctx.subset_check(n, small, large, count=False)
            assert any3 == (total > 0)
    finally:
        ixctx.set_subset_walk(0)


def test_walk_orders_prune_the_same_sets(ixctx):
    """The exact search with either walk: same length, same sets kept (the
    dominance flags do not depend on the order states are walked in)."""
    n, k = 300, 2
    out = {}
    try:
        for walk in (1, 2, 3, 4):
            ixctx.set_subset_walk(walk)
            d = gen.random_dfa(n, k, 8200)
            out[walk] = ixctx.exact(n, k, d, cfg=sw.ExactConfig(dominance_pairs=10**11))
    finally:
        ixctx.set_subset_walk(0)
    a = out[1]
    for w in (2, 3, 4):
        b = out[w]
        assert a.found and b.found and a.length == b.length
        assert a.stats["sets_forward"] == b.stats["sets_forward"]
        assert a.stats["sets_backward"] == b.stats["sets_backward"]
        assert replay(n, k, d, b.word) == 1


@pytest.mark.parametrize("walk", [1, 2, 3, 4])
@pytest.mark.parametrize("proper", [False, True])
@pytest.mark.parametrize("n", [64, 300])
def test_leaf_flags_with_either_walk(ixctx, walk, proper, n):
    """Per-tile-set flags (the leaf checker) through the indexed kernel; with
    the rarest-first walk both lists are sorted by size and each flag is
    written at its set's place in the unsorted list."""
    rng = np.random.default_rng(40 + n + 2 * walk + proper)
    W = words(n)
    small = np.concatenate([random_sets(rng, 200, n, 0.05), random_sets(rng, 201, n, 0.2)])
    small = small[rng.permutation(401)]
    large = np.concatenate([random_sets(rng, 2500, n, 0.3), random_sets(rng, 2500, n, 0.8)])
    large = large[rng.permutation(5000)]
    large[:40] = small[100:140]  # equal sets: contained but not properly
    ixctx.set_subset_walk(walk)
    leaf = ixctx.leaf(n, small, 5000)
    try:
        for pos, cnt in [(0, 401), (100, 40), (7, 1)]:
            got = leaf.run(pos, cnt, large, proper=proper)
            want = np.zeros(5000, dtype=bool)
            for j in range(5000):
                _, cc, _ = orc.subset_count(small[pos:pos + cnt], large[j:j + 1], W, proper=proper)
                want[j] = cc.sum() > 0
            np.testing.assert_array_equal(got, want)
    finally:
        leaf.close()
        ixctx.set_subset_walk(0)
