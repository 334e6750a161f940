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
        for seed in range(2):
            d = gen.random_dfa(n, k, 8100 + seed)
            a = ixctx.exact(n, k, d)
            b = bw.exact(n, k, d)
            assert a.found and b.found and a.length == b.length
            assert a.stats["sets_forward"] == b.stats["sets_forward"]
            assert a.stats["sets_backward"] == b.stats["sets_backward"]
    finally:
        bw.close()
