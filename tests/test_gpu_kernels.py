"""Kernel-level entry points on the GPU against the oracle (bit-exact)."""
import numpy as np
import pytest

from helpers import random_sets, slice_table, words
import oracle_lib as orc
from synchrowords_amd import gen

pytestmark = pytest.mark.gpu

SIZES = [1, 5, 63, 64, 65, 100, 128, 200, 256, 257, 300, 448, 512]


@pytest.mark.parametrize("n", SIZES)
def test_subset_check_counts(ctx, n):
    rng = np.random.default_rng(n)
    W = words(n)
    small = random_sets(rng, 700, n, 0.08 if n > 20 else 0.3)
    large = random_sets(rng, 3000, n, 0.75)
    # Plant exact subsets and equal sets.
    for i in range(0, 700, 50):
        j = int(rng.integers(0, 3000))
        small[i] = small[i] & large[j]
    small[7] = large[11]
    total, cnt, _ = orc.subset_count(small, large, W)
    res, c, anyf = ctx.subset_check(n, small, large, thresh=1, per_row=True)
    assert c == total
    assert anyf == (total > 0)
    np.testing.assert_array_equal(res, cnt >= 1)
    res2, c2, _ = ctx.subset_check(n, small, large, thresh=2, per_row=True)
    np.testing.assert_array_equal(res2, cnt >= 2)
    _, c3, any3 = ctx.subset_check(n, small, large, count=True)
    assert c3 == total and any3 == (total > 0)
    _, _, any4 = ctx.subset_check(n, small, large, count=False)
    assert any4 == (total > 0)


@pytest.mark.parametrize("shape", [(0, 10), (10, 0), (1, 1), (1, 4097), (4097, 1), (5000, 4500)])
def test_subset_check_shapes(ctx, shape):
    n = 150
    rng = np.random.default_rng(sum(shape))
    small = random_sets(rng, shape[0], n, 0.05)
    large = random_sets(rng, shape[1], n, 0.8)
    total, _, _ = orc.subset_count(small, large, words(n))
    _, c, anyf = ctx.subset_check(n, small, large, count=True)
    assert c == total and anyf == (total > 0)


def test_subset_check_edge_sets(ctx):
    n = 70
    W = words(n)
    full = np.array([[~np.uint64(0), (np.uint64(1) << np.uint64(6)) - np.uint64(1)]], dtype=np.uint64)
    empty = np.zeros((1, W), dtype=np.uint64)
    small = np.vstack([empty, full, empty])
    large = np.vstack([full, empty])
    total, cnt, _ = orc.subset_count(small, large, W)
    res, c, _ = ctx.subset_check(n, small, large, per_row=True)
    assert c == total == 2 + 1 + 2
    np.testing.assert_array_equal(res, cnt >= 1)


@pytest.mark.parametrize("n", [9, 64, 130, 300])
@pytest.mark.parametrize("proper", [False, True])
def test_leaf(ctx, n, proper):
    rng = np.random.default_rng(n + proper)
    W = words(n)
    small = random_sets(rng, 400, n, 0.1)
    large = random_sets(rng, 2500, n, 0.7)
    large[:50] = small[100:150]  # equal sets: contained but not properly
    leaf = ctx.leaf(n, small, 2500)
    for pos, cnt in [(0, 400), (100, 50), (399, 1), (10, 0)]:
        got = leaf.run(pos, cnt, large, proper=proper)
        # Oracle: large_j flagged iff some small in range is a (proper) subset.
        want = np.zeros(2500, dtype=bool)
        for j in range(2500):
            _, c, _ = orc.subset_count(small[pos:pos + cnt], large[j:j + 1], W, proper=proper)
            want[j] = c.sum() > 0
        np.testing.assert_array_equal(got, want)
    leaf.close()


@pytest.mark.parametrize("n,k", [(48, 2), (100, 3), (300, 2), (512, 2)])
def test_expand_and_tables(ctx, n, k):
    delta = gen.random_dfa(n, k, 17 + n)
    rng = np.random.default_rng(n)
    sets = random_sets(rng, 1500, n, 0.3)
    for backward in (False, True):
        got = ctx.expand(n, k, delta, sets, backward=backward)
        want = orc.apply(n, k, delta, sets, backward=backward)
        np.testing.assert_array_equal(got, want)
    table = slice_table(n, k, delta, 1, backward=True)
    got = ctx.transition_apply(n, table, sets)
    want = orc.apply(n, k, delta, sets, backward=True)[1500:3000]
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("n", [3, 64, 200, 512])
def test_count_bits(ctx, n):
    rng = np.random.default_rng(n)
    sets = random_sets(rng, 20000, n, 0.4)
    np.testing.assert_array_equal(ctx.count_bits(n, sets), orc.count_bits(sets, n))


@pytest.mark.parametrize("n", [100, 256, 300])
def test_every_launch_shape_gives_the_same_result(n):
    """The autotuner may pick any launch shape; each one must count the same pairs."""
    import synchrowords_amd as sw
    rng = np.random.default_rng(1000 + n)
    W = words(n)
    small = random_sets(rng, 1500, n, 0.07)
    large = random_sets(rng, 5000, n, 0.75)
    for i in range(0, 1500, 40):
        small[i] = small[i] & large[int(rng.integers(0, 5000))]
    total, cnt, _ = orc.subset_count(small, large, W)
    c = sw.Context(0)
    try:
        for mb in ([32, 64] if n <= 318 else [32]):
            for threads in (256, 512, 1024):
                for wgs in (1, 7, 256, 4096):
                    for ee in (0, 1):
                        c.set_subset_shape(n, 1500, 5000, threads, wgs, ee, mb)
                        res, got, anyf = c.subset_check(n, small, large, per_row=True)
                        assert got == total, (mb, threads, wgs, ee)
                        np.testing.assert_array_equal(res, cnt >= 1)
    finally:
        c.close()


def test_autotuned_shape_keeps_results():
    import synchrowords_amd as sw
    n = 200
    rng = np.random.default_rng(5)
    small = random_sets(rng, 4096, n, 0.08)
    large = random_sets(rng, 8192, n, 0.75)
    small[3] = large[17]
    total, _, _ = orc.subset_count(small, large, words(n))
    c = sw.Context(0)
    try:
        shape = c.autotune_subset(n, 4096, 8192, 0.08, 0.75, reps=1)
        assert shape["threads"] in (256, 512, 1024) and shape["mask_bits"] in (32, 64) and shape["ms"] > 0
        _, got, anyf = c.subset_check(n, small, large, count=True)
        assert got == total and anyf
    finally:
        c.close()


def test_microbench_count_matches_oracle_golden(ctx):
    """The full subset-check microbench (2^20 x 2^20 sets, n = 256): the
    device count of all 2^40 pairs equals the oracle's all-pairs count
    committed in tests/golden/c2_count.json, and the planted subsets make the
    existence flag true."""
    import json
    import bench
    from conftest import GOLDEN
    g = json.loads((GOLDEN / "c2_count.json").read_text())
    small, large, _ = bench.make_lists(g["count"], g["n"], g["p_small"], g["p_large"], seed=g["seed"])
    _, cnt, anyf = ctx.subset_check(g["n"], small, large, count=True)
    assert anyf and cnt == g["pair_count"]
    _, cnt, _ = ctx.subset_check(g["n"], small[:g["block_rows"]], large, count=True)
    assert cnt == g["block_counts"][0]
