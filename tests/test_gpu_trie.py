"""The device subset trie (sw_trie_*, the reference's SubsetsTrie,
synchrolib/data_structures/subsets_trie.hpp) against brute force: per query,
whether some trie set lies inside it (the reference's contains_subset_of) or
contains it, for lists that take the one-workgroup build, the level-by-level
build and the leaf of equal sets; and the trie's shape against the
reference's rules (division bit = the most frequent state not held by every
set of the node, lowest on ties, subsets_trie.hpp:299-350)."""
import numpy as np
import pytest

import oracle_lib as orc
import synchrowords_amd as sw
from helpers import random_sets, words
from synchrowords_amd import gen

pytestmark = pytest.mark.gpu

LEAF = 0xFFFF


def _subset_matrix(a, b):
    """m[i, j] = a_i is a subset of b_j (numpy, all pairs)."""
    m = np.ones((a.shape[0], b.shape[0]), dtype=bool)
    for w in range(a.shape[1]):
        m &= (a[:, w:w + 1] & ~b[None, :, w]) == 0
    return m


def _popcount(sets):
    return np.unpackbits(sets.view(np.uint8), axis=1).sum(axis=1)


def _check_shape(n, sets_in, trie):
    nodes, tsets, rows = trie.dump()
    m = sets_in.shape[0]
    assert sorted(rows.tolist()) == list(range(m))
    np.testing.assert_array_equal(tsets, sets_in[rows])
    # Walk from the root: every node's sets are a contiguous range; children
    # split it on the division bit, sets without it first.
    seen_leaf_sets = 0
    stack = [(0, 0, m, 0)]
    pops = _popcount(tsets)
    while stack:
        v, b, e, depth = stack.pop()
        zero, one, z, mx = (int(x) for x in nodes[v])
        bit, mn = z & 0xFFFF, z >> 16
        assert mn == pops[b:e].min() and mx == pops[b:e].max(), (v, depth)
        if bit == LEAF:
            assert zero == b and one == e - b
            assert e - b <= 8 or depth >= 48 or (tsets[b:e] == tsets[b]).all()
            seen_leaf_sets += e - b
            continue
        part = tsets[b:e]
        has = ((part[:, bit // 64] >> np.uint64(bit % 64)) & np.uint64(1)).astype(bool)
        counts = np.unpackbits(part.view(np.uint8), axis=1, bitorder="little").sum(axis=0)[:n]
        ok = (counts > 0) & (counts < e - b)
        assert ok.any() and e - b > 8 and depth < 48
        best = int(np.flatnonzero(ok & (counts == counts[ok].max()))[0])
        assert bit == best, (v, bit, best)
        zc = int((~has).sum())
        assert not has[:zc].any() and has[zc:].all()
        stack.append((zero, b, b + zc, depth + 1))
        stack.append((one, b + zc, e, depth + 1))
    assert seen_leaf_sets == m


@pytest.mark.parametrize("n,m,p_trie,p_query", [
    (5, 20, 0.3, 0.7), (64, 1, 0.2, 0.6), (64, 9, 0.2, 0.6), (65, 3000, 0.1, 0.7), (130, 2048, 0.1, 0.5),
    (200, 5000, 0.08, 0.6), (300, 1024, 0.1, 0.4), (300, 1500, 0.1, 0.4), (300, 20000, 0.1, 0.5),
    (512, 700, 0.05, 0.5), (512, 3000, 0.05, 0.5)])
def test_trie_queries_match_brute_force(ctx, n, m, p_trie, p_query):
    rng = np.random.default_rng(n * 1000 + m)
    tsets = random_sets(rng, m, n, p_trie)
    queries = random_sets(rng, 2500, n, p_query)
    queries[3] = tsets[m // 2]                       # an exact match
    queries[4] = tsets[0] | queries[4]               # a superset of a trie set
    queries[5] = 0                                   # the empty set
    t = sw.Trie(ctx, n, tsets)
    try:
        _check_shape(n, tsets, t)
        sub = _subset_matrix(tsets, queries)         # trie set inside query
        got, anyf, wit = t.query(queries, superset=False)
        np.testing.assert_array_equal(got, sub.any(axis=0))
        assert anyf == sub.any()
        _, anyf, wit = t.query(queries, superset=False, per_query=False)
        assert anyf == sub.any()
        if anyf:
            assert sub[wit & 0xFFFFFFFF, wit >> 32]
        sup = _subset_matrix(queries, tsets)         # query inside trie set
        got, anyf, _ = t.query(queries, superset=True)
        np.testing.assert_array_equal(got, sup.any(axis=1))
        # the oracle's count agrees on the superset flags
        _, cnt, _ = orc.subset_count(queries, tsets, words(n))
        np.testing.assert_array_equal(got, cnt > 0)
        _, anyf, wit = t.query(queries, superset=True, per_query=False)
        assert anyf == sup.any()
        if anyf:
            assert sup[wit >> 32, wit & 0xFFFFFFFF]
    finally:
        t.close()


def test_trie_of_equal_sets_and_empty(ctx):
    """More equal sets than one workgroup builds (one leaf over all of them),
    and an empty trie."""
    n = 40
    tsets = np.tile(np.array([[0x00F0F0F0F0]], dtype=np.uint64), (6000, 1))
    t = sw.Trie(ctx, n, tsets)
    try:
        assert t.info()["top_levels"] >= 1
        nodes, _, _ = t.dump()
        assert nodes[0, 2] & 0xFFFF == LEAF and nodes[0, 1] == 6000
        q = np.array([[0x00F0F0F0F0], [0x00F0F0F0F1], [0x0070F0F0F0], [0]], dtype=np.uint64)
        got, _, _ = t.query(q)
        assert got.tolist() == [True, True, False, False]
        got, _, _ = t.query(q, superset=True)
        assert got.tolist() == [True, False, True, True]
    finally:
        t.close()
    t = sw.Trie(ctx, n, np.zeros((0, 1), dtype=np.uint64))
    try:
        got, anyf, _ = t.query(np.zeros((5, 1), dtype=np.uint64))
        assert not got.any() and not anyf
    finally:
        t.close()


@pytest.mark.parametrize("seed", [1000, 1002])
def test_trie_on_search_frontiers(ctx, seed):
    """Forward images of Q and backward preimages of singletons of an n=300
    automaton (the lists the search's meet checks hold): flags of both
    query kinds against brute force."""
    n, k = 300, 2
    d = gen.random_dfa(n, k, seed)
    W = words(n)
    q_all = np.zeros((1, W), dtype=np.uint64)
    for s in range(n):
        q_all[0, s // 64] |= np.uint64(1) << np.uint64(s % 64)
    fwd = q_all
    for _ in range(12):
        fwd = np.unique(orc.apply(n, k, d, fwd), axis=0)
    dd = np.array(d).reshape(n, k)
    singles = []
    for s in range(n):
        if any((dd[:, a] == s).sum() > 1 for a in range(k)):
            x = np.zeros(W, dtype=np.uint64)
            x[s // 64] = np.uint64(1) << np.uint64(s % 64)
            singles.append(x)
    bwd = np.array(singles)
    for _ in range(10):
        bwd = orc.apply(n, k, d, bwd, backward=True)
        bwd = bwd[_popcount(bwd) > 1]
        bwd = np.unique(bwd, axis=0)
    t = sw.Trie(ctx, n, fwd)
    try:
        got, _, _ = t.query(bwd)
        np.testing.assert_array_equal(got, _subset_matrix(fwd, bwd).any(axis=0))
    finally:
        t.close()
    t = sw.Trie(ctx, n, bwd)
    try:
        got, _, _ = t.query(fwd, superset=True)
        np.testing.assert_array_equal(got, _subset_matrix(fwd, bwd).any(axis=1))
    finally:
        t.close()
