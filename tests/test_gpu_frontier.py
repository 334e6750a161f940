"""One search step (sw_frontier_step) against the reference's bfs_step with a
visited list (synchrolib/algorithm/exact/meet_in_the_middle.cpp:186-229):
the images (or preimages) of every frontier set under every letter
(oracle_apply), each distinct set once, none of the visited list.  The
reference returns the sorted list; the device returns candidate order with
each set at its first candidate, which is checked exactly here, and as a
sorted set against numpy's set difference."""
import numpy as np
import pytest

import oracle_lib as orc
import synchrowords_amd as sw
from helpers import random_sets, words
from synchrowords_amd import gen

pytestmark = pytest.mark.gpu


def _rows(a):
    """Rows as hashable bytes (numpy unique over structured views)."""
    a = np.ascontiguousarray(a)
    return a.view(np.dtype((np.void, a.dtype.itemsize * a.shape[1]))).reshape(-1)


def _expected(cand, visited):
    """First-occurrence order of the distinct candidates not in visited."""
    rc = _rows(cand)
    _, first = np.unique(rc, return_index=True)
    first.sort()
    keep = first[~np.isin(rc[first], _rows(visited))] if visited.shape[0] else first
    return keep


def _automaton(n, k, seed):
    """Random DFA whose letter 0 merges states 2q and 2q+1 (so that frontier
    sets differing inside such pairs have equal images)."""
    d = np.array(gen.random_dfa(n, k, seed), dtype=np.uint32).reshape(n, k)
    for q in range(0, n - 1, 2):
        d[q + 1, 0] = d[q, 0]
    return d.reshape(-1)


@pytest.mark.parametrize("n,k,m,p,backward", [
    (5, 2, 40, 0.4, False), (40, 2, 3000, 0.2, False), (64, 3, 5000, 0.1, False), (130, 3, 4000, 0.1, True),
    (300, 2, 20000, 0.05, False), (300, 2, 3000, 0.3, True), (512, 2, 2500, 0.05, False),
    # past 2^17 candidates: the bit-sliced expansion
    (64, 3, 50000, 0.1, False), (300, 2, 100000, 0.3, True), (200, 2, 70000, 0.7, True),
    (512, 2, 70000, 0.5, True), (130, 4, 40000, 0.2, False)])
def test_frontier_step_matches_reference_step(ctx, n, k, m, p, backward):
    rng = np.random.default_rng(n * 7 + m)
    d = _automaton(n, k, n + m)
    f = random_sets(rng, m, n, p)
    f[m // 3: m // 3 + m // 10] = f[: m // 10]          # repeated frontier sets
    cand = orc.apply(n, k, d, f, backward=backward)
    picks = rng.choice(cand.shape[0], size=cand.shape[0] // 3, replace=False)
    visited = np.concatenate([cand[picks], random_sets(rng, 500, n, p)])
    visited = visited[rng.permutation(visited.shape[0])]
    got, par, let, st = ctx.frontier_step(n, k, d, f, visited, backward=backward)
    keep = _expected(cand, visited)
    np.testing.assert_array_equal(got, cand[keep])
    np.testing.assert_array_equal(par, keep % m)
    np.testing.assert_array_equal(let, keep // m)
    # as sets: the reference's sorted unique candidates minus the visited list
    want = np.setdiff1d(np.unique(_rows(cand)), _rows(visited))
    np.testing.assert_array_equal(np.sort(_rows(got)), want)
    assert st["candidates"] == k * m and st["new_sets"] == got.shape[0]
    assert st["probes"] >= k * m and st["table_slots"] >= 2 * (k * m + visited.shape[0])


def test_frontier_step_edges(ctx):
    n, k = 20, 2
    d = _automaton(n, k, 3)
    W = words(n)
    empty = np.zeros((0, W), dtype=np.uint64)
    got, _, _, st = ctx.frontier_step(n, k, d, empty, empty)
    assert got.shape[0] == 0 and st["candidates"] == 0
    # every candidate already visited
    f = random_sets(np.random.default_rng(1), 100, n, 0.3)
    cand = orc.apply(n, k, d, f)
    got, _, _, _ = ctx.frontier_step(n, k, d, f, cand)
    assert got.shape[0] == 0
    # no visited list: plain deduplication, with duplicates in visited allowed
    got, _, _, _ = ctx.frontier_step(n, k, d, f, np.concatenate([cand[:5], cand[:5]]))
    np.testing.assert_array_equal(got, cand[_expected(cand, cand[:5])])


def test_frontier_step_past_a_million_tiles_of_scan(ctx):
    """2^23 candidates: the keep scan's tile sums take a scan of their own
    (4096 tiles of 2048 entries)."""
    n, k, m = 64, 2, 1 << 22
    rng = np.random.default_rng(5)
    d = _automaton(n, k, 11)
    f = rng.integers(0, 2**63, size=(m, 1), dtype=np.uint64) & rng.integers(0, 2**63, size=(m, 1), dtype=np.uint64)
    f &= rng.integers(0, 2**63, size=(m, 1), dtype=np.uint64)
    cand = orc.apply(n, k, d, f)
    visited = cand[rng.choice(cand.shape[0], size=1 << 20, replace=False)]
    got, par, let, st = ctx.frontier_step(n, k, d, f, visited)
    c = cand[:, 0]
    _, first = np.unique(c, return_index=True)
    first.sort()
    keep = first[~np.isin(c[first], visited[:, 0])]
    assert got.shape[0] == keep.shape[0], (got.shape[0], keep.shape[0])
    np.testing.assert_array_equal(got[:, 0], c[keep])
    np.testing.assert_array_equal(par, keep % m)
    np.testing.assert_array_equal(let, keep // m)
