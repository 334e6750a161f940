"""Eppstein's greedy bound on the device (eppstein.hpp:25-114 over
pairs_tree.hpp:25-65) against the reference's bounds (golden files written
by the reference binary) and the oracle's pair BFS and greedy word."""
import numpy as np
import pytest

import oracle_lib as orc
from conftest import golden
from synchrowords_amd import gen, replay

pytestmark = pytest.mark.gpu

GOLDEN = ["hard.json", "cerny.json", "slowly_sink.json", "readme.json", "random_misc.json", "c1_random48.json",
          "c3_random200.json", "c5_random300.json"]


def _golden_cases():
    out = []
    for f in GOLDEN:
        for c in golden(f):
            # With n <= 15 the plan's Brute settles the length first and
            # Eppstein does not run.
            if c["n"] > 15 or c["non_synchro"]:
                out.append(c)
    return out


@pytest.mark.parametrize("case", _golden_cases(), ids=lambda c: c["name"])
def test_eppstein_matches_reference_bound(ctx, case):
    n, k, d = case["n"], case["k"], case["delta"]
    up = (n - 1) ** 2 + 1
    ns, bound, word = ctx.eppstein(n, k, d, up)
    if case["non_synchro"]:
        assert ns
        return
    assert not ns and bound == case["eppstein"]
    ons, obound, oword = orc.eppstein(n, k, d, up)
    assert (ons, obound) == (ns, bound)
    if word is not None:
        assert word == oword and len(word) == bound
        assert replay(n, k, d, word) == 1


@pytest.mark.parametrize("n,k", [(2, 2), (3, 1), (7, 2), (16, 3), (33, 2), (64, 4), (129, 2), (300, 2), (512, 2),
                                 (100, 7)])
def test_eppstein_random_against_oracle(ctx, n, k):
    for seed in range(3):
        d = gen.random_dfa(n, k, 4100 + 13 * n + seed)
        up = (n - 1) ** 2 + 1
        got = ctx.eppstein(n, k, d, up)
        want = orc.eppstein(n, k, d, up)
        assert got[0] == want[0] and got[1] == want[1], seed
        if not got[0]:
            assert (got[2] or []) == want[2]


def _two_sinks(n, k=2):
    """Two sink cycles 0..n/2-1 and n/2..n-1: no pair across them merges."""
    h2 = n // 2
    d = [0] * (n * k)
    for q in range(n):
        h = (q // h2) * h2
        d[q * k + 0] = h + (q + 1) % h2
        d[q * k + 1] = h + (q * 7) % h2
    return d


@pytest.mark.parametrize("n", [40, 80])  # one workgroup; the whole device (one automaton of 64 states or more)
def test_eppstein_non_synchronizing(ctx, n):
    k = 2
    d = _two_sinks(n, k)
    ns, bound, word = ctx.eppstein(n, k, d, 100)
    assert ns and bound == 100 and word is None
    assert orc.eppstein(n, k, d, 100)[0]


def test_eppstein_single_and_batch_paths_agree(ctx):
    """One automaton of 128 states or more runs its pair BFS over the whole
    device; in a batch the same automaton runs in one workgroup.  Flags,
    bounds and words are the same."""
    rng = np.random.default_rng(21)
    inst = [(int(rng.integers(128, 300)), int(rng.integers(2, 4)), None) for _ in range(12)]
    inst = [(n, k, gen.random_dfa(n, k, 9100 + j)) for j, (n, k, _) in enumerate(inst)]
    inst += [(128, 2, gen.cerny(128)), (64, 2, gen.cerny(64)), (160, 2, _two_sinks(160))]
    ups = [(n - 1) ** 2 + 1 for n, _, _ in inst]
    bns, bbounds, bwords = ctx.eppstein_batch(inst, ups, word_cap=1 << 16)
    for j, (n, k, d) in enumerate(inst):
        ns, bound, word = ctx.eppstein(n, k, d, ups[j])
        assert bool(ns) == bool(bns[j]) and bound == int(bbounds[j]), j
        assert (list(word) if word else None) == (bwords[j] or None), j  # no word: not improved
        if word:
            assert replay(n, k, d, word) == 1


def test_eppstein_upper_not_improved(ctx):
    n, k = 64, 2
    d = gen.cerny(n)
    ns, bound, word = ctx.eppstein(n, k, d, 100)  # greedy length is far above 100
    assert not ns and bound == 100 and word is None
    assert orc.eppstein(n, k, d, 100)[1] == 100


def test_eppstein_batch(ctx):
    rng = np.random.default_rng(8)
    inst = []
    for j in range(200):
        n = int(rng.integers(1, 120))
        k = int(rng.integers(1, 4))
        inst.append((n, k, gen.random_dfa(n, k, 8800 + j)))
    inst.append((30, 2, gen.cerny(30)))
    ups = [(n - 1) ** 2 + 1 for n, _, _ in inst]
    ups[-1] = 50  # not improved
    ns, bounds, words = ctx.eppstein_batch(inst, ups, word_cap=4096)
    for j, (n, k, d) in enumerate(inst):
        w = orc.eppstein(n, k, d, ups[j])
        assert bool(ns[j]) == w[0] and int(bounds[j]) == w[1], j
        if not ns[j] and w[2]:
            assert words[j] == w[2], j


def test_eppstein_word_matches_reference_docs(ctx):
    """README automaton 3: the reference prints Eppstein's 25-letter word
    (docs/install.md:104); the device pair tree and greedy phase give the
    same word letter for letter."""
    case = [c for c in golden("readme.json") if c["name"] == "readme3"][0]
    n, k, d = case["n"], case["k"], case["delta"]
    ns, bound, word = ctx.eppstein(n, k, d, (n - 1) ** 2 + 1)
    assert not ns and bound == 25
    assert word == case["docs_eppstein_word"]
