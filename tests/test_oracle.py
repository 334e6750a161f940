"""The oracle pinned against the reference's golden vectors, brute force and
closed forms (CPU)."""
import numpy as np
import pytest

from conftest import golden
from helpers import random_sets, words
import oracle_lib as orc
from synchrowords_amd import gen


def _all(*names):
    out = []
    for nm in names:
        try:
            out += [pytest.param(c, id=c["name"]) for c in golden(nm)]
        except FileNotFoundError:
            pass
    return out


@pytest.mark.parametrize("case", _all("c1_random48.json", "readme.json", "cerny.json", "hard.json",
                                      "slowly_sink.json", "random_misc.json", "c3_random200.json"))
def test_oracle_pipeline_matches_reference(case):
    o = orc.solve(case["n"], case["k"], case["delta"])
    assert o["non_synchro"] == case["non_synchro"]
    if case["non_synchro"]:
        return
    assert (o["lower"], o["upper"]) == (case["lower"], case["upper"])
    if case["n"] > 15:
        assert o["eppstein"] == case["eppstein"]
        assert o["beam"] == case["beam"]


@pytest.mark.parametrize("case", [c for c in _all("c3_random200.json") if "trace" in c.values[0]])
def test_oracle_mitm_trace_matches_reference(case):
    """Per-step (phase, length, |F|, |B|) and the DFS entry equal the reference's log."""
    o = orc.solve(case["n"], case["k"], case["delta"], trace=True)
    steps = [s for s in o["trace"] if s[0] != "DFS"]
    assert steps == case["trace"]
    dfs = [s for s in o["trace"] if s[0] == "DFS"]
    if case["dfs_entry"]:
        assert dfs and dfs[0][1] == case["dfs_entry"][0] and dfs[0][2] == case["dfs_entry"][2]


def test_known_answers_hold_in_fixtures():
    for c in golden("cerny.json") + golden("slowly_sink.json"):
        assert c["upper"] == c["lower"] == c["closed_form"]
    for c in golden("hard.json"):
        assert c["upper"] == c["table_value"]
    for c in golden("readme.json"):
        if c["docs"] is None:
            assert c["non_synchro"]
        else:
            assert [c["lower"], c["upper"]] == c["docs"]


@pytest.mark.parametrize("n", [4, 6, 8, 10, 12])
def test_brute_force_and_exact_agree(n):
    for s in range(15):
        k = 2 + s % 2
        d = gen.random_dfa(n, k, 900 + 31 * n + s)
        b = orc.brute(n, k, d)
        o = orc.solve(n, k, d)
        if b < 0:
            assert o["non_synchro"]
        else:
            assert o["upper"] == b


def test_exact_agrees_with_brute_force_above_brute_range():
    """n in 16..20: the pipeline's Exact (Brute skipped) against power-set BFS."""
    for n in (16, 18, 20):
        for s in range(4):
            d = gen.random_dfa(n, 2, 4000 + n * 10 + s)
            b = orc.brute(n, 2, d)
            o = orc.solve(n, 2, d)
            if b < 0:
                assert o["non_synchro"]
            else:
                assert o["upper"] == o["lower"] == b


def test_cerny_closed_form_small():
    for n in range(3, 12):
        assert orc.brute(n, 2, gen.cerny(n)) == (n - 1) ** 2


def test_eppstein_word_synchronizes():
    for s in range(10):
        d = gen.random_dfa(30, 2, s)
        ns, bound, w = orc.eppstein(30, 2, d, 10**6)
        if not ns:
            assert len(w) == bound and orc.replay(30, 2, d, w) == 1


def test_subset_count_against_numpy():
    rng = np.random.default_rng(5)
    for n in (7, 64, 100, 200):
        W = words(n)
        a = random_sets(rng, 60, n, 0.1)
        b = random_sets(rng, 80, n, 0.7)
        a[3] = b[4]
        total, cnt, _ = orc.subset_count(a, b, W)
        # numpy: a_i subset b_j iff a_i & ~b_j == 0 in every word
        m = np.ones((60, 80), dtype=bool)
        for w in range(W):
            m &= (a[:, w:w + 1] & ~b[None, :, w]) == 0
        assert total == m.sum()
        np.testing.assert_array_equal(cnt, m.sum(axis=1))
        tp, cp, _ = orc.subset_count(a, b, W, proper=True)
        eq = np.ones((60, 80), dtype=bool)
        for w in range(W):
            eq &= a[:, w:w + 1] == b[None, :, w]
        assert tp == (m & ~eq).sum()


def test_apply_against_definition():
    n, k = 70, 3
    d = np.array(gen.random_dfa(n, k, 3)).reshape(n, k)
    rng = np.random.default_rng(1)
    sets = random_sets(rng, 40, n, 0.4)
    img = orc.apply(n, k, d.ravel(), sets)
    pre = orc.apply(n, k, d.ravel(), sets, backward=True)
    W = words(n)
    for i in range(40):
        members = [q for q in range(n) if (int(sets[i, q // 64]) >> (q % 64)) & 1]
        for a in range(k):
            im = {int(d[q, a]) for q in members}
            got = {q for q in range(n) if (int(img[a * 40 + i, q // 64]) >> (q % 64)) & 1}
            assert got == im
            pr = {p for p in range(n) if int(d[p, a]) in set(members)}
            got = {q for q in range(n) if (int(pre[a * 40 + i, q // 64]) >> (q % 64)) & 1}
            assert got == pr


def test_count_bits_against_unpackbits():
    """oracle_count_bits (subset.hpp:123-136) against numpy's bit unpacking,
    across word boundaries and with the tail bits above n clear."""
    rng = np.random.default_rng(11)
    for n in (1, 5, 63, 64, 65, 130, 256, 300, 512):
        sets = random_sets(rng, 257, n, 0.37)
        got = orc.count_bits(sets, n)
        bits = np.unpackbits(sets.view(np.uint8), axis=1, bitorder="little")
        want = bits.sum(axis=0, dtype=np.uint64)[:n]
        np.testing.assert_array_equal(got, want)
        assert bits[:, n:].sum() == 0


def test_eppstein_word_matches_reference_docs():
    """The reference prints Eppstein's greedy word for README automaton 3
    (docs/install.md:104, 25 letters): the oracle's pair tree and greedy
    choice give the same word, letter for letter."""
    case = [c for c in golden("readme.json") if c["name"] == "readme3"][0]
    n, k, d = case["n"], case["k"], case["delta"]
    ns, bound, w = orc.eppstein(n, k, d, (n - 1) ** 2 + 1)
    assert not ns
    assert w == case["docs_eppstein_word"]
    assert bound == len(w) == 25 == case["eppstein"]


def test_subset_count_mt_against_numpy():
    """The threaded all-pairs count (the microbench's full count) against a
    numpy all-pairs test and the scalar count, per row."""
    rng = np.random.default_rng(12)
    for n in (7, 64, 200, 256, 300):
        W = words(n)
        a = random_sets(rng, 150, n, 0.08)
        b = random_sets(rng, 3000, n, 0.75)
        a[7] = b[11]
        a[8] = 0
        m = np.ones((150, 3000), dtype=bool)
        for w in range(W):
            m &= (a[:, w:w + 1] & ~b[None, :, w]) == 0
        for threads in (1, 3, 0):
            total, per_row = orc.subset_count_mt(a, b, threads=threads, per_row=True)
            assert total == m.sum()
            np.testing.assert_array_equal(per_row, m.sum(axis=1))
        assert per_row[8] == 3000  # the empty set lies in every set


@pytest.mark.parametrize("name", ["random200_s0", "random200_s2"])
@pytest.mark.parametrize("mb", [1, 5, 10])
def test_oracle_memory_budget_keeps_the_length(name, mb):
    """The oracle's meet in the middle under a small memory budget switches
    to its DFS (dfs.cpp:63-347 restated) and still gives the reference's
    length."""
    c = [c for c in golden("c3_random200.json") if c["name"] == name][0]
    o = orc.solve(c["n"], c["k"], c["delta"], max_memory_mb=mb)
    assert o["entered_dfs"]
    assert (o["lower"], o["upper"]) == (c["upper"], c["upper"])


def test_reference_reduce_plan_keeps_the_exact_lengths():
    """The reference's configs/exact_reduce.json plan (its Reduce shrank
    every one of these automata, reduce.hpp:31-103, and a second build for
    the reduced N finished the search) ends at the same bounds as its
    configs/exact.json plan: the reduction never changes a length."""
    red = {c["name"]: c for c in golden("reduce_plan.json")}
    assert len(red) >= 20 and all(c["reduced_n"] and c["reduced_n"] < c["n"] for c in red.values())
    seen = 0
    for name in ("c3_random200.json", "c5_random300.json", "random_misc.json", "readme.json"):
        for c in golden(name):
            if c["name"] in red:
                r = red[c["name"]]
                assert r["delta"] == c["delta"]
                assert (r["non_synchro"], r["lower"], r["upper"]) == (int(bool(c["non_synchro"])), c["lower"],
                                                                      c["upper"]), name
                seen += 1
    assert seen == len(red)


@pytest.mark.parametrize("case", [pytest.param(c, id=c["name"]) for c in golden("reduce_plan.json")])
def test_oracle_reduce_matches_reference(case):
    """The oracle's Reduce (reduce.hpp:31-103 restated) takes as many
    forward steps as the reference's and reduces to the same number of
    states, from the plan's Beam bound with the plan's list_size_threshold
    (N * K * 8).  A restatement that skipped the re-sort before the
    self-reduction (the implicit trie needs its lists sorted) stopped one
    step early on ten of these automata."""
    n, k = case["n"], case["k"]
    up = orc.solve(n, k, case["delta"], run_exact=False)["upper"]
    steps, decided, _, rn = orc.reduce(n, k, case["delta"], up, n * k * 8)
    assert not decided
    assert (steps, rn) == (case["reduce_bfs_steps"], case["reduced_n"])
