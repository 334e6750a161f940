"""Exact search and plan pipeline on the GPU against the reference's golden
vectors (tests/golden/, written by the compiled reference) and the oracle."""
import pytest

from conftest import golden
from synchrowords_amd import gen, replay

pytestmark = pytest.mark.gpu


def _check_word(case, r_len, word):
    assert word is not None, "no word returned"
    assert len(word) == r_len
    assert replay(case["n"], case["k"], case["delta"], word) == 1


def _cases(*names):
    out = []
    for nm in names:
        out += [pytest.param(c, id=c["name"]) for c in golden(nm) if not c["non_synchro"]]
    return out


@pytest.mark.parametrize("case", _cases("c1_random48.json", "cerny.json", "hard.json", "slowly_sink.json",
                                        "readme.json"))
def test_exact_trivial_bound(ctx, case):
    """Exact alone, from the trivial bound (n-1)^2+1: finds the reference's length and a word."""
    n, k = case["n"], case["k"]
    if n == 1:
        pytest.skip("n=1 has the empty word")
    r = ctx.exact(n, k, case["delta"])
    assert r.found
    assert r.length == case["upper"]
    if r.length < (n - 1) ** 2 + 1:
        _check_word(case, r.length, r.word)


@pytest.mark.parametrize("case", _cases("c1_random48.json", "readme.json", "c3_random200.json",
                                        "c5_random300.json"))
def test_exact_with_beam_bound(ctx, case):
    """Exact given the reference's Beam bound: proves it optimal (or finds shorter)."""
    n, k = case["n"], case["k"]
    if n <= 15:
        pytest.skip("Brute decides n <= 15")
    r = ctx.exact(n, k, case["delta"], upper=case["beam"])
    assert r.found and r.length == case["upper"]
    if not r.bound_word:
        _check_word(case, r.length, r.word)


def _pipeline_cases():
    out = []
    for nm in ("c1_random48.json", "readme.json", "cerny.json", "hard.json", "slowly_sink.json",
               "random_misc.json", "c3_random200.json", "c5_random300.json"):
        try:
            out += [pytest.param(c, id=c["name"]) for c in golden(nm)]
        except FileNotFoundError:
            pass
    return out


@pytest.mark.parametrize("case", _pipeline_cases())
def test_pipeline_matches_reference(ctx, case):
    """configs/exact.json plan: bounds of every stage and the final length."""
    n, k = case["n"], case["k"]
    r = ctx.pipeline(n, k, case["delta"])
    assert r.non_synchro == bool(case["non_synchro"])
    if case["non_synchro"]:
        return
    assert (r.lower, r.upper) == (case["lower"], case["upper"])
    if n > 15:
        assert r.eppstein_bound == case["eppstein"]
        assert r.beam_bound == case["beam"]
    if r.upper > 0 and r.word is not None:
        _check_word(case, r.upper, r.word)
    if n > 15:
        assert r.word is not None


def test_cerny64_word(ctx):
    d = gen.cerny(64)
    r = ctx.exact(64, 2, d)
    assert r.found and r.length == 63 * 63
    assert replay(64, 2, d, r.word) == 1


@pytest.mark.parametrize("case", _cases("c5_random300.json") + _cases("c3_random200.json")[:3])
def test_bound_planned_side_order_keeps_the_length(ctx, case):
    """The side order planned by the upper bound (side_rule 2: the remaining
    steps below the Beam bound ordered by a cost model) reaches the
    reference's length, like the balanced order: the order of the steps
    never changes the answer."""
    import synchrowords_amd as sw
    n, k = case["n"], case["k"]
    r = ctx.exact(n, k, case["delta"], upper=case["beam"], cfg=sw.ExactConfig(side_rule=2))
    assert r.found and r.length == case["upper"]
    if not r.bound_word:
        _check_word(case, r.length, r.word)


@pytest.mark.parametrize("n,k", [(40, 3), (90, 2), (90, 3), (150, 2), (260, 2), (300, 2)])
def test_dominance_pruning_keeps_the_length(ctx, n, k):
    """Dropping dominated sets (both sides) never changes the shortest length:
    the search with pruning against the search with pruning switched off."""
    import synchrowords_amd as sw
    for seed in range(3):
        d = gen.random_dfa(n, k, 7000 + seed)
        off = ctx.exact(n, k, d, cfg=sw.ExactConfig(dominance_pairs=1))
        on = ctx.exact(n, k, d)
        every = ctx.exact(n, k, d, cfg=sw.ExactConfig(dominance_pairs=10**11))  # small passes too
        assert on.found and off.found and every.found
        assert on.length == off.length == every.length
        for r in (on, off, every):
            assert replay(n, k, d, r.word) == 1


@pytest.mark.parametrize("n,k,width", [(100, 2, 9000), (300, 2, 12000), (100, 3, 6000), (300, 2, 100000),
                                       (200, 3, 70000)])
def test_device_beam_wide_crosses_the_counted_order(ctx, n, k, width):
    """Widths whose depths hold more than 16384 distinct candidates select
    the kept sets by radix selection there (one workgroup up to 2^17
    candidates, over the whole device past that) and count the order below:
    the oracle's bound."""
    import oracle_lib as orc
    d = gen.random_dfa(n, k, 9300 + n)
    cap = (n - 1) ** 2 + 1
    want = orc.beam_width(n, k, d, width, cap)
    b, w = ctx.beam(n, k, d, width, cap)
    assert b == want
    if b < cap:
        assert len(w) == b and replay(n, k, d, w) == 1


@pytest.mark.parametrize("n,k,width", [(64, 3, 3000), (50, 4, 1500), (64, 2, 3000), (33, 16, 400), (64, 2, 1800),
                                       (48, 2, 268)])
def test_device_beam_small_automata_hand_over(ctx, n, k, width):
    """Automata of at most 64 states run the Beam's depths in one workgroup
    until a depth's candidates outgrow its LDS (4096) or it keeps more than
    2048 sets, then the device-wide beam continues from the list and history
    it hands back; narrower beams (1800, and the plan's 268 at n = 48) stay
    in the workgroup and keep each depth's sets by selection: the oracle's
    bound at the same width, and a word that replays."""
    import oracle_lib as orc
    for seed in range(2):
        d = gen.random_dfa(n, k, 9500 + seed + n)
        cap = (n - 1) ** 2 + 1
        want = orc.beam_width(n, k, d, width, cap)
        b, w = ctx.beam(n, k, d, width, cap)
        assert b == want, (seed, b, want)
        if b < cap:
            assert len(w) == b and replay(n, k, d, w) == 1


@pytest.mark.parametrize("n,k", [(20, 2), (48, 2), (64, 3), (100, 2), (200, 2), (300, 2), (130, 4), (512, 2)])
@pytest.mark.parametrize("width", [1, 7, 64, 500])
def test_device_beam_matches_reference_order(ctx, n, k, width):
    """The device beam in the reference's order gives the oracle's Beam bound
    (restated from beam.hpp:98-188) at the same width, and its word."""
    import oracle_lib as orc
    for seed in range(2):
        d = gen.random_dfa(n, k, 300 + 7 * seed + width)
        cap = (n - 1) ** 2 + 1
        want = orc.beam_width(n, k, d, width, cap)
        b, w = ctx.beam(n, k, d, width, cap)
        assert b == want, (seed, b, want)
        if b < cap:
            assert len(w) == b and replay(n, k, d, w) == 1


@pytest.mark.parametrize("case", _cases("c5_random300.json"))
def test_learned_bound_exact_search(ctx, case):
    """The learned-bound search of the n=300 workload: the committed network
    orders a device beam (width 64), its bound and word cap the exact
    search, which ends at the reference's length with a word that replays."""
    import synchrowords_amd as sw
    from pathlib import Path
    from synchrowords_amd.train_scorer import load_scorer
    n, k, d = case["n"], case["k"], case["delta"]
    m = load_scorer(ctx, Path(sw.__file__).parent / "models" / "scorer_n300_k2.npz")
    try:
        bound, bword = m.beam(d, 64, (n - 1) ** 2 + 1)
    finally:
        m.close()
    assert len(bword) == bound and replay(n, k, d, bword) == 1
    assert bound >= case["upper"]
    r = ctx.exact(n, k, d, upper=bound, upper_word=bword)
    assert r.found and r.length == case["upper"]
    _check_word(case, r.length, r.word)


def _trie_cases():
    """(mode, case): every mode on C1 and C3; modes 1 and 2 on C5 (superset
    walks over a backward trie visit most of its nodes, long at n = 300)."""
    out = []
    for nm in ("c1_random48.json", "c3_random200.json", "c5_random300.json"):
        cs = [c for c in golden(nm) if not c["non_synchro"]]
        cs = cs[:12] if nm == "c1_random48.json" else cs
        for mode in ((1, 2) if nm == "c5_random300.json" else (1, 2, 3)):
            out += [pytest.param(mode, c, id=f"{mode}-{c['name']}") for c in cs]
    return out


@pytest.mark.parametrize("mode,case", _trie_cases())
def test_meet_checks_by_kernel_or_trie(ctx, mode, case):
    """Meet checks by the vertical subset kernel (1), by the device subset
    trie over the forward frontier (2), or by tries over the frontier that
    did not step (3: superset queries after forward steps), from the trivial
    bound so that the search itself finds the meet: the reference's length,
    and the word built from the witness replays."""
    import synchrowords_amd as sw
    n, k, d = case["n"], case["k"], case["delta"]
    r = ctx.exact(n, k, d, cfg=sw.ExactConfig(goal_check=mode))
    assert r.found and not r.bound_word and r.length == case["upper"]
    _check_word(case, r.length, r.word)
    if mode >= 2:
        assert r.stats["trie_builds"] > 0 and r.stats["trie_queries"] > 0
    else:
        assert r.stats["trie_builds"] == 0


def _budget_cases():
    out = [pytest.param(c, id=c["name"]) for c in golden("c3_random200.json")[:5]]
    out += [pytest.param(c, id=c["name"]) for c in golden("c5_random300.json")[:2]]
    return out


def _search_mb(r, n):
    """Device MB an unconstrained search held (sets of both sides, row plus
    parent and letter, their hash tables and frontiers), rounded up."""
    W = (n + 63) // 64
    sets = r.stats["sets_forward"] + r.stats["sets_backward"]
    return max(1, (sets * (8 * W + 5 + 16 + 8 * W)) >> 20)


@pytest.mark.parametrize("case", _budget_cases())
@pytest.mark.parametrize("fraction,goal", [(2, 0), (4, 0), (4, 2)])
def test_memory_budget_deepens_the_backward_side(ctx, case, fraction, goal):
    """With a device budget of a half or a quarter of what the search's
    levels take, a step passes the budget and the search deepens the
    backward frontier in parts (the reference's DFS, dfs.cpp:63-347), each
    list checked against the forward frontier by the subset kernel (goal 0)
    or the forward trie (goal 2): it still ends at the reference's length,
    with a word that replays unless the Beam bound's word is the answer."""
    import synchrowords_amd as sw
    n, k, d = case["n"], case["k"], case["delta"]
    full = ctx.exact(n, k, d, upper=case["beam"])
    mb = max(1, _search_mb(full, n) // fraction)
    r = ctx.exact(n, k, d, upper=case["beam"],
                  cfg=sw.ExactConfig(max_memory_mb=mb, dfs_min_list_size=1024, goal_check=goal))
    assert r.stats["dfs_entered"] == 1, (mb, r.stats)
    assert (r.stats["trie_queries"] > 0) == (goal == 2)
    assert r.found and not r.stats["out_of_memory"] and r.length == case["upper"]
    if not r.bound_word:
        _check_word(case, r.length, r.word)


@pytest.mark.parametrize("case", _budget_cases()[:3])
def test_memory_budget_without_deepening_gives_a_lower_bound(ctx, case):
    """With the deepening off (dfs: false) or a strict budget too small for
    its smallest parts, the search ends at the budget with a lower bound on
    the length ("mlsw >= x", exact.cpp:75-96 of the reference), never an
    error."""
    import synchrowords_amd as sw
    n, k, d = case["n"], case["k"], case["delta"]
    full = ctx.exact(n, k, d, upper=case["beam"])
    mb = max(1, _search_mb(full, n) // 8)
    for cfg in (sw.ExactConfig(max_memory_mb=mb, dfs=False),
                sw.ExactConfig(max_memory_mb=1, strict_memory_limit=True, dfs_min_list_size=1 << 26)):
        r = ctx.exact(n, k, d, upper=case["beam"], cfg=cfg)
        assert not r.found and r.stats["out_of_memory"] == 1
        assert 1 <= r.length <= case["upper"]


@pytest.mark.parametrize("case", [pytest.param(c, id=c["name"]) for c in golden("reduce_plan.json")])
def test_pipeline_with_the_reduce_plan(ctx, case):
    """configs/exact_reduce.json (Reduce before Exact): the device Reduce
    takes the reference's number of forward steps and reduces to its number
    of states, and the Exact search on the reduced automaton from the
    Reduce's list ends at the bounds the reference printed after its second
    build for the reduced N (tests/golden/reduce_plan.json), with a word
    (the Reduce's prefix, then the reduced search's letters) that replays on
    the original automaton."""
    import json
    from cli_helpers import EXACT_REDUCE_PLAN
    n, k = case["n"], case["k"]
    r = ctx.pipeline(n, k, case["delta"], plan=json.dumps(EXACT_REDUCE_PLAN))
    assert r.non_synchro == bool(case["non_synchro"])
    assert (r.reduced_n, r.reduce_steps) == (case["reduced_n"], case["reduce_bfs_steps"])
    assert (r.lower, r.upper) == (case["lower"], case["upper"])
    assert "Reduce" in [name for name, _ in r.times_ms]
    _check_word(case, r.upper, r.word)


def _reduce_plan(threshold: str) -> str:
    import json
    from cli_helpers import EXACT_REDUCE_PLAN
    plan = json.loads(json.dumps(EXACT_REDUCE_PLAN))
    for a in plan["algorithms"]:
        if a["name"] == "Reduce":
            a["config"]["list_size_threshold"] = threshold
    return json.dumps(plan)


@pytest.mark.parametrize("n,k,seed", [(n, k, s) for n, k in ((80, 2), (96, 3), (130, 2), (160, 3), (220, 2))
                                      for s in range(3)])
def test_reduce_matches_oracle(ctx, n, k, seed):
    """Random automata through the Reduce plan: the device Reduce's steps and
    reduced number of states equal the oracle's (reduce.hpp restated, itself
    checked against the reference's golden runs) from the same Beam bound,
    and the final length equals the one of the plan without Reduce."""
    import oracle_lib as orc
    d = gen.random_dfa(n, k, 7000 + seed)
    r = ctx.pipeline(n, k, d, plan=_reduce_plan("AUT_N * AUT_K * 8"))
    plain = ctx.pipeline(n, k, d)
    assert r.beam_bound == plain.beam_bound
    steps, decided, length, rn = orc.reduce(n, k, d, plain.beam_bound, n * k * 8)
    if decided:
        assert (r.lower, r.upper) == (length, length) and r.reduced_n == 0
    else:
        assert (r.reduced_n, r.reduce_steps) == (rn, steps if rn else 0)
    assert (r.lower, r.upper) == (plain.lower, plain.upper)
    assert r.word is not None and len(r.word) == r.upper and replay(n, k, d, r.word) == 1


@pytest.mark.parametrize("case", [c for c in _cases("random_misc.json") if c.values[0]["n"] >= 80])
def test_reduce_settles_the_length(ctx, case):
    """A plan without Beam (so the upper bound is Eppstein's, above the
    length) and a list threshold no level reaches: the Reduce's own forward
    search runs until a level holds a singleton, which settles the length
    (reduce.hpp:55-67).  The reference's length, the oracle's Reduce agrees,
    and the word is the Reduce's own (a path to that singleton)."""
    import json
    import oracle_lib as orc
    n, k = case["n"], case["k"]
    plan = json.loads(_reduce_plan("AUT_N * 100000"))
    plan["algorithms"] = [a for a in plan["algorithms"] if a["name"] != "Beam"]
    r = ctx.pipeline(n, k, case["delta"], plan=json.dumps(plan))
    assert r.eppstein_bound > case["upper"]
    assert (r.lower, r.upper) == (case["upper"], case["upper"])
    assert r.reduced_n == 0 and r.exact["found"] == 0
    _check_word(case, r.upper, r.word)
    _, decided, length, _ = orc.reduce(n, k, case["delta"], r.eppstein_bound, n * 100000)
    assert decided and length == case["upper"]



def _images(n, k, d, sets):
    """Distinct images of `sets` under every letter (oracle apply)."""
    import numpy as np
    import oracle_lib as orc
    out = orc.apply(n, k, d, sets)
    return np.unique(out, axis=0)


@pytest.mark.parametrize("case", _cases("c1_random48.json")[:6] + _cases("c3_random200.json")[:3])
@pytest.mark.parametrize("depth", [3, 7])
def test_exact_from_a_list(ctx, case, depth):
    """The search from a list of sets (sw_exact_solve_from_list) on the
    images of Q under all words of one length d: any reset word passes
    through one of them, so the shortest word from the list has the
    reference's length minus d; the returned word takes the set the search
    names to a singleton."""
    import numpy as np
    import oracle_lib as orc
    n, k, d = case["n"], case["k"], case["delta"]
    if depth >= case["upper"]:
        pytest.skip("the length is below the depth")
    W = (n + 63) // 64
    full = np.zeros((1, W), dtype=np.uint64)
    for q in range(n):
        full[0, q // 64] |= np.uint64(1) << np.uint64(q % 64)
    lst = full
    for _ in range(depth):
        lst = _images(n, k, d, lst)
    # a repeat of the first set: counted once
    lst = np.concatenate([lst, lst[:1]])
    r, idx = ctx.exact_from_list(n, k, d, lst, upper=case["upper"] - depth + 1)
    assert r.found and r.length == case["upper"] - depth
    if r.word is not None:
        assert len(r.word) == r.length and 0 <= idx < len(lst)
        cur = lst[idx:idx + 1]
        for a in r.word:
            cur = orc.apply(n, k, d, cur)[a:a + 1]
        assert int(sum(bin(int(x)).count("1") for x in cur[0])) == 1



@pytest.mark.parametrize("n,k", [(80, 4), (130, 3), (260, 2), (330, 2), (400, 2), (512, 2), (512, 3)])
def test_reduce_stage_across_word_counts(ctx, n, k):
    """The Reduce stage alone (a plan of Eppstein, Beam and Reduce, no
    Exact) on random automata of 2 to 8 words per set and 2 to 4 letters:
    the device's steps and reduced number of states equal the oracle's from
    the same Beam bound."""
    import json
    import oracle_lib as orc
    d = gen.random_dfa(n, k, 7300 + n + k)
    plan = {"upper_bound": "1ULL * (AUT_N - 1) * (AUT_N - 1) + 1", "algorithms": [
        {"name": "Eppstein", "config": {}},
        {"name": "Beam", "config": {"presort": "indeg", "dynamic": True, "min_beam_size": "AUT_N * std::log2(AUT_N)",
                                    "max_beam_size": "AUT_N * AUT_N * std::log2(AUT_N)", "beam_exact_ratio": 0.01}},
        {"name": "Reduce", "config": {"min_n": "80", "list_size_threshold": "AUT_N * AUT_K * 8"}}]}
    r = ctx.pipeline(n, k, d, plan=json.dumps(plan))
    steps, decided, length, rn = orc.reduce(n, k, d, r.beam_bound, n * k * 8)
    if decided:
        assert (r.lower, r.upper) == (length, length) and r.reduced_n == 0
    else:
        assert (r.reduced_n, r.reduce_steps) == (rn, steps if rn else 0)
        assert r.lower < r.upper == r.beam_bound
