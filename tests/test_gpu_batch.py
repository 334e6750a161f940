"""Batched exact search (one persistent launch) against the reference."""
import pytest

from conftest import golden
from synchrowords_amd import gen, replay

pytestmark = pytest.mark.gpu


def _instances(cases):
    return [(c["n"], c["k"], c["delta"]) for c in cases]


def test_batch_small_fixtures(ctx):
    cases = golden("c1_random48.json") + golden("random_misc.json") + golden("hard.json") + \
        golden("slowly_sink.json") + golden("readme.json")
    res, st = ctx.batch(_instances(cases))
    for c, r in zip(cases, res):
        assert r["non_synchro"] == bool(c["non_synchro"]), c["name"]
        if c["non_synchro"]:
            continue
        assert r["length"] == c["upper"], c["name"]
        if c["n"] > 1:
            assert r["word"] is not None and len(r["word"]) == r["length"], c["name"]
            assert replay(c["n"], c["k"], c["delta"], r["word"]) == 1, c["name"]


def test_batch_overflow_goes_to_device_search(ctx):
    """Tiny workgroup scratch forces every instance through the overflow path."""
    cases = golden("c1_random48.json")[:10]
    res, st = ctx.batch(_instances(cases), cap_rows=64, cap_cand=64)
    assert st["overflowed"] > 0
    for c, r in zip(cases, res):
        assert r["length"] == c["upper"]
        assert replay(c["n"], c["k"], c["delta"], r["word"]) == 1


def test_batch_c4_full(ctx):
    """All 4096 instances of the batch workload (n in [16,160], k in {2,3}) in
    one call: every NON SYNCHRO flag and length equals the reference's, and
    every word replays to a singleton."""
    cases = golden("c4_batch.json")
    assert len(cases) == 4096
    inst = [(c["n"], c["k"], gen.random_dfa(c["n"], c["k"], c["seed"])) for c in cases]
    res, st = ctx.batch(inst)
    for c, (n, k, d), r in zip(cases, inst, res):
        assert r["non_synchro"] == bool(c["non_synchro"])
        if not c["non_synchro"]:
            assert r["length"] == c["upper"], c
            assert replay(n, k, d, r["word"]) == 1
