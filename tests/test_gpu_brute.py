"""Power-set BFS on the device (the reference's Brute, brute.hpp:28-122)
against the reference's own lengths (golden files written by the reference
binary), closed forms and the oracle's BFS."""
import numpy as np
import pytest

import oracle_lib as orc
from conftest import golden
from synchrowords_amd import gen

pytestmark = pytest.mark.gpu


def _cases(name, max_n):
    return [c for c in golden(name) if c["n"] <= max_n]


@pytest.mark.parametrize("case", _cases("hard.json", 20) + _cases("cerny.json", 16) + _cases("slowly_sink.json", 32)
                         + _cases("readme.json", 32) + _cases("random_misc.json", 26), ids=lambda c: c["name"])
def test_brute_matches_reference_lengths(ctx, case):
    n, k = case["n"], case["k"]
    got = ctx.brute(n, k, case["delta"])
    if case["non_synchro"]:
        assert got == -1
    else:
        assert got == case["upper"]


@pytest.mark.parametrize("n", [2, 3, 5, 9, 13, 17, 20])
def test_cerny_closed_form(ctx, n):
    """C_n resets in (n-1)^2 letters; above 16 states past the reference's 8-bit distances."""
    assert ctx.brute(n, 2, gen.cerny(n)) == (n - 1) ** 2


def test_non_synchronizing(ctx):
    # Two cycles that never merge: a permutation automaton.
    n, k = 12, 2
    d = [0] * (n * k)
    for q in range(n):
        d[q * k + 0] = (q + 1) % n
        d[q * k + 1] = (q + 5) % n
    assert ctx.brute(n, k, d) == -1
    # One letter merges only within halves.
    d2 = [0] * (n * k)
    for q in range(n):
        d2[q * k + 0] = (q // 6) * 6
        d2[q * k + 1] = (q // 6) * 6 + (q + 1) % 6
    assert ctx.brute(n, k, d2) == orc.brute(n, k, d2) == -1


@pytest.mark.parametrize("n,k", [(6, 2), (11, 3), (15, 2), (18, 4), (20, 2), (21, 2), (23, 3), (24, 2)])
def test_random_against_oracle(ctx, n, k):
    """Both the one-workgroup kernel (bitmap in LDS) and the per-level kernel (n > 20)."""
    for seed in range(3):
        d = gen.random_dfa(n, k, 900 + 31 * n + seed)
        assert ctx.brute(n, k, d) == orc.brute(n, k, d), seed


def test_batch_mixed_sizes(ctx):
    rng = np.random.default_rng(5)
    inst = []
    for j in range(300):
        n = int(rng.integers(1, 19))
        k = int(rng.integers(1, 5))
        inst.append((n, k, gen.random_dfa(n, k, 7700 + j)))
    inst.append((22, 2, gen.random_dfa(22, 2, 1)))  # per-level path inside a batch
    inst.append((10, 2, gen.cerny(10)))
    got = ctx.brute_batch(inst)
    want = [orc.brute(n, k, d) for n, k, d in inst]
    np.testing.assert_array_equal(got, want)
    assert len(ctx.brute_batch([])) == 0
