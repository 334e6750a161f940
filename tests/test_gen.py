"""The seeded generator (CPU)."""
from synchrowords_amd import gen


def test_splitmix64_reference_vector():
    # splitmix64 from state 0: the published first outputs.
    st, a = gen.splitmix64(0)
    st, b = gen.splitmix64(st)
    assert a == 0xE220A8397B1DCDAF
    assert b == 0x6E789E6AA1B965F4


def test_random_dfa_deterministic_and_in_range():
    d1 = gen.random_dfa(300, 2, 1)
    d2 = gen.random_dfa(300, 2, 1)
    assert d1 == d2 and len(d1) == 600 and all(0 <= x < 300 for x in d1)
    assert gen.random_dfa(300, 2, 2) != d1


def test_bounded_is_uniform_enough():
    r = gen.Rng(7)
    counts = [0] * 5
    for _ in range(50000):
        counts[r.bounded(5)] += 1
    assert all(abs(c - 10000) < 400 for c in counts)


def test_batch_specs_ranges():
    specs = gen.batch_specs(500, 42)
    assert all(16 <= n <= 160 and k in (2, 3) for n, k, _ in specs)
    assert len({n for n, _, _ in specs}) > 50


def test_cerny_definition():
    d = gen.cerny(5)
    assert [d[2 * q] for q in range(5)] == [1, 2, 3, 4, 0]
    assert [d[2 * q + 1] for q in range(5)] == [1, 1, 2, 3, 4]


def test_native_generator_matches_python():
    import numpy as np
    import synchrowords_amd as sw
    for n, k, seed in [(48, 2, 0), (300, 2, 1), (160, 3, 2**63 + 5), (17, 4, 99)]:
        assert np.array_equal(sw.random_dfa(n, k, seed), np.array(gen.random_dfa(n, k, seed), dtype=np.uint32))
