"""The subset-check microbench's golden count (tests/golden/c2_count.json,
written by tests/golden/make_c2_count.py with the oracle's all-pairs count):
the generator still makes the same lists, and the oracle's count of the
first rows equals the fixture's."""
import json

import numpy as np
import pytest

import oracle_lib as orc
from conftest import GOLDEN

import bench
from golden.make_c2_count import checksum


@pytest.fixture(scope="module")
def c2():
    g = json.loads((GOLDEN / "c2_count.json").read_text())
    small, large, planted = bench.make_lists(g["count"], g["n"], g["p_small"], g["p_large"], seed=g["seed"])
    return g, small, large, planted


def test_generator_makes_the_fixture_lists(c2):
    g, small, large, planted = c2
    assert checksum(small) == g["checksum_small"]
    assert checksum(large) == g["checksum_large"]
    assert [list(p) for p in planted] == g["planted"]


def test_oracle_count_of_first_rows(c2):
    g, small, large, _ = c2
    assert orc.subset_count_mt(small[:4096], large) == g["first_4096_rows"]
    assert sum(g["block_counts"]) == g["pair_count"]
    # About 2^40 * 0.98^256 pairs, the expectation for these densities.
    expect = float(g["count"]) ** 2 * (1 - 0.08 * 0.25) ** g["n"]
    assert abs(g["pair_count"] / expect - 1) < 0.01

