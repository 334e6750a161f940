"""Random test data (numpy) shared by the tests."""
from __future__ import annotations

import numpy as np


def words(n: int) -> int:
    return (n + 63) // 64


def random_sets(rng: np.random.Generator, count: int, n: int, p: float) -> np.ndarray:
    """count sets over n states, each state present with probability p."""
    W = words(n)
    bits = rng.random((count, W * 64)) < p
    bits[:, n:] = False
    out = np.zeros((count, W), dtype=np.uint64)
    for w in range(W):
        chunk = bits[:, w * 64:(w + 1) * 64].astype(np.uint64)
        out[:, w] = (chunk << np.arange(64, dtype=np.uint64)).sum(axis=1, dtype=np.uint64)
    return out


def slice_table(n: int, k: int, delta, letter: int, backward: bool) -> np.ndarray:
    """The reference's PreprocessedTransition table for one letter, built
    directly from its definition (preprocessed_transitions.cpp:19-53):
    [ceil(n/8)][256][W] words."""
    W = words(n)
    S = (n + 7) // 8
    d = np.asarray(delta, dtype=np.int64).reshape(n, k)
    t = np.zeros((S, 256, W), dtype=np.uint64)
    for s in range(S):
        for b in range(256):
            row = np.zeros(W, dtype=np.uint64)
            for j in range(8):
                q = s * 8 + j
                if q >= n or not (b >> j) & 1:
                    continue
                targets = np.nonzero(d[:, letter] == q)[0] if backward else [d[q, letter]]
                for x in targets:
                    row[x // 64] |= np.uint64(1) << np.uint64(x % 64)
            t[s, b] = row
    return t
