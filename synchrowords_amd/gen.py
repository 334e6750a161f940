"""Seeded synthetic inputs: random DFAs, Cerny automata, batches, subset lists.

The generator is the project's own (the reference ships a numpy-based script,
scripts/generate.py:19-52, whose PCG64 stream is not reproduced).  The stream
is fixed so that every input is reproducible from its seed, in Python here and
in C++ in ``csrc/sw_gen.hpp``:

* ``splitmix64`` (Steele, Lea, Flood 2014) expands a 64-bit seed;
* ``xoshiro256**`` (Blackman, Vigna 2018) seeded with four splitmix64 outputs;
* unbiased bounded draws by Lemire's multiply-and-reject method (2019).

A random DFA draws ``delta(q, a)`` in q-major, a-minor order, the order of the
reference input format (docs/install.md:47-50, app/io.hpp:55-78).
"""
from __future__ import annotations

MASK = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15


def splitmix64(state: int) -> tuple[int, int]:
    """One splitmix64 step: returns (new_state, output)."""
    state = (state + GOLDEN) & MASK
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return state, z ^ (z >> 31)


def _rotl(x: int, k: int) -> int:
    return ((x << k) | (x >> (64 - k))) & MASK


class Rng:
    """xoshiro256** with Lemire bounded draws."""

    def __init__(self, seed: int):
        st = seed & MASK
        s = []
        for _ in range(4):
            st, out = splitmix64(st)
            s.append(out)
        self.s = s

    def next(self) -> int:
        s = self.s
        result = (_rotl((s[1] * 5) & MASK, 7) * 9) & MASK
        t = (s[1] << 17) & MASK
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = _rotl(s[3], 45)
        return result

    def bounded(self, n: int) -> int:
        """Uniform integer in [0, n)."""
        m = self.next() * n
        low = m & MASK
        if low < n:
            t = ((1 << 64) - n) % n
            while low < t:
                m = self.next() * n
                low = m & MASK
        return m >> 64


def instance_seed(base: int, index: int) -> int:
    """Seed of batch instance ``index`` derived from a batch seed."""
    _, out = splitmix64((base ^ ((index * 0xD1B54A32D192ED03) & MASK)) & MASK)
    return out


def random_dfa(n: int, k: int, seed: int) -> list[int]:
    """delta as a flat q-major list: delta[q*k + a] uniform in [0, n)."""
    rng = Rng(seed)
    return [rng.bounded(n) for _ in range(n * k)]


def cerny(n: int) -> list[int]:
    """Cerny automaton C_n: a: q -> q+1 mod n; b fixes every state but 0 -> 1.

    Same automaton as the reference's ``generate_cerny(0)``
    (automaton/automaton.hpp:84-95); its shortest reset word has length
    (n-1)^2.
    """
    d = []
    for q in range(n):
        d += [(q + 1) % n, q]
    d[0 * 2 + 1] = 1
    return d


def hard(n: int) -> list[int]:
    """The reference's ``generate_hard`` family (automaton/automaton.hpp:102-115)."""
    d = [0] * (2 * n)
    for i in range(n - 1):
        d[2 * i] = d[2 * i + 1] = i + 1
    d[2 * (n - 1)] = 0
    d[2 * (n - 1) + 1] = 1
    d[2 * 1 + 1] = 0
    d[2 * 0 + 1] = 3
    d[2 * 2 + 1] = 3
    return d


def slowly_sink(n: int) -> tuple[int, list[int]]:
    """Slowly synchronizing family with k = n-1 letters and reset length n(n-1)/2.

    The family of scripts/generate.py:35-52: letter 0 merges state 1 into
    state 0 and fixes the rest; letter j >= 1 swaps states j and j+1 and fixes
    the rest.  Returns (k, delta).
    """
    k = n - 1
    d = [0] * (n * k)
    for j in range(k):
        for q in range(n):
            d[q * k + j] = q
        if j == 0:
            d[1 * k + 0] = 0
        else:
            d[j * k + j] = j + 1
            d[(j + 1) * k + j] = j
    return k, d


def batch_specs(count: int, base: int, n_lo: int = 16, n_hi: int = 160):
    """(n, k, seed) of each batch instance: n uniform in [n_lo, n_hi], k in {2, 3}."""
    out = []
    for i in range(count):
        s = instance_seed(base, i)
        rng = Rng(s)
        n = n_lo + rng.bounded(n_hi - n_lo + 1)
        k = 2 + rng.bounded(2)
        out.append((n, k, rng.next()))
    return out


def to_text(n: int, k: int, delta: list[int]) -> str:
    """Reference input format: "K N" then the N*K transitions."""
    return f"{k} {n}\n" + " ".join(str(x) for x in delta) + "\n"
