"""ctypes access to the oracle (oracle/liboracle.so) — test infrastructure.

The oracle is the checker of the HIP path, never part of it.  Builds
oracle/liboracle.so with `make -C oracle` when it is missing (g++ only).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "oracle" / "liboracle.so"
_lib = None


class Result(C.Structure):
    _fields_ = [("non_synchro", C.c_int32), ("exact_found", C.c_int32), ("entered_dfs", C.c_int32),
                ("pad", C.c_int32), ("lo", C.c_uint64), ("hi", C.c_uint64), ("eppstein", C.c_uint64),
                ("beam", C.c_uint64)]


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists() or LIB.stat().st_mtime < (ROOT / "oracle" / "oracle.cpp").stat().st_mtime:
            subprocess.run(["make", "-C", str(ROOT / "oracle")], check=True, capture_output=True)
        _lib = C.CDLL(str(LIB))
        _lib.oracle_solve.restype = C.c_int
        _lib.oracle_brute.restype = C.c_long
        _lib.oracle_beam.restype = C.c_uint64
        _lib.oracle_beam_width.restype = C.c_uint64
        _lib.oracle_replay.restype = C.c_uint32
        _lib.oracle_subset_count.restype = C.c_uint64
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None and a.size else None


def solve(n, k, delta, run_exact=True, trace=False, max_memory_mb=0):
    d = np.ascontiguousarray(delta, dtype=np.uint32)
    r = Result()
    cap = 100000 if trace else 0
    ph = np.zeros(cap, dtype=np.int32)
    ln = np.zeros(cap, dtype=np.uint64)
    f = np.zeros(cap, dtype=np.uint64)
    b = np.zeros(cap, dtype=np.uint64)
    tl = C.c_uint64()
    rc = lib().oracle_solve(n, k, _p(d), int(run_exact), C.c_uint64(max_memory_mb), C.byref(r), _p(ph), _p(ln),
                            _p(f), _p(b), C.c_uint64(cap), C.byref(tl))
    assert rc == 0, rc
    out = {"non_synchro": r.non_synchro, "lower": r.lo, "upper": r.hi, "eppstein": r.eppstein, "beam": r.beam,
           "exact_found": r.exact_found, "entered_dfs": r.entered_dfs}
    if trace:
        m = min(tl.value, cap)
        out["trace"] = [[["BFS", "IBFS", "DFS"][ph[i]], int(ln[i]), int(f[i]), int(b[i])] for i in range(m)]
    return out


def reduce(n, k, delta, upper, threshold):
    """The Reduce stage: (steps, decided, length, reduced_n); reduced_n 0 when
    the automaton was not reduced."""
    d = np.ascontiguousarray(delta, dtype=np.uint32)
    steps, dec, length, rn = C.c_uint64(), C.c_int32(), C.c_uint64(), C.c_uint32()
    rc = lib().oracle_reduce(n, k, _p(d), C.c_uint64(upper), C.c_uint64(threshold), C.byref(steps), C.byref(dec),
                             C.byref(length), C.byref(rn))
    assert rc == 0, rc
    return int(steps.value), bool(dec.value), int(length.value), int(rn.value)


def brute(n, k, delta) -> int:
    d = np.ascontiguousarray(delta, dtype=np.uint32)
    return int(lib().oracle_brute(n, k, _p(d)))


def eppstein(n, k, delta, upper):
    d = np.ascontiguousarray(delta, dtype=np.uint32)
    ns = C.c_int32()
    bound = C.c_uint64()
    w = np.zeros(max(upper, 1) if upper < (1 << 22) else (1 << 22), dtype=np.uint32)
    wl = C.c_uint64()
    lib().oracle_eppstein(n, k, _p(d), C.c_uint64(upper), C.byref(ns), C.byref(bound), _p(w), C.c_uint64(w.size),
                          C.byref(wl))
    return bool(ns.value), int(bound.value), [int(x) for x in w[: wl.value]]


def beam(n, k, delta, upper) -> int:
    d = np.ascontiguousarray(delta, dtype=np.uint32)
    return int(lib().oracle_beam(n, k, _p(d), C.c_uint64(upper)))


def beam_width(n, k, delta, width, max_len) -> int:
    """One fixed-width Beam run on the automaton as given (no presort)."""
    d = np.ascontiguousarray(delta, dtype=np.uint32)
    return int(lib().oracle_beam_width(n, k, _p(d), C.c_uint32(width), C.c_uint64(max_len)))


def replay(n, k, delta, word) -> int:
    d = np.ascontiguousarray(delta, dtype=np.uint32)
    w = np.ascontiguousarray(word, dtype=np.uint32)
    return int(lib().oracle_replay(n, k, _p(d), _p(w), C.c_uint64(w.size)))


def apply(n, k, delta, sets, backward=False):
    d = np.ascontiguousarray(delta, dtype=np.uint32)
    s = np.ascontiguousarray(sets, dtype=np.uint64)
    cnt = s.shape[0]
    out = np.zeros((k * cnt, (n + 63) // 64), dtype=np.uint64)
    lib().oracle_apply(n, k, _p(d), int(backward), _p(s), C.c_uint64(cnt), _p(out))
    return out


def subset_count(small, large, W, proper=False, thresh=1):
    s = np.ascontiguousarray(small, dtype=np.uint64)
    lg = np.ascontiguousarray(large, dtype=np.uint64)
    ns = s.shape[0]
    cnt = np.zeros(ns, dtype=np.uint32)
    flags = np.zeros(ns, dtype=np.uint8)
    total = lib().oracle_subset_count(_p(s), C.c_uint64(ns), _p(lg), C.c_uint64(lg.shape[0]), W, int(proper),
                                      thresh, _p(cnt), _p(flags))
    return int(total), cnt, flags.astype(bool)


def count_bits(sets, n):
    s = np.ascontiguousarray(sets, dtype=np.uint64)
    acc = np.zeros(n, dtype=np.uint64)
    lib().oracle_count_bits(_p(s), C.c_uint64(s.shape[0]), n, _p(acc))
    return acc


def subset_count_mt(small, large, threads=0, per_row=False):
    """All-pairs count over host threads (same semantics as subset_count, not
    proper): the total, and per-row counts when asked."""
    s = np.ascontiguousarray(small, dtype=np.uint64)
    lg = np.ascontiguousarray(large, dtype=np.uint64)
    W = s.shape[1]
    cnt = np.zeros(s.shape[0], dtype=np.uint32) if per_row else None
    f = lib().oracle_subset_count_mt
    f.restype = C.c_uint64
    total = f(_p(s), C.c_uint64(s.shape[0]), _p(lg), C.c_uint64(lg.shape[0]), W, threads, _p(cnt))
    return (int(total), cnt) if per_row else int(total)
