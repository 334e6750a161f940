"""synchrowords_amd — MI355X-native shortest-reset-word search (SynchroLib's Exact path).

Python mirror of the C ABI in ``include/synchrowords_amd.h`` (ctypes).  The
HIP library ``libsynchrowords_amd.so`` is built in-tree by
``python -m synchrowords_amd.build``; importing this package does not need it,
but every call that computes does, and fails loudly when it is missing: there
is no CPU path behind these functions.

Sets are numpy ``uint64`` arrays of shape ``(count, W)``, ``W = ceil(n/64)``,
state q at bit q%64 of word q/64 — the reference's ``Subset<N>`` layout.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

__all__ = [
    "Context", "ExactConfig", "ExactResult", "PipelineResult", "Leaf", "load", "words", "DeviceArray",
]

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libsynchrowords_amd.so"

SW_HOST = 0
SW_DEVICE = 1

_lib = None


class SwError(RuntimeError):
    """A failed call into the HIP library (message from sw_last_error)."""


def words(n: int) -> int:
    return (n + 63) // 64


class _ExactConfig(C.Structure):
    _fields_ = [("max_memory_mb", C.c_uint64), ("dominance_pairs", C.c_uint64), ("forward_cap", C.c_uint64),
                ("dfs", C.c_int32), ("verbose", C.c_int32), ("dfs_min_list_size", C.c_uint64),
                ("strict_memory_limit", C.c_int32), ("goal_check", C.c_int32), ("side_rule", C.c_int32)]


class _ExactResult(C.Structure):
    _fields_ = [("found", C.c_int32), ("bound_word", C.c_int32), ("length", C.c_uint64), ("word_len", C.c_uint64),
                ("steps_forward", C.c_uint64), ("steps_backward", C.c_uint64), ("sets_forward", C.c_uint64),
                ("sets_backward", C.c_uint64), ("pair_tests", C.c_uint64), ("hash_probes", C.c_uint64),
                ("ms_total", C.c_double), ("ms_expand", C.c_double), ("ms_dominance", C.c_double),
                ("ms_goal", C.c_double), ("dfs_entered", C.c_int32), ("out_of_memory", C.c_int32),
                ("dfs_calls", C.c_uint64), ("dfs_sets", C.c_uint64), ("trie_builds", C.c_uint64),
                ("trie_queries", C.c_uint64), ("ms_dfs", C.c_double)]


class _Result(C.Structure):
    _fields_ = [("non_synchro", C.c_int32), ("n_algorithms", C.c_int32), ("lower", C.c_uint64),
                ("upper", C.c_uint64), ("word_len", C.c_uint64), ("eppstein_bound", C.c_uint64),
                ("beam_bound", C.c_uint64), ("ms", C.c_double * 8), ("names", (C.c_char * 16) * 8),
                ("exact", _ExactResult), ("reduced_n", C.c_uint32), ("reduce_steps", C.c_uint32)]


class _BatchConfig(C.Structure):
    _fields_ = [("grid", C.c_uint32), ("cap_rows", C.c_uint32), ("cap_cand", C.c_uint32),
                ("overflow_workers", C.c_uint32), ("max_pairs", C.c_uint64)]


class _FrontierStats(C.Structure):
    _fields_ = [("candidates", C.c_uint64), ("new_sets", C.c_uint64), ("probes", C.c_uint64),
                ("table_slots", C.c_uint64), ("ms_expand", C.c_double), ("ms_insert", C.c_double),
                ("ms_scan", C.c_double), ("ms_commit", C.c_double)]


class _BatchStats(C.Structure):
    _fields_ = [("ms_host_prep", C.c_double), ("ms_kernel", C.c_double), ("ms_overflow", C.c_double),
                ("ms_total", C.c_double), ("non_synchro", C.c_uint32), ("found", C.c_uint32),
                ("bound", C.c_uint32), ("overflowed", C.c_uint32), ("grid", C.c_uint32), ("pad", C.c_uint32),
                ("scratch_bytes", C.c_uint64), ("over_cand", C.c_uint32), ("over_rows", C.c_uint32),
                ("over_word", C.c_uint32), ("over_hash", C.c_uint32), ("over_pairs", C.c_uint32),
                ("pad2", C.c_uint32), ("ms_overflow_search_sum", C.c_double)]


_P = C.c_void_p
_U64P = C.POINTER(C.c_uint64)
_U32P = C.POINTER(C.c_uint32)

_SIGS = {
    "sw_last_error": (C.c_char_p, []),
    "sw_version": (C.c_int, []),
    "sw_gen_random_dfa": (None, [C.c_uint32, C.c_uint32, C.c_uint64, _P]),
    "sw_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "sw_ctx_destroy": (None, [_P]),
    "sw_device_info": (C.c_int, [_P, C.c_char_p, C.c_size_t, C.POINTER(C.c_int), _U64P]),
    "sw_dev_alloc": (C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    "sw_dev_free": (C.c_int, [_P, _P]),
    "sw_memcpy_h2d": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "sw_memcpy_d2h": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "sw_sync": (C.c_int, [_P]),
    "sw_ctx_reset": (C.c_int, [_P]),
    "sw_stream_timer": (C.c_int, [_P, C.c_int, C.POINTER(C.c_float)]),
    "sw_dev_copy": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "sw_transition_apply": (C.c_int, [_P, C.c_uint32, _P, _P, _P, C.c_uint64, C.c_int]),
    "sw_expand": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, C.c_int, _P, C.c_uint64, _P, C.c_int]),
    "sw_frontier_step": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, C.c_int, _P, C.c_uint64, _P, C.c_uint64, _P, _P, _P,
                                   _U64P, C.c_int, _P]),
    "sw_leaf_create": (C.c_int, [_P, C.c_uint32, _P, C.c_uint32, C.c_uint32, C.POINTER(_P)]),
    "sw_leaf_run": (C.c_int, [_P, C.c_int, C.c_uint32, C.c_uint32, _P, C.c_uint32, _P]),
    "sw_leaf_destroy": (None, [_P]),
    "sw_subset_check": (C.c_int, [_P, C.c_uint32, _P, C.c_uint64, _P, C.c_uint64, C.c_uint32, C.c_int, _P, _U64P,
                                  C.POINTER(C.c_int32)]),
    "sw_subset_check_noalloc": (C.c_int, [_P, C.c_uint32, _P, C.c_uint64, _P, C.c_uint64, C.c_uint32, _P, _P,
                                          _P]),
    "sw_count_bits": (C.c_int, [_P, C.c_uint32, _P, C.c_uint64, _P, C.c_int]),
    "sw_trie_build": (C.c_int, [_P, C.c_uint32, _P, C.c_uint64, C.c_int, C.POINTER(_P)]),
    "sw_trie_query": (C.c_int, [_P, C.c_int, _P, C.c_uint64, C.c_int, _P, C.POINTER(C.c_int32), _U64P]),
    "sw_trie_info": (C.c_int, [_P, _U64P, _U64P, _U32P]),
    "sw_trie_nodes": (C.c_int, [_P, _P, C.c_uint64, _P, _P]),
    "sw_trie_destroy": (None, [_P]),
    "sw_exact_solve": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, C.c_uint64, C.c_uint64, _P, C.c_uint64,
                                 C.POINTER(_ExactConfig), C.POINTER(_ExactResult), _P, C.c_uint64]),
    "sw_exact_solve_from_list": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, _P, C.c_uint64, C.c_uint64, C.c_uint64,
                                           C.POINTER(_ExactConfig), C.POINTER(_ExactResult), _P, C.c_uint64,
                                           C.POINTER(C.c_int64)]),
    "sw_pipeline_solve": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, C.c_char_p, C.POINTER(_Result), _P,
                                    C.c_uint64]),
    "sw_pipeline_solve_result_word": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, C.c_char_p, C.POINTER(_Result), _P,
                                                C.c_uint64, _P, C.c_uint64, _U64P]),
    "sw_batch_solve": (C.c_int, [_P, C.c_uint32, _P, _P, _P, _P, C.POINTER(_BatchConfig), _P, _P, _P, _P,
                                 C.c_uint32, C.POINTER(_BatchStats)]),
    "sw_mlp_create": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.POINTER(_P)]),
    "sw_mlp_destroy": (None, [_P]),
    "sw_mlp_shape": (C.c_int, [_P, _U32P, _U32P, _U64P]),
    "sw_mlp_get_params": (C.c_int, [_P, _P, C.c_uint64]),
    "sw_mlp_set_params": (C.c_int, [_P, _P, C.c_uint64]),
    "sw_mlp_features": (C.c_int, [_P, _P, _P, C.c_uint64, _P]),
    "sw_mlp_forward": (C.c_int, [_P, _P, C.c_uint64, _P]),
    "sw_mlp_gradients": (C.c_int, [_P, _P, _P, C.c_uint64, _P, C.POINTER(C.c_float)]),
    "sw_mlp_set_loss": (C.c_int, [_P, C.c_uint32]),
    "sw_mlp_train": (C.c_int, [_P, _P, _P, C.c_uint64, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64, _P]),
    "sw_mlp_samples": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                 _P, _P, C.c_uint64, _U64P]),
    "sw_beam_device": (C.c_int, [_P, _P, C.c_uint32, C.c_uint32, _P, C.c_uint32, C.c_uint64, _U64P, _P, C.c_uint64,
                                 _U64P]),
    "sw_autotune_subset": (C.c_int, [_P, C.c_uint32, C.c_uint64, C.c_uint64, C.c_float, C.c_float, C.c_uint32, _U32P,
                                     _U32P, _U32P, C.POINTER(C.c_float)]),
    "sw_set_subset_kernel": (C.c_int, [_P, C.c_int, C.c_uint64]),
    "sw_set_subset_walk": (C.c_int, [_P, C.c_int]),
    "sw_brute": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, _P]),
    "sw_brute_batch": (C.c_int, [_P, C.c_uint32, _P, _P, _P, _P, _P]),
    "sw_eppstein": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, C.c_uint64, _P, _P, _P, C.c_uint64, _P]),
    "sw_eppstein_batch": (C.c_int, [_P, C.c_uint32, _P, _P, _P, _P, _P, C.c_uint32, _P, _P, _P, _P]),
    "sw_calibrate_lds": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_double)]),
    "sw_set_subset_shape": (C.c_int, [_P, C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]),
}


def load(path: str | os.PathLike | None = None):
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise SwError(f"{p} is missing: build it with `python -m synchrowords_amd.build` "
                      "(or __graft_entry__.build()); there is no CPU fallback")
    lib = C.CDLL(str(p))
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != 0:
        msg = load().sw_last_error()
        raise SwError(f"synchrowords_amd error {rc}: {msg.decode() if msg else ''}")


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p) if a is not None and a.size else None


def _sets(a, n: int) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint64)
    if a.ndim == 1:
        a = a.reshape(-1, words(n)) if a.size else a.reshape(0, words(n))
    if a.shape[1] != words(n):
        raise ValueError(f"sets must have {words(n)} words per row for n={n}")
    return a


def _delta(n: int, k: int, delta) -> np.ndarray:
    d = np.ascontiguousarray(delta, dtype=np.uint32).reshape(-1)
    if d.size != n * k:
        raise ValueError("delta must have n*k entries")
    return d


@dataclass
class ExactConfig:
    max_memory_mb: int = 0
    dominance_pairs: int = 0
    forward_cap: int = 0
    dfs: bool = True
    verbose: bool = False
    dfs_min_list_size: int = 0
    strict_memory_limit: bool = False
    goal_check: int = 0  # 0 automatic, 1 vertical subset kernel, 2 device subset trie
    side_rule: int = 0   # 0 default, 1 the smaller frontier steps, 2 steps planned by the upper bound

    def _c(self) -> _ExactConfig:
        return _ExactConfig(self.max_memory_mb, self.dominance_pairs, self.forward_cap, int(self.dfs),
                            int(self.verbose), self.dfs_min_list_size, int(self.strict_memory_limit),
                            self.goal_check, self.side_rule)


@dataclass
class ExactResult:
    found: bool
    length: int
    word: list[int] | None
    bound_word: bool
    stats: dict = field(default_factory=dict)


@dataclass
class PipelineResult:
    non_synchro: bool
    lower: int
    upper: int
    word: list[int] | None
    eppstein_bound: int
    beam_bound: int
    times_ms: list[tuple[str, float]]
    exact: dict
    reduced_n: int = 0      # states Exact searched after a Reduce stage shrank the automaton (0: none)
    reduce_steps: int = 0   # forward steps that Reduce took first


class DeviceArray:
    """A device buffer owned by a context (for inputs resident in HBM)."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _check(load().sw_dev_alloc(ctx.handle, self.nbytes, C.byref(p)))
        self.ptr = p

    @classmethod
    def from_host(cls, ctx: "Context", a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.nbytes)
        if a.nbytes:
            _check(load().sw_memcpy_h2d(ctx.handle, d.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes))
        return d

    def to_host(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        _check(load().sw_memcpy_d2h(self.ctx.handle, out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes))
        return out

    def free(self) -> None:
        if self.ptr:
            _check(load().sw_dev_free(self.ctx.handle, self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.handle:
                load().sw_dev_free(self.ctx.handle, self.ptr)
        except Exception:
            pass


class Leaf:
    """The reference's SubsetsImplicitTrieKernel (allocate / run / deallocate)."""

    def __init__(self, ctx: "Context", n: int, small, large_cnt: int):
        self.ctx = ctx
        self.n = n
        s = _sets(small, n)
        h = C.c_void_p()
        _check(load().sw_leaf_create(ctx.handle, n, _ptr(s), s.shape[0], large_cnt, C.byref(h)))
        self.handle = h

    def run(self, small_pos: int, small_cnt: int, large, proper: bool = False) -> np.ndarray:
        lg = _sets(large, self.n)
        ret = np.zeros(lg.shape[0], dtype=np.uint8)
        _check(load().sw_leaf_run(self.handle, int(proper), small_pos, small_cnt, _ptr(lg), lg.shape[0], _ptr(ret)))
        return ret.astype(bool)

    def close(self) -> None:
        if self.handle and self.ctx.handle:
            load().sw_leaf_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Trie:
    """Device subset trie over a list of sets (sw_trie_build)."""

    def __init__(self, ctx: "Context", n: int, sets):
        self.ctx, self.n = ctx, n
        s = _sets(sets, n)
        h = _P()
        _check(load().sw_trie_build(ctx.handle, n, _ptr(s), s.shape[0], SW_HOST, C.byref(h)))
        self.handle = h

    def query(self, queries, superset: bool = False, per_query: bool = True):
        """(per-query flags or None, any, witness) for host queries."""
        q = _sets(queries, self.n)
        res = np.zeros(q.shape[0], dtype=np.uint8) if per_query else None
        anyf = C.c_int32()
        wit = C.c_uint64()
        _check(load().sw_trie_query(self.handle, int(superset), _ptr(q), q.shape[0], SW_HOST,
                                    _ptr(res) if per_query and q.shape[0] else None, C.byref(anyf), C.byref(wit)))
        return (res.astype(bool) if per_query else None), bool(anyf.value), int(wit.value)

    def info(self) -> dict:
        s, nd, lv = C.c_uint64(), C.c_uint64(), C.c_uint32()
        _check(load().sw_trie_info(self.handle, C.byref(s), C.byref(nd), C.byref(lv)))
        return {"sets": s.value, "nodes": nd.value, "top_levels": lv.value}

    def dump(self):
        """(nodes [count, 4] u32, sets in trie order, their rows in the built list)."""
        inf = self.info()
        nodes = np.zeros((inf["nodes"], 4), dtype=np.uint32)
        sets = np.zeros((inf["sets"], words(self.n)), dtype=np.uint64)
        rows = np.zeros(inf["sets"], dtype=np.uint32)
        _check(load().sw_trie_nodes(self.handle, _ptr(nodes), inf["nodes"], _ptr(sets), _ptr(rows)))
        return nodes, sets, rows

    def close(self):
        if getattr(self, "handle", None):
            load().sw_trie_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Mlp:
    """The learned scorer (sw_mlp): BF16 MLP on MFMA, FP32 accumulation."""

    def __init__(self, ctx: "Context", n: int, k: int, hidden: int = 512, seed: int = 1):
        self.ctx, self.n, self.k = ctx, n, k
        h = C.c_void_p()
        _check(load().sw_mlp_create(ctx.handle, n, k, hidden, seed, C.byref(h)))
        self.handle = h
        d, hh, npar = C.c_uint32(), C.c_uint32(), C.c_uint64()
        _check(load().sw_mlp_shape(h, C.byref(d), C.byref(hh), C.byref(npar)))
        self.in_dim, self.hidden, self.nparams = d.value, hh.value, npar.value

    def close(self):
        # A network must go before its context; once the context is closed
        # its device memory went with it.
        if self.handle and self.ctx.handle:
            load().sw_mlp_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.zeros(self.nparams, dtype=np.float32)
        _check(load().sw_mlp_get_params(self.handle, _ptr(out), out.size))
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(load().sw_mlp_set_params(self.handle, _ptr(p), p.size))

    def features(self, delta, sets) -> np.ndarray:
        s = _sets(sets, self.n)
        d = _delta(self.n, self.k, delta)
        x = np.zeros((s.shape[0], self.in_dim), dtype=np.uint16)
        _check(load().sw_mlp_features(self.handle, _ptr(d), _ptr(s), s.shape[0], _ptr(x)))
        return x

    def forward(self, x) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.uint16)
        y = np.zeros(x.shape[0], dtype=np.float32)
        _check(load().sw_mlp_forward(self.handle, _ptr(x), x.shape[0], _ptr(y)))
        return y

    def set_loss(self, rank_group: int = 0) -> None:
        """0: squared error on the targets; G: pairwise ranking within groups of G samples."""
        _check(load().sw_mlp_set_loss(self.handle, rank_group))

    def gradients(self, x, target):
        x = np.ascontiguousarray(x, dtype=np.uint16)
        t = np.ascontiguousarray(target, dtype=np.float32)
        g = np.zeros(self.nparams, dtype=np.float32)
        loss = C.c_float()
        _check(load().sw_mlp_gradients(self.handle, _ptr(x), _ptr(t), x.shape[0], _ptr(g), C.byref(loss)))
        return g, float(loss.value)

    def train(self, x, target, epochs=10, batch=512, lr=1e-3, seed=0):
        x = np.ascontiguousarray(x, dtype=np.uint16)
        t = np.ascontiguousarray(target, dtype=np.float32)
        curve = np.zeros(epochs, dtype=np.float32)
        _check(load().sw_mlp_train(self.handle, _ptr(x), _ptr(t), x.shape[0], epochs, batch, lr, seed, _ptr(curve)))
        return curve

    def beam(self, delta, width: int, upper: int):
        """(bound, word) of the beam scored by this network."""
        return self.ctx.beam(self.n, self.k, delta, width, upper, mlp=self)


def mlp_samples(ctx: "Context", n: int, k: int, delta, fdepth: int = 14, width: int = 256, bdepth: int = 60,
                per_depth: int = 32, cap: int = 1 << 16):
    """(sets, distances) labelled on the search's forward levels."""
    d = _delta(n, k, delta)
    sets = np.zeros((cap, words(n)), dtype=np.uint64)
    dist = np.zeros(cap, dtype=np.float32)
    cnt = C.c_uint64()
    _check(load().sw_mlp_samples(ctx.handle, n, k, _ptr(d), fdepth, width, bdepth, per_depth, _ptr(sets), _ptr(dist),
                                 cap, C.byref(cnt)))
    return sets[: cnt.value], dist[: cnt.value]


class Context:
    """One device, one stream (sw_ctx)."""

    def __init__(self, device: int = 0):
        lib = load()
        h = C.c_void_p()
        _check(lib.sw_ctx_create(device, C.byref(h)))
        self.handle = h

    def close(self) -> None:
        if self.handle:
            load().sw_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def info(self) -> dict:
        name = C.create_string_buffer(64)
        cus = C.c_int()
        mem = C.c_uint64()
        _check(load().sw_device_info(self.handle, name, 64, C.byref(cus), C.byref(mem)))
        return {"name": name.value.decode(), "cus": cus.value, "mem_bytes": mem.value}

    def sync(self) -> None:
        _check(load().sw_sync(self.handle))

    def reset(self) -> None:
        """Free the context's device buffers (sw_ctx_reset); later calls allocate again."""
        _check(load().sw_ctx_reset(self.handle))

    def timer_start(self) -> None:
        """Record a HIP event on the context's stream."""
        _check(load().sw_stream_timer(self.handle, 0, None))

    def timer_stop(self) -> float:
        """Milliseconds on the context's stream since timer_start()."""
        ms = C.c_float()
        _check(load().sw_stream_timer(self.handle, 1, C.byref(ms)))
        return float(ms.value)

    # ------------------------------------------------ kernel-level calls
    def transition_apply(self, n: int, table, sets) -> np.ndarray:
        s = _sets(sets, n)
        t = np.ascontiguousarray(table, dtype=np.uint64)
        out = np.zeros_like(s)
        _check(load().sw_transition_apply(self.handle, n, _ptr(t), _ptr(s), _ptr(out), s.shape[0], SW_HOST))
        return out

    def expand(self, n: int, k: int, delta, sets, backward: bool = False) -> np.ndarray:
        s = _sets(sets, n)
        d = _delta(n, k, delta)
        out = np.zeros((k * s.shape[0], words(n)), dtype=np.uint64)
        _check(load().sw_expand(self.handle, n, k, _ptr(d), int(backward), _ptr(s), s.shape[0], _ptr(out), SW_HOST))
        return out

    def frontier_step(self, n: int, k: int, delta, frontier, visited=None, backward: bool = False):
        """One search step (sw_frontier_step): (new sets, parent index, letter,
        stats dict), new sets in candidate order, each at its first candidate."""
        f = _sets(frontier, n)
        v = _sets(visited if visited is not None else np.zeros((0, words(n)), dtype=np.uint64), n)
        d = _delta(n, k, delta)
        cap = k * f.shape[0]
        out = np.zeros((max(cap, 1), words(n)), dtype=np.uint64)
        par = np.zeros(max(cap, 1), dtype=np.uint32)
        let = np.zeros(max(cap, 1), dtype=np.uint8)
        cnt = C.c_uint64(0)
        st = _FrontierStats()
        _check(load().sw_frontier_step(self.handle, n, k, _ptr(d), int(backward), _ptr(v), v.shape[0], _ptr(f),
                                       f.shape[0], _ptr(out), _ptr(par), _ptr(let), C.byref(cnt), SW_HOST,
                                       C.byref(st)))
        m = cnt.value
        stats = {name: getattr(st, name) for name, _ in _FrontierStats._fields_}
        return out[:m], par[:m], let[:m], stats

    def subset_check(self, n: int, small, large, thresh: int = 1, per_row: bool = False, count: bool = True):
        """(results or None, pair_count or None, any) of the list-vs-list checker."""
        s = _sets(small, n)
        lg = _sets(large, n)
        res = np.zeros(s.shape[0], dtype=np.uint8) if per_row else None
        cnt = C.c_uint64()
        anyf = C.c_int32()
        _check(load().sw_subset_check(self.handle, n, _ptr(s), s.shape[0], _ptr(lg), lg.shape[0], thresh, SW_HOST,
                                      _ptr(res) if per_row else None,
                                      C.byref(cnt) if (count or per_row) else None, C.byref(anyf)))
        return (res.astype(bool) if per_row else None), (cnt.value if (count or per_row) else None), bool(anyf.value)

    def subset_check_device(self, n: int, small: DeviceArray, small_cnt: int, large: DeviceArray, large_cnt: int,
                            count: bool = True):
        cnt = C.c_uint64()
        anyf = C.c_int32()
        _check(load().sw_subset_check(self.handle, n, small.ptr, small_cnt, large.ptr, large_cnt, 1, SW_DEVICE,
                                      None, C.byref(cnt) if count else None, C.byref(anyf)))
        return (cnt.value if count else None), bool(anyf.value)

    def frontier_step_device(self, n: int, k: int, delta, frontier: DeviceArray, frontier_cnt: int,
                             visited: DeviceArray | None, visited_cnt: int, out: DeviceArray,
                             parent: DeviceArray | None = None, letter: DeviceArray | None = None,
                             backward: bool = False):
        """sw_frontier_step on device-resident lists: (new-set count, stats)."""
        d = _delta(n, k, delta)
        cnt = C.c_uint64(0)
        st = _FrontierStats()
        _check(load().sw_frontier_step(self.handle, n, k, _ptr(d), int(backward), visited.ptr if visited else None,
                                       visited_cnt, frontier.ptr, frontier_cnt, out.ptr,
                                       parent.ptr if parent else None, letter.ptr if letter else None,
                                       C.byref(cnt), SW_DEVICE, C.byref(st)))
        return cnt.value, {name: getattr(st, name) for name, _ in _FrontierStats._fields_}

    def subset_check_noalloc(self, n: int, small: DeviceArray, small_cnt: int, large: DeviceArray, large_cnt: int,
                             thresh: int = 1, results: DeviceArray | None = None,
                             pair_count: DeviceArray | None = None, any_out: DeviceArray | None = None) -> None:
        """The checker without allocation (sw_subset_check_noalloc): every
        output is device memory, written asynchronously on the stream."""
        _check(load().sw_subset_check_noalloc(self.handle, n, small.ptr, small_cnt, large.ptr, large_cnt, thresh,
                                              results.ptr if results else None,
                                              pair_count.ptr if pair_count else None,
                                              any_out.ptr if any_out else None))

    def beam(self, n: int, k: int, delta, width: int, upper: int, mlp: "Mlp | None" = None):
        """(bound, word) of a device beam: cardinality order, or scored by `mlp`."""
        d = _delta(n, k, delta)
        cap = max(int(upper), 1)
        word = np.zeros(min(cap, 1 << 20), dtype=np.uint32)
        b, wl = C.c_uint64(), C.c_uint64()
        _check(load().sw_beam_device(self.handle, mlp.handle if mlp else None, n, k, _ptr(d), width, upper,
                                     C.byref(b), _ptr(word), word.size, C.byref(wl)))
        return int(b.value), [int(v) for v in word[: wl.value]]

    def autotune_subset(self, n: int, s_count: int, t_count: int, p_stream: float, p_tile: float, reps: int = 3):
        """Tune the subset kernel's launch shape for this size class (kept in the context)."""
        th, wg, ee, ms = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_float()
        _check(load().sw_autotune_subset(self.handle, n, s_count, t_count, p_stream, p_tile, reps, C.byref(th),
                                         C.byref(wg), C.byref(ee), C.byref(ms)))
        return {"threads": th.value, "target_wgs": wg.value, "early_exit": ee.value & 255,
                "mask_bits": ee.value >> 8, "ms": ms.value}

    def set_subset_kernel(self, indexed: bool = True, min_pairs: int = 0) -> None:
        """Indexed subset kernel for launches of at least min_pairs pairs (0: 2^30), or never."""
        _check(load().sw_set_subset_kernel(self.handle, int(indexed), min_pairs))

    def set_subset_walk(self, walk: int = 0) -> None:
        """How the indexed kernel walks stream sets: as each launch asks (0),
        set order without early exit (1), rarest first with early exit over
        size-sorted lists (2), size-sorted lists in set order (3), anchor
        buckets where the mode allows (4)."""
        _check(load().sw_set_subset_walk(self.handle, walk))

    def calibrate_lds(self, bytes_per_lane: int = 8) -> float:
        """Measured LDS read bandwidth of the device, TB/s."""
        v = C.c_double()
        _check(load().sw_calibrate_lds(self.handle, bytes_per_lane, C.byref(v)))
        return v.value

    def set_subset_shape(self, n: int, s_count: int, t_count: int, threads: int, target_wgs: int,
                         early_exit: int, mask_bits: int):
        """Record a launch shape (as autotune_subset reports it) for this size class."""
        _check(load().sw_set_subset_shape(self.handle, n, s_count, t_count, threads, target_wgs,
                                          (early_exit & 1) | (mask_bits << 8)))

    def count_bits(self, n: int, sets) -> np.ndarray:
        s = _sets(sets, n)
        acc = np.zeros(n, dtype=np.uint64)
        _check(load().sw_count_bits(self.handle, n, _ptr(s), s.shape[0], _ptr(acc), SW_HOST))
        return acc

    def leaf(self, n: int, small, large_cnt: int) -> Leaf:
        return Leaf(self, n, small, large_cnt)

    # ------------------------------------------------- algorithm-level calls
    def exact(self, n: int, k: int, delta, lower: int = 0, upper: int | None = None, upper_word=None,
              cfg: ExactConfig | None = None) -> ExactResult:
        d = _delta(n, k, delta)
        if upper is None:
            upper = (n - 1) * (n - 1) + 1
        uw = np.ascontiguousarray(upper_word if upper_word is not None else [], dtype=np.uint32)
        cap = max(int(upper), 1)
        word = np.zeros(min(cap, 1 << 24), dtype=np.uint32)
        r = _ExactResult()
        c = (cfg or ExactConfig())._c()
        _check(load().sw_exact_solve(self.handle, n, k, _ptr(d), lower, upper, _ptr(uw), uw.size, C.byref(c),
                                     C.byref(r), _ptr(word), word.size))
        w = [int(x) for x in word[: r.word_len]] if r.word_len else None
        stats = {f: getattr(r, f) for f, _ in _ExactResult._fields_}
        return ExactResult(bool(r.found), int(r.length), w, bool(r.bound_word), stats)

    def exact_from_list(self, n: int, k: int, delta, sets, lower: int = 0, upper: int | None = None,
                        cfg: ExactConfig | None = None) -> tuple[ExactResult, int]:
        """The exact search from a list of sets (each of at least two states)
        instead of {Q}: the shortest length mapping some set to a singleton,
        and the list index its word starts from (-1 without a word)."""
        d = _delta(n, k, delta)
        s = _sets(sets, n)
        if upper is None:
            upper = (n - 1) * (n - 1) + 1
        word = np.zeros(min(max(int(upper), 1), 1 << 24), dtype=np.uint32)
        r = _ExactResult()
        c = (cfg or ExactConfig())._c()
        li = C.c_int64()
        _check(load().sw_exact_solve_from_list(self.handle, n, k, _ptr(d), _ptr(s), s.shape[0], lower, upper,
                                               C.byref(c), C.byref(r), _ptr(word), word.size, C.byref(li)))
        w = [int(x) for x in word[: r.word_len]] if r.word_len else None
        stats = {f: getattr(r, f) for f, _ in _ExactResult._fields_}
        return ExactResult(bool(r.found), int(r.length), w, bool(r.bound_word), stats), int(li.value)

    def brute(self, n: int, k: int, delta) -> int:
        """Shortest reset length by power-set BFS (n <= 32), -1 if not synchronizing."""
        d = _delta(n, k, delta)
        out = np.zeros(1, dtype=np.int64)
        _check(load().sw_brute(self.handle, n, k, _ptr(d), _ptr(out)))
        return int(out[0])

    def brute_batch(self, instances) -> np.ndarray:
        """Power-set BFS lengths of many small automata (list of (n, k, delta))."""
        cnt = len(instances)
        ns = np.array([i[0] for i in instances], dtype=np.uint32)
        ks = np.array([i[1] for i in instances], dtype=np.uint32)
        offs = np.zeros(cnt, dtype=np.uint64)
        parts = []
        o = 0
        for j, (n, k, d) in enumerate(instances):
            offs[j] = o
            dd = _delta(n, k, d)
            parts.append(dd)
            o += dd.size
        deltas = np.concatenate(parts) if parts else np.zeros(1, dtype=np.uint32)
        out = np.zeros(max(cnt, 1), dtype=np.int64)
        _check(load().sw_brute_batch(self.handle, cnt, _ptr(ns), _ptr(ks), _ptr(offs), _ptr(deltas), _ptr(out)))
        return out[:cnt]

    def eppstein(self, n: int, k: int, delta, upper: int | None = None):
        """Eppstein's greedy bound on the device: (non_synchro, bound, word or None)."""
        d = _delta(n, k, delta)
        up = (n - 1) * (n - 1) + 1 if upper is None else upper
        cap = min(up, 1 << 22)
        ns = np.zeros(1, dtype=np.int32)
        b = np.zeros(1, dtype=np.uint64)
        wl = np.zeros(1, dtype=np.uint64)
        w = np.zeros(max(cap, 1), dtype=np.uint32)
        _check(load().sw_eppstein(self.handle, n, k, _ptr(d), up, _ptr(ns), _ptr(b), _ptr(w), cap, _ptr(wl)))
        word = [int(x) for x in w[:int(wl[0])]] if wl[0] else None
        return bool(ns[0]), int(b[0]), word

    def eppstein_batch(self, instances, uppers=None, word_cap: int = 0):
        """Eppstein bounds of many automata in one launch: (non_synchro, bounds, words)."""
        cnt = len(instances)
        ns = np.array([i[0] for i in instances], dtype=np.uint32)
        ks = np.array([i[1] for i in instances], dtype=np.uint32)
        offs = np.zeros(cnt, dtype=np.uint64)
        parts = []
        o = 0
        for j, (n, k, d) in enumerate(instances):
            offs[j] = o
            dd = _delta(n, k, d)
            parts.append(dd)
            o += dd.size
        deltas = np.concatenate(parts) if parts else np.zeros(1, dtype=np.uint32)
        if uppers is None:
            uppers = [(int(n) - 1) ** 2 + 1 for n in ns]
        up = np.array(uppers, dtype=np.uint64) if cnt else np.zeros(1, dtype=np.uint64)
        nsync = np.zeros(max(cnt, 1), dtype=np.int32)
        b = np.zeros(max(cnt, 1), dtype=np.uint64)
        wl = np.zeros(max(cnt, 1), dtype=np.uint64)
        w = np.zeros(max(cnt * word_cap, 1), dtype=np.uint32)
        _check(load().sw_eppstein_batch(self.handle, cnt, _ptr(ns), _ptr(ks), _ptr(offs), _ptr(deltas), _ptr(up),
                                        word_cap, _ptr(nsync), _ptr(b), _ptr(wl), _ptr(w)))
        words = [[int(x) for x in w[j * word_cap: j * word_cap + min(int(wl[j]), word_cap)]] if wl[j] else None
                 for j in range(cnt)]
        return nsync[:cnt].astype(bool), b[:cnt], words

    def batch(self, instances, grid: int = 0, cap_rows: int = 0, cap_cand: int = 0, max_pairs: int = 0,
              word_cap: int = 512, overflow_workers: int = 0):
        """Exact lengths of many automata in one persistent launch.

        instances: list of (n, k, delta).  Returns (results, stats): results[i]
        = {"non_synchro", "length", "word"}.
        """
        cnt = len(instances)
        ns = np.array([i[0] for i in instances], dtype=np.uint32)
        ks = np.array([i[1] for i in instances], dtype=np.uint32)
        offs = np.zeros(cnt, dtype=np.uint64)
        parts = []
        o = 0
        for j, (n, k, d) in enumerate(instances):
            offs[j] = o
            dd = _delta(n, k, d)
            parts.append(dd)
            o += dd.size
        deltas = np.concatenate(parts) if parts else np.zeros(1, dtype=np.uint32)
        nsync = np.zeros(cnt, dtype=np.int32)
        lengths = np.zeros(cnt, dtype=np.uint64)
        wl = np.zeros(cnt, dtype=np.uint32)
        words_buf = np.zeros(cnt * word_cap, dtype=np.uint32)
        cfg = _BatchConfig(grid, cap_rows, cap_cand, overflow_workers, max_pairs)
        st = _BatchStats()
        _check(load().sw_batch_solve(self.handle, cnt, _ptr(ns), _ptr(ks), _ptr(offs), _ptr(deltas), C.byref(cfg),
                                     _ptr(nsync), _ptr(lengths), _ptr(wl), _ptr(words_buf), word_cap, C.byref(st)))
        res = []
        for j in range(cnt):
            w = [int(x) for x in words_buf[j * word_cap: j * word_cap + wl[j]]] if wl[j] else None
            res.append({"non_synchro": bool(nsync[j]), "length": int(lengths[j]), "word": w})
        stats = {f: getattr(st, f) for f, _ in _BatchStats._fields_ if not f.startswith("pad")}
        return res, stats

    def pipeline(self, n: int, k: int, delta, plan: dict | str | None = None) -> PipelineResult:
        d = _delta(n, k, delta)
        if isinstance(plan, dict):
            plan = json.dumps(plan)
        word = np.zeros(max((n - 1) * (n - 1) + 1, 1), dtype=np.uint32)
        r = _Result()
        _check(load().sw_pipeline_solve(self.handle, n, k, _ptr(d), plan.encode() if plan else None, C.byref(r),
                                        _ptr(word), word.size))
        names = [bytes(r.names[i]).split(b"\0")[0].decode() for i in range(r.n_algorithms)]
        times = [(names[i], float(r.ms[i])) for i in range(r.n_algorithms)]
        w = [int(x) for x in word[: r.word_len]] if r.word_len else None
        ex = {f: getattr(r.exact, f) for f, _ in _ExactResult._fields_}
        return PipelineResult(bool(r.non_synchro), int(r.lower), int(r.upper), w, int(r.eppstein_bound),
                              int(r.beam_bound), times, ex, int(r.reduced_n), int(r.reduce_steps))


def random_dfa(n: int, k: int, seed: int) -> np.ndarray:
    """gen.random_dfa computed by the library (same stream, native speed)."""
    out = np.zeros(n * k, dtype=np.uint32)
    load().sw_gen_random_dfa(n, k, seed, _ptr(out))
    return out


def replay(n: int, k: int, delta, word) -> int:
    """|delta(Q, word)| computed directly (host check of a returned word)."""
    d = np.asarray(delta, dtype=np.int64).reshape(n, k)
    cur = np.ones(n, dtype=bool)
    for a in word:
        nxt = np.zeros(n, dtype=bool)
        nxt[d[cur, int(a)]] = True
        cur = nxt
    return int(cur.sum())
