/*
 * synchrowords_amd — C ABI of the MI355X (gfx950) shortest-reset-word search.
 *
 * Plain C: pointers, sizes, integer status codes.  Every function returns
 * SW_OK (0) or a negative SW_ERR_* code; sw_last_error() gives the message of
 * the calling thread's last failure.  Nothing here calls exit().
 *
 * State sets use the reference's Subset<N> layout
 * (synchrolib/data_structures/subset.hpp:21-28): W = ceil(N/64) 64-bit words
 * per set, state q at bit q%64 of word q/64, bits >= N zero; lists are
 * row-major (set i = words [i*W, i*W + W)).  An automaton is delta[q*K + a]
 * (synchrolib/data_structures/automaton/automaton.hpp:13-33), states and
 * letters zero-based.
 *
 * Where an argument `where` is SW_HOST the list pointers are host memory and
 * are copied in and out (the contract of the reference's CUDA entry points);
 * with SW_DEVICE they are device pointers (from sw_dev_alloc or any HIP
 * allocation on the context's device) and results stay resident.
 *
 * Threading: a context belongs to one host thread at a time; it owns one HIP
 * stream and its scratch memory.  Calls on one context are serialised by the
 * caller.
 */
#ifndef SYNCHROWORDS_AMD_H
#define SYNCHROWORDS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SW_OK 0
#define SW_ERR_INVALID (-1)     /* bad argument */
#define SW_ERR_HIP (-2)         /* HIP runtime error */
#define SW_ERR_NOMEM (-3)       /* device memory exhausted */
#define SW_ERR_UNSUPPORTED (-4) /* outside the compiled range (N > 512, K > 255) */

#define SW_HOST 0
#define SW_DEVICE 1

typedef struct sw_ctx sw_ctx;
typedef struct sw_leaf sw_leaf;

/* ------------------------------------------------------------ housekeeping */
const char* sw_last_error(void);
int sw_version(void);
/* Create a context on HIP device `device` (its own stream and scratch). */
int sw_ctx_create(int device, sw_ctx** out);
void sw_ctx_destroy(sw_ctx* ctx);
/* Device name, compute-unit count and global memory bytes. */
int sw_device_info(sw_ctx* ctx, char* name, size_t name_cap, int* cus, uint64_t* mem_bytes);
/* Device memory for callers without their own allocator. */
int sw_dev_alloc(sw_ctx* ctx, size_t bytes, void** out);
int sw_dev_free(sw_ctx* ctx, void* p);
int sw_memcpy_h2d(sw_ctx* ctx, void* dst, const void* src, size_t bytes);
int sw_memcpy_d2h(sw_ctx* ctx, void* dst, const void* src, size_t bytes);
int sw_sync(sw_ctx* ctx);
/* HIP events on the context's stream, for timing work launched by later
 * calls: op 0 records the start event; op 1 records the stop event, waits for
 * it and writes the elapsed milliseconds to *ms. */
int sw_stream_timer(sw_ctx* ctx, int op, float* ms);

/* --------------------------------------------- kernel-level entry points
 * Each replaces one of the reference's CUDA entry points (the de-facto FFI
 * between SynchroLib's C++ and its device code).
 */

/* Batch transition apply: to[i] = OR over 8-state slices s of
 * table[s][byte s of from[i]], for `cnt` sets.  `table` has the layout of the
 * reference's PreprocessedTransition<N,K>::trans (Subset<N>[ceil(N/8)][256],
 * synchrolib/data_structures/preprocessed_transition.hpp:18-29) and is always
 * host memory.
 * Replaces preprocessed_transition_kernel<N,K>(trans, from, to, cnt)
 * (synchrolib/data_structures/cuda/preprocessed_transition_kernel.hpp:10-11). */
int sw_transition_apply(sw_ctx* ctx, uint32_t n, const uint64_t* table, const uint64_t* from, uint64_t* to,
    uint64_t cnt, int where);

/* Same, building the tables from the automaton: out[a*cnt + i] = image
 * (dir 0) or preimage (dir 1) of in[i] under letter a, for every letter. */
int sw_expand(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta, int dir, const uint64_t* in,
    uint64_t cnt, uint64_t* out, int where);

/* Implicit-trie leaf test, the reference's SubsetsImplicitTrieKernel<N,Proper>
 * (synchrolib/data_structures/cuda/subsets_implicit_trie_kernel.hpp:10-34):
 * create == allocate(small, small_cnt, large_cnt) (copies `small`, host);
 * run == run(small_pos, small_cnt, large, large_cnt, ret): ret[j] = 1 iff some
 * small[i], small_pos <= i < small_pos + small_cnt, is a subset of large[j]
 * (a proper subset if `proper`); large and ret are host memory;
 * destroy == deallocate(). */
int sw_leaf_create(sw_ctx* ctx, uint32_t n, const uint64_t* small, uint32_t small_cnt, uint32_t large_cnt,
    sw_leaf** out);
int sw_leaf_run(sw_leaf* leaf, int proper, uint32_t small_pos, uint32_t small_cnt, const uint64_t* large,
    uint32_t large_cnt, uint8_t* ret);
void sw_leaf_destroy(sw_leaf* leaf);

/* List-vs-list checker, the reference's check_subset_intersection_kernel_impl
 * (synchrolib/data_structures/cuda/subsets_checker_kernel.hpp:7-8, semantics
 * cuda/subsets_checker_kernel.cu:8-33): cnt_i = #{j : small[i] subset of
 * large[j]}; results[i] = cnt_i >= thresh (host, may be NULL; computing it
 * costs a per-row count).  Also returns the total pair count
 * (pair_count, may be NULL) and whether any pair exists (any, may be NULL).
 * With results == NULL and pair_count == NULL the kernel stops at the first
 * pair found. */
int sw_subset_check(sw_ctx* ctx, uint32_t n, const uint64_t* small, uint64_t small_cnt, const uint64_t* large,
    uint64_t large_cnt, uint32_t thresh, int where, uint8_t* results, uint64_t* pair_count, int32_t* any);

/* Per-state frequency over a list: acc[q] = #{i : q in sets[i]} (acc host,
 * n entries).  Replaces count_bits_kernel<N>(subsets, acc, cnt)
 * (synchrolib/data_structures/cuda/count_bits_kernel.hpp:9-10), with 64-bit
 * counts. */
int sw_count_bits(sw_ctx* ctx, uint32_t n, const uint64_t* sets, uint64_t cnt, uint64_t* acc, int where);

/* ------------------------------------------------ algorithm-level entry points */

/* Exact search settings (the reference's ExactConfig keys,
 * synchrolib/algorithm/exact/config.hpp:24-46, plus device settings). */
typedef struct sw_exact_config {
  uint64_t max_memory_mb;     /* device bytes the search may hold (0: 64 GiB) */
  uint64_t dominance_pairs;   /* pair budget of one dominance pass (0: default) */
  uint64_t forward_cap;       /* forward frontier size that freezes the forward side (0: default) */
  int32_t dfs;                /* allow the backward-only phase after the forward side freezes */
  int32_t verbose;            /* per-step log on stderr */
} sw_exact_config;

typedef struct sw_exact_result {
  int32_t found;        /* 1: `length` is the exact shortest length */
  int32_t bound_word;   /* 1: the search proved the supplied upper bound optimal */
  uint64_t length;      /* exact length (found) or a lower bound */
  uint64_t word_len;    /* letters written to `word` (0 if none) */
  uint64_t steps_forward, steps_backward;
  uint64_t sets_forward, sets_backward;  /* distinct sets generated per side */
  uint64_t pair_tests;  /* subset pairs tested (goal checks and dominance) */
  uint64_t hash_probes; /* probes made by the dedup hash */
  double ms_total;      /* wall time of the search */
} sw_exact_result;

/* Shortest reset word of a synchronizing automaton, given bounds
 * lower <= length <= upper (upper: the length of a known reset word, e.g. from
 * Beam; `upper_word` of that length may be passed and is returned when it is
 * optimal).  Replaces Exact<N,K>::run(AlgoData&)
 * (synchrolib/algorithm/exact/exact.cpp:36-97), the plugin the reference's
 * jit/jitmain.cpp:38-39 calls.  `word` (host, may be NULL) receives a
 * shortest reset word when word_cap suffices. */
int sw_exact_solve(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta, uint64_t lower, uint64_t upper,
    const uint32_t* upper_word, uint64_t upper_word_len, const sw_exact_config* cfg, sw_exact_result* out,
    uint32_t* word, uint64_t word_cap);

/* Result of a plan run: the reference's AlgoResult
 * (synchrolib/algorithm/algorithm.hpp:19-42). */
typedef struct sw_result {
  int32_t non_synchro;
  int32_t n_algorithms;
  uint64_t lower, upper;
  uint64_t word_len;            /* 0: no word */
  uint64_t eppstein_bound, beam_bound;
  double ms[8];                 /* per algorithm, plan order */
  char names[8][16];
  sw_exact_result exact;
} sw_result;

/* Run a plan (the reference's config JSON, e.g. configs/exact.json) on one
 * automaton, as the reference's run() does (jit/jitmain.cpp:11-49):
 * Brute, Eppstein, Beam and Exact in plan order, tightening [lower, upper].
 * `plan_json` NULL means configs/exact.json. */
int sw_pipeline_solve(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta, const char* plan_json,
    sw_result* out, uint32_t* word, uint64_t word_cap);

#ifdef __cplusplus
}
#endif

#endif /* SYNCHROWORDS_AMD_H */
