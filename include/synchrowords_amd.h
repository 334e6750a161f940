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
/* Seeded random DFA (the generator of synchrowords_amd/gen.py):
 * out[q*k + a] uniform in [0, n).  Host only. */
void sw_gen_random_dfa(uint32_t n, uint32_t k, uint64_t seed, uint32_t* out);
/* Create a context on HIP device `device` (its own stream and scratch). */
int sw_ctx_create(int device, sw_ctx** out);
void sw_ctx_destroy(sw_ctx* ctx);
/* Free every device buffer the context holds (scratch of the entry points,
 * search, beam and batch workspaces), keeping the context and its stream;
 * later calls allocate again.  Replaces reset_cuda_device_impl()
 * (synchrolib/data_structures/cuda/subsets_checker_kernel.hpp:8, a
 * cudaDeviceReset at subsets_checker_kernel.cu:64-66), which frees every
 * allocation on the device; this frees only the library's own, so other
 * users of the GPU in the process keep theirs. */
int sw_ctx_reset(sw_ctx* ctx);
/* Device name, compute-unit count and global memory bytes. */
int sw_device_info(sw_ctx* ctx, char* name, size_t name_cap, int* cus, uint64_t* mem_bytes);
/* Device memory for callers without their own allocator. */
int sw_dev_alloc(sw_ctx* ctx, size_t bytes, void** out);
int sw_dev_free(sw_ctx* ctx, void* p);
int sw_memcpy_h2d(sw_ctx* ctx, void* dst, const void* src, size_t bytes);
int sw_memcpy_d2h(sw_ctx* ctx, void* dst, const void* src, size_t bytes);
int sw_sync(sw_ctx* ctx);
/* Device-to-device copy by a plain streaming kernel (bytes % 16 == 0): the
 * known-size kernel that calibrates hardware traffic counters. */
int sw_dev_copy(sw_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Measured LDS read bandwidth of the device (TB/s): every CU reads whole
 * 64-lane rows of bytes_per_lane (8 or 16) bytes per lane, eight reads in
 * flight, for a few milliseconds.  The peak the subset kernel's LDS roofline is
 * quoted against.  No reference counterpart. */
int sw_calibrate_lds(sw_ctx* ctx, uint32_t bytes_per_lane, double* tb_per_s);
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

/* One breadth-first step of the meet-in-the-middle search with a visited
 * list: the images (dir 0) or preimages (dir 1) of every frontier set under
 * every letter, each distinct set once, none that `visited` holds.  Replaces
 * the apply, sort_keep_unique and visited-merge part of
 * MeetInTheMiddle::bfs_step with bfs_novisited false
 * (synchrolib/algorithm/exact/meet_in_the_middle.cpp:186-229), without its
 * dominance reduction; dir 1 applies preimages, as invbfs_step does
 * (:245-263), with the same deduplication.  Order: candidate order
 * (candidate a*cnt + i is frontier[i] under letter a), each set at its first
 * candidate; the reference returns them sorted.  `out` holds k*cnt sets;
 * out_parent[j] (frontier index) and out_letter[j] may be NULL.  visited may
 * hold duplicates.  where: SW_HOST or SW_DEVICE for every list.  stats may be
 * NULL; its kernel times are HIP events on the context's stream. */
typedef struct sw_frontier_stats {
  uint64_t candidates, new_sets, probes, table_slots;
  double ms_expand, ms_insert, ms_scan, ms_commit;
} sw_frontier_stats;
int sw_frontier_step(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta, int dir, const uint64_t* visited,
    uint64_t visited_cnt, const uint64_t* frontier, uint64_t cnt, uint64_t* out, uint32_t* out_parent,
    uint8_t* out_letter, uint64_t* out_count, int where, sw_frontier_stats* stats);

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

/* The same checker for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  small and large are device
 * pointers; the outputs are device pointers too and each may be NULL:
 * results[i] = (cnt_i >= thresh) as one byte per small set, *pair_count = the
 * total count, *any = whether any pair exists (with results and pair_count
 * both NULL the kernel stops at the first pair).  Asynchronous on the
 * context's stream (read the outputs after sw_sync); it calls no hipMalloc,
 * hipFree, hipMemset or synchronisation.  Its scratch is a static device
 * array of 16 MiB + 256 B per GPU, held until the process ends, so
 * small_cnt is limited to 4194304 (2^22) when results are asked for
 * (SW_ERR_INVALID beyond).  Every call on one GPU shares that scratch: calls
 * that overlap on one GPU race.  Uses the bit-walking kernel (no stream
 * index). */
int sw_subset_check_noalloc(sw_ctx* ctx, uint32_t n, const uint64_t* small, uint64_t small_cnt,
    const uint64_t* large, uint64_t large_cnt, uint32_t thresh, uint8_t* results, uint64_t* pair_count,
    int32_t* any);

/* Per-state frequency over a list: acc[q] = #{i : q in sets[i]} (acc host,
 * n entries).  Replaces count_bits_kernel<N>(subsets, acc, cnt)
 * (synchrolib/data_structures/cuda/count_bits_kernel.hpp:9-10), with 64-bit
 * counts. */
int sw_count_bits(sw_ctx* ctx, uint32_t n, const uint64_t* sets, uint64_t cnt, uint64_t* acc, int where);

/* Subset trie over a list of sets, resident on the device: the reference's
 * SubsetsTrie<N,Threads> (synchrolib/data_structures/subsets_trie.hpp:18-468),
 * which its DFS builds over the forward frontier (algorithm/exact/dfs.cpp:155)
 * and queries with backward sets.  Nodes split on the state held by the
 * most of their sets that not all hold (lowest state on ties,
 * subsets_trie.hpp:299-350); leaves hold at most 8 sets (any number at depth
 * 48); each node keeps its smallest and largest cardinality.
 * sw_trie_build == SubsetsTrie::build(vec) (subsets_trie.hpp:57-70), on the
 * device; the sets are copied (host) or read (device), not kept. */
typedef struct sw_trie sw_trie;
int sw_trie_build(sw_ctx* ctx, uint32_t n, const uint64_t* sets, uint64_t count, int where, sw_trie** out);
/* For each query, whether some trie set lies inside it (superset == 0: the
 * reference's contains_subset_of, subsets_trie.hpp:76-89,353-468, per query)
 * or contains it (superset == 1).  results (host, one byte per query) may be
 * NULL: then the walk stops at the first hit, *any says whether there was one
 * and *witness = (query index << 32) | row of the trie set in the built list
 * (~0 when none).  With results, *any = OR of the results. */
int sw_trie_query(sw_trie* t, int superset, const uint64_t* queries, uint64_t count, int where, uint8_t* results,
    int32_t* any, uint64_t* witness);
/* Sets, node ids used and level-synchronous build passes of the last build. */
int sw_trie_info(sw_trie* t, uint64_t* sets, uint64_t* nodes, uint32_t* top_levels);
/* Copies of the trie (host): node records as 4 x u32 each (internal: zero
 * child, one child, bit | minpop << 16, maxpop; leaf: first set, count,
 * 0xffff | minpop << 16, maxpop), the sets in trie order and their rows in
 * the built list.  Any output may be NULL. */
int sw_trie_nodes(sw_trie* t, uint32_t* nodes, uint64_t cap_nodes, uint64_t* sets, uint32_t* rows);
void sw_trie_destroy(sw_trie* t);

/* ------------------------------------------------ algorithm-level entry points */

/* Exact search settings (the reference's ExactConfig keys,
 * synchrolib/algorithm/exact/config.hpp:24-46, plus device settings). */
typedef struct sw_exact_config {
  uint64_t max_memory_mb;     /* device bytes the search may hold (0: 64 GiB) */
  uint64_t dominance_pairs;   /* dominance passes of at most this many set pairs
                                 (0: default, passes of 5e10 to 1e11 pairs with
                                 at least 10 steps left below the upper bound;
                                 for n <= 192 only once a side holds 2 GiB) */
  uint64_t forward_cap;       /* forward frontier size that freezes the forward side (0: default) */
  int32_t dfs;                /* when a step would pass the memory budget, deepen the backward side
                                 in parts that fit (the reference's DFS, dfs.cpp:63-347) instead of
                                 ending with a lower bound */
  int32_t verbose;            /* per-step log on stderr */
  uint64_t dfs_min_list_size; /* smallest backward part of that phase (0: 8192, the reference's
                                 default, config.hpp:24-46) */
  int32_t strict_memory_limit;/* 1: when even parts of dfs_min_list_size sets do not fit the budget,
                                 end with a lower bound (the reference's STRICT_MEMORY_LIMIT);
                                 0: use that size anyway */
  int32_t goal_check;         /* meet checks: 0 automatic (the subset kernel; the trie over the
                                 frozen forward frontier, where one build serves many checks),
                                 1 always the vertical subset kernel, 2 always the forward trie
                                 (the deepening phase's checks too), 3 tries over both frontiers
                                 (the one that did not step; the forward trie when deepening) */
  int32_t side_rule;          /* which side steps: 0 default (1), 1 the smaller frontier, 2 planned
                                 by the upper bound: the cheapest order of the steps left below it
                                 under a cost model of expansion and meet checks, first move taken */
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
  double ms_expand;     /* expansion, dedup and commit of new levels */
  double ms_dominance;  /* dominance pruning of new levels */
  double ms_goal;       /* meet checks between the frontiers */
  int32_t dfs_entered;  /* 1: the search deepened the backward side in parts (memory budget) */
  int32_t out_of_memory;/* 1: it ended at the budget; `length` is a lower bound */
  uint64_t dfs_calls;   /* parts expanded by that phase */
  uint64_t dfs_sets;    /* sets it generated */
  uint64_t trie_builds; /* device subset tries built for meet checks */
  uint64_t trie_queries;/* sets queried against them */
  double ms_dfs;        /* wall time of that phase */
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

/* The same search from a list of sets instead of {Q}: the length of a
 * shortest word mapping some set of the list to a singleton, within
 * [lower, upper] (upper: a known bound for that question).  `list` holds
 * `list_count` host sets in the Subset<N> layout (ceil(n/64) words each,
 * row-major), each of at least two states; repeated sets count once.  This
 * is the Exact plugin's run after a Reduce (exact.cpp:131-140 of the
 * reference: initialize_bfs_lists from data.result.reduce->list_bfs); the
 * caller adds the Reduce's bfs_steps back (exact.cpp:83-87).  The word, when
 * `word` is given and word_cap suffices, starts from a list set; its index
 * is written to *list_index (may be NULL, -1 without a word). */
int sw_exact_solve_from_list(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta, const uint64_t* list,
    uint64_t list_count, uint64_t lower, uint64_t upper, const sw_exact_config* cfg, sw_exact_result* out,
    uint32_t* word, uint64_t word_cap, int64_t* list_index);

/* Shortest reset length by breadth-first search over the subsets of Q,
 * the reference's Brute<N,K>::run (synchrolib/algorithm/brute/brute.hpp:28-122;
 * its BFS, :88-122): *length = the length, or -1 when the automaton is not
 * synchronizing.  n <= 32 (k <= 8 once the 2^n-bit visited set no longer fits
 * LDS, above about 20 states).  Unlike the reference, whose distances are
 * 8-bit, lengths of 255 and more are found.  Host arrays. */
int sw_brute(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta, int64_t* length);

/* The same for `count` automata (automaton i: ns[i] states, ks[i] letters,
 * transitions at deltas + delta_offsets[i]); small ones run one workgroup
 * each in one launch.  No reference counterpart for the batch form (the
 * reference runs one automaton per plan). */
int sw_brute_batch(sw_ctx* ctx, uint32_t count, const uint32_t* ns, const uint32_t* ks, const uint64_t* delta_offsets,
    const uint32_t* deltas, int64_t* lengths);

/* Eppstein's greedy bound, the reference's Eppstein<N,K>::run
 * (synchrolib/algorithm/eppstein/eppstein.hpp:25-114) over its pair tree
 * (data_structures/pairs_tree.hpp:25-65), on the device: *non_synchro = 1
 * when some pair of states never merges; otherwise, when the greedy word is
 * no longer than `upper`, *bound = its length and word[0 .. *word_len) = the
 * word (up to word_cap letters), else *bound = upper and *word_len = 0.
 * Same queue order and first letters as the reference, hence the same word.
 * n <= 512; host arrays; word and word_len may be NULL. */
int sw_eppstein(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta, uint64_t upper, int32_t* non_synchro,
    uint64_t* bound, uint32_t* word, uint64_t word_cap, uint64_t* word_len);

/* The same for `count` automata in one launch (one workgroup each):
 * automaton i at deltas + delta_offsets[i] with bound uppers[i]; its word at
 * words + i * word_cap.  words and word_lens may be NULL.  No reference
 * counterpart for the batch form. */
int sw_eppstein_batch(sw_ctx* ctx, uint32_t count, const uint32_t* ns, const uint32_t* ks,
    const uint64_t* delta_offsets, const uint32_t* deltas, const uint64_t* uppers, uint32_t word_cap,
    int32_t* non_synchro, uint64_t* bounds, uint64_t* word_lens, uint32_t* words);

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
  uint32_t reduced_n;           /* states of the automaton Exact searched after a Reduce (0: not reduced) */
  uint32_t reduce_steps;        /* forward steps the Reduce took before handing over (0: not reduced) */
} sw_result;

/* Run a plan (the reference's config JSON, e.g. configs/exact.json) on one
 * automaton, as the reference's run() does (jit/jitmain.cpp:11-49):
 * Brute, Eppstein, Beam and Exact in plan order, tightening [lower, upper].
 * `plan_json` NULL means configs/exact.json. */
int sw_pipeline_solve(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta, const char* plan_json,
    sw_result* out, uint32_t* word, uint64_t word_cap);

/* The same run, also giving the word the reference's AlgoResult.word holds
 * after the plan (algorithm/algorithm.hpp:19-42): only Eppstein sets it,
 * when the plan's Eppstein has find_word true and transition_tables false
 * and it lowered the upper bound (eppstein.hpp:70-72).  *result_word_len is
 * 0 when there is none; the word is written to result_word when
 * result_word_cap suffices.  `word` keeps sw_pipeline_solve's meaning (a
 * word of length `upper`, the shortest when Exact ran). */
int sw_pipeline_solve_result_word(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta,
    const char* plan_json, sw_result* out, uint32_t* word, uint64_t word_cap, uint32_t* result_word,
    uint64_t result_word_cap, uint64_t* result_word_len);

/* ----------------------------------------------------------- batched search */

typedef struct sw_batch_config {
  uint32_t grid;       /* workgroups of the persistent launch (0: 4 per CU, 3 when n > 128) */
  uint32_t cap_rows;   /* sets per side a workgroup may hold (0: 65536) */
  uint32_t cap_cand;   /* candidates of one step (0: 98304) */
  uint32_t overflow_workers; /* host threads finishing overflowed instances (0: 8) */
  uint64_t max_pairs;  /* largest goal check done inside a workgroup (0: 2^32) */
} sw_batch_config;

typedef struct sw_batch_stats {
  double ms_host_prep, ms_kernel, ms_overflow, ms_total;
  uint32_t non_synchro, found, bound, overflowed;
  uint32_t grid;
  uint32_t pad;
  uint64_t scratch_bytes;
  /* instances that left the workgroup, by reason: candidates of one step,
   * rows of a side, word buffer, hash table, goal-check pairs */
  uint32_t over_cand, over_rows, over_word, over_hash, over_pairs, pad2;
  double ms_overflow_search_sum;  /* summed wall time of the overflowed searches */
} sw_batch_stats;

/* Shortest reset words of `count` automata in one persistent launch.
 * Instance i: ns[i] states, ks[i] letters, transitions at
 * deltas + delta_offsets[i] (row-major, n*k entries).  Outputs per instance:
 * non_synchro[i]; lengths[i] (exact length when synchronizing); word_lens[i]
 * and the word at words + i*word_cap (both may be NULL).  Instances that do
 * not fit one workgroup finish on the whole device.  The reference runs such
 * a file one automaton at a time (app/main.cpp:49-78). */
int sw_batch_solve(sw_ctx* ctx, uint32_t count, const uint32_t* ns, const uint32_t* ks, const uint64_t* delta_offsets,
    const uint32_t* deltas, const sw_batch_config* cfg, int32_t* non_synchro, uint64_t* lengths,
    uint32_t* word_lens, uint32_t* words, uint32_t word_cap, sw_batch_stats* stats);

/* --------------------------------------------------------------- autotuner
 * Time every launch shape of the subset kernel (tile mask width 64/32,
 * workgroup size, workgroups aimed at, early exit; *early_exit returns
 * early_exit | mask_bits << 8; for a size class that takes the indexed
 * kernel, see sw_set_subset_kernel, the workgroups aimed at and the early
 * exit) on seeded
 * random lists of the given sizes and densities on the device; the fastest
 * is kept in the context and used by later calls of the same size class
 * (exact search, batch, the seams).
 * No reference counterpart (the reference's launch shapes are fixed,
 * cuda/subsets_implicit_trie_kernel.cu:8,196). */
int sw_autotune_subset(sw_ctx* ctx, uint32_t n, uint64_t s_count, uint64_t t_count, float p_stream, float p_tile,
    uint32_t reps, uint32_t* threads, uint32_t* target_wgs, uint32_t* early_exit, float* ms);

/* Record a launch shape (as sw_autotune_subset returns it; early_exit packs
 * early_exit | mask_bits << 8) for the size class of (n, s_count, t_count),
 * so a shape tuned once can be reused without timing again.  No reference
 * counterpart. */
int sw_set_subset_shape(sw_ctx* ctx, uint32_t n, uint64_t s_count, uint64_t t_count, uint32_t threads,
    uint32_t target_wgs, uint32_t early_exit);

/* Which subset kernel the context's later calls launch (exact search, seams):
 * with indexed != 0 (the default), launches of at least min_pairs pairs
 * (0: 2^30) list the states of every stream set once and use the indexed
 * kernel (two stream sets per wave, 128-bit LDS rows); the others, and all
 * launches when indexed == 0, walk the bits of each stream set per tile.
 * Both give the same results.  No reference counterpart. */
int sw_set_subset_kernel(sw_ctx* ctx, int indexed, uint64_t min_pairs);

/* How the indexed kernel walks the states of each stream set.  walk = 0 (the
 * default): as each launch asks — the exact search's goal checks and
 * dominance passes list every stream set's states rarest first (fewest sets
 * of the other list hold them), drop a stream set as soon as no set of the
 * other list can still hold it, and test it only against the sets holding
 * its rarest state (anchor buckets) or, where the mode does not allow that,
 * against size-sorted lists; large counts sort both lists by size; other
 * launches take the tuned shape.  1: set order, unsorted, no early exit, for
 * every launch.  2: sorted by size, rarest first with early exit.  3: sorted
 * by size, set order, no early exit.  4: anchor buckets, rarest first with
 * early exit, where the mode allows (not counts), else as 2.  Results are the
 * same in every mode.  No reference counterpart. */
int sw_set_subset_walk(sw_ctx* ctx, int walk);

/* -------------------------------------------------------- learned upper bound
 * A BF16 MLP (MFMA, FP32 accumulation) that scores beam candidates; lower
 * scores are kept.  No reference counterpart: it replaces the cardinality
 * order of the reference's Beam (algorithm/beam/beam.hpp:98-188) with a
 * learned one; the bound stays a real word's length.
 * Inputs of a set S of an automaton (in_dim = ceil((n+k+5)/32)*32 columns):
 *   column perm[q] = 1 for each state q in S, where perm is the automaton's
 *     canonical state order: states by how often they occur among the
 *     distinct images of Q first reached at depth 8 (breadth-first), most
 *     frequent first, ties by state number;
 *   column n+a = 8 |delta_a(S)|/n per letter;
 *   column n+k+j = 8 min |A \ S| / n over the images A of Q of depth <= 2(j+1), j < 4;
 *   column n+k+4 = 8 |S|/n; zero elsewhere.
 * Network in_dim -> H -> H -> 1, ReLU.  Parameters flattened as
 * [W1 (H x in_dim), b1, W2 (H x H), b2, w3 (H), b3].
 */
typedef struct sw_mlp sw_mlp;
int sw_mlp_create(sw_ctx* ctx, uint32_t n, uint32_t k, uint32_t hidden, uint64_t seed, sw_mlp** out);
void sw_mlp_destroy(sw_mlp* m);
int sw_mlp_shape(sw_mlp* m, uint32_t* in_dim, uint32_t* hidden, uint64_t* nparams);
int sw_mlp_get_params(sw_mlp* m, float* out, uint64_t cap);
int sw_mlp_set_params(sw_mlp* m, const float* in, uint64_t count);
/* BF16 features x[count][in_dim] of host sets for automaton delta. */
int sw_mlp_features(sw_mlp* m, const uint32_t* delta, const uint64_t* sets, uint64_t count, uint16_t* x);
/* y[count] = network(x[count][in_dim]). */
int sw_mlp_forward(sw_mlp* m, const uint16_t* x, uint64_t count, float* y);
/* Loss of sw_mlp_gradients / sw_mlp_train: rank_group 0 = mean squared error
 * against the targets (default); G in [2, 16] = pairwise ranking within
 * consecutive groups of G samples: sum over pairs (a, b) of a group with
 * t_a < t_b of softplus(y_a - y_b), divided by groups * G(G-1)/2. */
int sw_mlp_set_loss(sw_mlp* m, uint32_t rank_group);
/* Loss of one batch and its gradient (flattened like the parameters). */
int sw_mlp_gradients(sw_mlp* m, const uint16_t* x, const float* target, uint64_t count, float* grads, float* loss);
/* Adam, `epochs` passes of shuffled minibatches; loss_curve[e] = mean batch loss of epoch e (may be NULL). */
int sw_mlp_train(sw_mlp* m, const uint16_t* x, const float* target, uint64_t count, uint32_t epochs, uint32_t batch,
    float lr, uint64_t seed, float* loss_curve);
/* Training samples: beam candidate sets labelled with their forward distance
 * measured on the search's forward levels up to fdepth (fdepth+1 if beyond),
 * in groups of per_depth candidates of one beam depth (depths with fewer
 * candidates contribute none). */
int sw_mlp_samples(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta, uint32_t fdepth, uint32_t width,
    uint32_t bdepth, uint32_t per_depth, uint64_t* sets_out, float* dist_out, uint64_t cap, uint64_t* count);
/* Beam of `width` on the device, from the reference's start (sink-component
 * singletons with two preimages under a letter, beam.hpp:104-116), ordered by
 * the network m, or in the reference's order when m is NULL (larger sets
 * first, then word by word ascending; beam.hpp:145-162): *bound = first depth
 * producing Q (upper if none below it), with that word.  With m == NULL and
 * the automaton's states relabelled as the plan's presort does, the bound is
 * the reference Beam's bound for that width. */
int sw_beam_device(sw_ctx* ctx, sw_mlp* m, uint32_t n, uint32_t k, const uint32_t* delta, uint32_t width,
    uint64_t upper, uint64_t* bound, uint32_t* word, uint64_t word_cap, uint64_t* word_len);

#ifdef __cplusplus
}
#endif

#endif /* SYNCHROWORDS_AMD_H */
