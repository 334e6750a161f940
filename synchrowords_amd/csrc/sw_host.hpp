// Host-side internals: context, device buffers, error reporting.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/synchrowords_amd.h"

namespace sw {

void set_error(const std::string& msg);

// Thrown inside the library and turned into a status code at the C ABI.
struct Error {
  int code;
  std::string msg;
};

#define SW_CHECK_HIP(expr)                                                                  \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      throw ::sw::Error{e_ == hipErrorOutOfMemory ? SW_ERR_NOMEM : SW_ERR_HIP,              \
          std::string(#expr) + ": " + hipGetErrorString(e_) + " (" + __FILE__ + ":" +        \
              std::to_string(__LINE__) + ")"};                                              \
  } while (0)

#define SW_REQUIRE(cond, msg) \
  do {                        \
    if (!(cond)) throw ::sw::Error{SW_ERR_INVALID, (msg)}; \
  } while (0)

// Growable device allocation (contents kept on growth only if asked).
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  // Ensure at least `bytes`; when growing, keep the first `keep` bytes.
  void ensure(size_t bytes, hipStream_t st = nullptr, size_t keep = 0) {
    if (bytes <= cap) return;
    size_t nc = cap ? cap : 4096;
    while (nc < bytes) nc = nc * 3 / 2 + 4096;
    void* np = nullptr;
    SW_CHECK_HIP(hipMalloc(&np, nc));
    if (p && keep) {
      SW_CHECK_HIP(hipMemcpyAsync(np, p, keep < cap ? keep : cap, hipMemcpyDeviceToDevice, st));
      SW_CHECK_HIP(hipStreamSynchronize(st));
    }
    if (p) SW_CHECK_HIP(hipFree(p));
    p = np;
    cap = nc;
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// Buffers of the Reduce stage (reduce.cpp), reused between calls.
struct ReduceBufs {
  DevBuf masks, tables, list, next, cand, uniq, uidx, visited, flags, keep, pos, idx, scan, small, map, table, slot;
  size_t scan_bytes = 0;
  uint64_t* pinned = nullptr;  // read-back landing zone (64 bytes)
  uint64_t* host() {
    if (!pinned) SW_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&pinned), 64, hipHostMallocDefault));
    return pinned;
  }
  ReduceBufs() = default;
  ReduceBufs(const ReduceBufs&) = delete;
  ReduceBufs& operator=(const ReduceBufs&) = delete;
  ~ReduceBufs() {
    if (pinned) (void)hipHostFree(pinned);
  }
};

}  // namespace sw

struct sw_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string name;
  int cus = 0;
  uint64_t mem = 0;
  // Scratch of the kernel-level entry points.
  sw::DevBuf a, b, c, d, e;
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  // Subset-kernel launch shapes chosen by the autotuner, keyed by
  // (W, floor(log2 stream sets), floor(log2 tile sets)).
  struct Tune {
    uint32_t threads, target_wgs, early_exit, mask_bits;
    float ms;
  };
  std::map<uint32_t, Tune> tune;
  // Buffers of the device exact search, reused between searches.
  void* exact_ws = nullptr;
  void (*exact_ws_free)(void*) = nullptr;
  // Buffers of the device beam, reused between calls.
  void* beam_ws = nullptr;
  void (*beam_ws_free)(void*) = nullptr;
  // Buffers and events of sw_frontier_step, reused between calls.
  void* frontier_ws = nullptr;
  void (*frontier_ws_free)(void*) = nullptr;
  // Stream index of the indexed subset kernel, reused between launches, and
  // when it is used: never (0), or for launches of at least ix_min_pairs pairs.
  sw::DevBuf ix_padded, ix_size, ix_start, ix_list, ix_tmp;
  sw::DevBuf ix_order;  // state frequencies in the tile, ranks and rank order (3 x 512 words)
  sw::DevBuf ix_sort, ix_sperm, ix_tperm, ix_head;
  sw::DevBuf ix_anchor, ix_tlist, ix_work;  // anchor buckets: anchors, inverted lists, work items
  std::vector<uint32_t> ix_counts;         // host copy of the bucket sizes
  // One-workgroup exact search of a small automaton: instance, scratch, result and word.
  sw::DevBuf wg_in, wg_scratch, wg_out;
  // Power-set BFS batches: jobs, packed automata, frontiers, lengths.
  sw::DevBuf brute_jobs, brute_delta, brute_frontier, brute_out;
  // Eppstein batches: jobs, packed automata, pair-BFS scratch, words, results;
  // for a single automaton, the whole-device BFS's offsets, tile sums and
  // level state.
  sw::DevBuf epp_jobs, epp_delta, epp_scratch, epp_words, epp_out;
  sw::DevBuf epp_off, epp_tmp, epp_small;
  // How the indexed kernel walks stream sets: as each launch asks (0), always
  // in set order without early exit (1), always rarest first with early exit
  // over size-sorted lists (2), always size-sorted lists walked in set order
  // without early exit (3), anchor buckets where the mode allows, else 2 (4).
  int ix_walk = 0;
  size_t ix_tmp_bytes = 0;
  uint32_t* ix_host = nullptr;  // pinned: the index length read back
  // Pinned staging of the one-automaton stages' small copies (stage_bytes).
  void* stage = nullptr;
  size_t stage_cap = 0;
  int ix_enabled = 1;
  uint64_t ix_min_pairs = 1ull << 30;
  // The plan's Reduce stage.
  sw::ReduceBufs reduce;
  // The plan's helper (pipeline.cpp): a host thread with a context of its
  // own that runs the Beam's first pass while Eppstein runs on this one;
  // the stop cell that Eppstein's bound is written to, and the generation
  // of its latest use.
  void* helper = nullptr;
  void (*helper_free)(void*) = nullptr;
  unsigned long long* stop_cell = nullptr;  // pinned host memory, read by the beam's kernel
  uint32_t stop_gen = 0;
  bool helper_ctx = false;     // this context is such a helper's own
  bool helper_failed = false;  // no helper could be made: the stages run in order
};

namespace sw {

// Make the context's device current.
void use(sw_ctx* ctx);

// At least `bytes` of the context's pinned staging memory, for the small
// copies of one call; call it while the context's stream holds no copy from
// it (growing frees the old buffer).
void* stage_bytes(sw_ctx* ctx, size_t bytes);

// Contexts alive in this process that callers made (helpers' own not
// counted), and marking a context as a helper's own.
int user_contexts();
void mark_helper_context(sw_ctx* ctx);

// A stop cell for a beam that runs beside other work: the cell, in pinned
// host memory, holds (generation << 32) | depth, written by the host.
struct BeamStop {
  const unsigned long long* cell;
  uint32_t gen;
};
// sw_beam_device with an optional stop cell (mlp.cpp).
int beam_device(sw_ctx* ctx, sw_mlp* m, uint32_t n, uint32_t k, const uint32_t* delta, uint32_t width,
    uint64_t upper, uint64_t* bound, uint32_t* word, uint64_t word_cap, uint64_t* word_len, const BeamStop* stop);

// Look up the device address of the no-allocation entry point's static
// scratch on `device` (current); once per device (noalloc.hip).
void noalloc_init(int device);

struct VsubParams;
// Fill the launch shape of a subset-kernel call from the context's tuning
// table (the nearest tuned size class of the same W), if any.
void apply_tuning(sw_ctx* ctx, int W, VsubParams& p);

// Launch a subset test on the context's stream: the indexed kernel when it
// is enabled, fits and the launch is large enough (building the stream index
// first), else the bit-walking kernel.
// Eppstein's greedy bound for a batch: per automaton non_synchro, improved
// (greedy length <= uppers[i]), bounds (uppers[i] unless improved), the word
// length and the word (words + i * word_cap).  Any output may be NULL.
void eppstein_batch(sw_ctx* ctx, uint32_t count, const uint32_t* ns, const uint32_t* ks, const uint64_t* delta_offsets,
    const uint32_t* deltas, const uint64_t* uppers, uint64_t word_cap, int32_t* non_synchro, int32_t* improved,
    uint64_t* bounds, uint64_t* word_lens, uint32_t* words);
// Power-set BFS lengths (-1: not synchronizing) for a batch of automata.
void brute_batch(sw_ctx* ctx, uint32_t count, const uint32_t* ns, const uint32_t* ks, const uint64_t* delta_offsets,
    const uint32_t* deltas, int64_t* lengths);
void run_subset(sw_ctx* ctx, int W, VsubParams& p);

// Reference-layout slice tables for every letter: [k][ceil(n/8)][256][W]
// words; dir 0 = image, 1 = preimage.
std::vector<uint64_t> build_tables(uint32_t n, uint32_t k, const uint32_t* delta, int dir);
// The same tables built on the device into `out` from per-state masks
// computed here (kept in `host` and uploaded to `masks` until the next call).
void build_tables_device(uint32_t n, uint32_t k, const uint32_t* delta, int dir, std::vector<uint64_t>& host,
    DevBuf& masks, DevBuf& out, hipStream_t st);
// targets[a * n + q] = delta(q, a) as 16-bit words (the bit-sliced
// expansion's table), staged in `host` and uploaded to `out`.
void upload_targets(uint32_t n, uint32_t k, const uint32_t* delta, std::vector<uint16_t>& host, DevBuf& out,
    hipStream_t st);

// States of the sink component (exact.cpp).
std::vector<char> sink_states(uint32_t n, uint32_t k, const uint32_t* delta);

// Result of a search (exact.cpp).
struct ExactOutcome {
  sw_exact_result r{};
  std::vector<uint32_t> word;
  int64_t root = -1;  // the forward level-0 row the word starts from (-1: no word)
};
// A forward level 0 other than {Q}: `count` distinct device rows (the list a
// Reduce stage hands to Exact, exact.cpp:131-140 of the reference).
struct ExactSeed {
  const uint64_t* rows = nullptr;
  uint64_t count = 0;
};
// Forward levels of the exact search (images of Q up to `depth`).
void forward_levels(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta, uint32_t depth, uint64_t cap_rows,
    DevBuf& out, std::vector<uint64_t>& level_end);
// allow_workgroup: automata of at most 64 states first try the one-workgroup
// search (batch.cpp, exact_search_workgroup).
// seed: start the forward side from these rows instead of {Q}.
ExactOutcome exact_search(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta, uint64_t lower, uint64_t upper,
    const sw_exact_config& cfg, bool allow_workgroup = true, const ExactSeed* seed = nullptr);
bool exact_search_workgroup(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta, uint64_t upper,
    ExactOutcome& out);

}  // namespace sw
