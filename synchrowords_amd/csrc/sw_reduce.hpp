// The plan's Reduce stage (reduce.cpp).
#pragma once
#include <cstdint>
#include <vector>

#include "sw_host.hpp"

namespace sw {

// Outcome of the Reduce stage's forward search.
struct ReduceState {
  bool ran = false;      // the stage searched (n >= min_n and the bounds were open)
  bool decided = false;  // the search settled the length: lower = upper = length
  uint64_t length = 0;
  std::vector<uint32_t> word;  // decided by a singleton: a word of that length
  uint64_t steps = 0;          // forward steps taken (the reference's bfs_steps)
  bool reduced = false;        // fewer states reachable: Exact runs on the reduced automaton
  uint32_t n_red = 0;
  std::vector<uint32_t> delta_red;  // reduced transitions, delta_red[q * k + a]
  DevBuf list;                      // the last level on the reduced states (distinct sets, same order)
  uint64_t list_count = 0;
  // Per level, the kept candidates' indices (c = a * m + i: set i of the
  // level before under letter a) and that level's size m.
  std::vector<std::vector<uint32_t>> level_cand;
  std::vector<uint64_t> level_m;
  // The word from Q to set e of the last level.
  std::vector<uint32_t> prefix(uint32_t e) const;
};

// Run the stage on an automaton (upper: the current upper bound; threshold:
// the plan's list_size_threshold).
void reduce_stage(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* d, uint64_t upper, uint64_t threshold,
    ReduceState& rs);

// The distinct sets of a host list on the device (the first of equal sets
// kept, in list order): out_count rows in `out`, src[j] = the list index of
// row j.
void distinct_rows(sw_ctx* ctx, uint32_t W, const uint64_t* host_rows, uint64_t count, DevBuf& out,
    uint64_t& out_count, std::vector<uint32_t>& src);

}  // namespace sw
