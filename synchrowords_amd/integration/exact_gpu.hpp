// ExactGpu<N,K>: the reference's Exact plugin on the device.  Same interface
// (Algorithm<N,K>::run(AlgoData&), synchrolib/algorithm/algorithm.hpp:58-63)
// and the same effect on the result as Exact<N,K>::run
// (synchrolib/algorithm/exact/exact.cpp:36-97): nothing to do when the
// automaton is not synchronizing or the bounds already meet; n = 1 has
// length 0; otherwise the lower bound rises to what the search proved, and
// the upper bound becomes the length when the search found it.  The search
// is sw_exact_solve; the plan's Exact settings map onto sw_exact_config:
//   dfs                                    -> cfg.dfs
//   max_memory_mb                          -> cfg.max_memory_mb (the search's
//                                             budget, here in device memory)
//   dfs_min_list_size, strict_memory_limit -> the same fields
// (exact/config.hpp:24-46; the macros the JIT defines from the plan).
// When the budget ends the search ("mlsw >= x"), a word the deepening found
// on the way still lowers the upper bound.  Unlike the reference's Exact, it
// also stores the shortest word it found in result.word.  After a Reduce
// (data.result.reduce done: a smaller automaton and the Reduce's list), the
// search starts from that list (sw_exact_solve_from_list, as the
// reference's initialize_bfs_lists, exact.cpp:131-140) and the Reduce's
// steps are added back to the bounds (exact.cpp:83-87); its word starts at a
// list set, not at Q, so none is stored then.
#pragma once
#include <synchrolib/algorithm/algorithm.hpp>

#include <algorithm>
#include <vector>

#include "sw_context.hpp"

namespace synchrowords_amd {

template <synchrolib::uint N, synchrolib::uint K>
class ExactGpu : public synchrolib::Algorithm<N, K> {
 public:
  void run(synchrolib::AlgoData<N, K>& data) override {
    auto& r = data.result;
    if (r.non_synchro || r.mlsw_lower_bound == r.mlsw_upper_bound) return;
    if (N == 1) {
      r.mlsw_lower_bound = r.mlsw_upper_bound = 0;
      return;
    }
    std::vector<uint32_t> delta(static_cast<size_t>(N) * K);
    for (synchrolib::uint q = 0; q < N; ++q)
      for (synchrolib::uint a = 0; a < K; ++a) delta[static_cast<size_t>(q) * K + a] = data.aut[q][a];
    // A known word of the upper bound's length is returned when it is optimal.
    std::vector<uint32_t> known;
    if (r.word && r.word->size() == r.mlsw_upper_bound) known.assign(r.word->begin(), r.word->end());
    const uint64_t cap = std::min<uint64_t>(r.mlsw_upper_bound, uint64_t(1) << 24);
    std::vector<uint32_t> word(cap ? cap : 1);
    sw_exact_config cfg{};
    cfg.dfs = ExactGpu::kDfs;
    cfg.max_memory_mb = ExactGpu::kMaxMemoryMb;
    cfg.dfs_min_list_size = ExactGpu::kDfsMinListSize;
    cfg.strict_memory_limit = ExactGpu::kStrict;
    sw_exact_result out{};
    if (r.reduce && r.reduce->done) {
      constexpr uint64_t W = (N + 63) / 64;
      std::vector<uint64_t> list;
      list.reserve(r.reduce->list_bfs.size() * W);
      for (const auto& v : r.reduce->list_bfs) {
        uint64_t row[W] = {};
        for (size_t q = 0; q < v.size() && q < N; ++q)
          if (v[q]) row[q / 64] |= uint64_t(1) << (q % 64);
        list.insert(list.end(), row, row + W);
      }
      check(sw_exact_solve_from_list(context(), N, K, delta.data(), list.data(), r.reduce->list_bfs.size(),
          r.mlsw_lower_bound, r.mlsw_upper_bound, &cfg, &out, nullptr, 0, nullptr));
      const uint64_t steps = r.reduce->bfs_steps;
      r.mlsw_lower_bound += steps;
      r.mlsw_upper_bound += steps;
      const uint64_t length = out.length + steps;
      r.mlsw_lower_bound = std::max<uint64_t>(r.mlsw_lower_bound, length);
      if (out.found) r.mlsw_upper_bound = length;
      return;
    }
    check(sw_exact_solve(context(), N, K, delta.data(), r.mlsw_lower_bound, r.mlsw_upper_bound,
        known.empty() ? nullptr : known.data(), known.size(), &cfg, &out, word.data(), word.size()));
    r.mlsw_lower_bound = std::max<uint64_t>(r.mlsw_lower_bound, out.length);
    if (out.found) {
      r.mlsw_upper_bound = out.length;
      if (out.word_len) r.word = synchrolib::FastVector<synchrolib::uint>(word.begin(), word.begin() + out.word_len);
    } else if (out.word_len && out.word_len < r.mlsw_upper_bound) {
      r.mlsw_upper_bound = out.word_len;
      r.word = synchrolib::FastVector<synchrolib::uint>(word.begin(), word.begin() + out.word_len);
    }
  }

 private:
  // The plan's settings as the reference's JIT defines them (DFS,
  // MAX_MEMORY in MB, DFS_MIN_LIST_SIZE, STRICT_MEMORY_LIMIT); the
  // library's defaults where a build leaves them undefined.
#ifdef DFS
  static constexpr int kDfs = (DFS) ? 1 : 0;
#else
  static constexpr int kDfs = 1;
#endif
#if defined(MAX_MEMORY)
  static constexpr uint64_t kMaxMemoryMb = uint64_t(MAX_MEMORY);
#else
  static constexpr uint64_t kMaxMemoryMb = 0;
#endif
#ifdef DFS_MIN_LIST_SIZE
  static constexpr uint64_t kDfsMinListSize = uint64_t(DFS_MIN_LIST_SIZE);
#else
  static constexpr uint64_t kDfsMinListSize = 0;
#endif
#ifdef STRICT_MEMORY_LIMIT
  static constexpr int kStrict = (STRICT_MEMORY_LIMIT) ? 1 : 0;
#else
  static constexpr int kStrict = 0;
#endif
};

}  // namespace synchrowords_amd
