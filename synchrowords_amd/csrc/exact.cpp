// Device-resident exact search for the shortest reset word.
//
// The search is the reference's meet in the middle (Exact<N,K>,
// synchrolib/algorithm/exact/exact.cpp:36-97; meet_in_the_middle.cpp:57-344)
// re-planned for one GPU:
//
//   forward side   F_d = images of Q under words of length d,
//   backward side  B_d = preimages of sink-component singletons under words
//                  of length d, kept complemented (C = Q \ B; preimage commutes
//                  with complement, so the same tables expand C),
//   one step       expand the chosen side through slice tables, drop exact
//                  duplicates of anything the side ever produced (device hash
//                  set), then — while the pair budget allows — drop sets
//                  dominated by a produced set of the same side (forward:
//                  a proper subset; backward: a proper superset),
//   goal check     some forward set inside some backward set, i.e. disjoint
//                  from its complement (the vertical subset kernel, ANY mode).
//
// Every set the search ever keeps stays in a per-side arena with its parent
// row and letter, so the meeting pair (A, B) is turned into a word u.v: u
// spells the forward path to A, v the backward path from B to its singleton.
//
// Why the pruning is exact: a set X dropped at depth d has a witness Y kept at
// depth <= d with Y inside X (forward) — any word continuing X continues Y at
// no greater length — so a shortest word, if it passed through X, yields one
// through Y of no greater length; induction on the depth keeps a witness in
// each frontier.  Dropping backward sets of at most one state is the
// reference's own rule (dfs.cpp:285-287): on a shortest word every backward
// set strictly contains the forward image it meets, which has >= 2 states.
//
// Meet checks of large frontiers walk a device subset trie (k_trie.hip)
// built over the frontier that did not just change: a trie over the forward
// frontier answers "is some forward set inside this backward set", one over
// the backward frontier "does some backward set contain this forward set";
// the tries are rebuilt only when their side steps.
//
// The reference's cost model (meet_in_the_middle.cpp:346-551) picks sides
// and the switch to its DFS (dfs.cpp) to stay inside host memory.  Here the
// device keeps whole levels while they fit the budget (max_memory_mb),
// expands the smaller frontier, and freezes the forward side past a size
// cap (backward-only from then on).  When a step would pass the budget, the
// search deepens the backward frontier in parts instead, as the reference's
// DFS does (dfs.cpp:63-347): each part's preimages are deduplicated (against
// the backward side's rows and among themselves), checked against the
// forward frontier (the vertical subset kernel; the forward trie when the
// goal-check mode asks for it: on the budget cases the kernel took a fifth
// to a third of the trie's time, whose subset queries are the dense
// backward sets), and deepened in parts again while below the best length
// found; a find lowers that bound (dfs.cpp:179-189).  If even
// that does not fit, the result is a lower bound (exact.cpp:75-96 of the
// reference: "mlsw >= x").  Choices of side never change the answer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <climits>
#include <limits>
#include <cmath>
#include <string>

#include <memory>

#include "sw_host.hpp"
#include "sw_kernels.hpp"
#include "sw_trie.hpp"

namespace sw {

namespace {

struct Side {
  DevBuf arena, parent, letter, table, front, front_idx, front_idx_next;
  uint64_t rows = 0;
  uint64_t tcap = 0;
  std::vector<uint64_t> level_end;  // arena rows after each level
  uint64_t fcount = 0, prev_fcount = 0;
  uint64_t steps = 0;
  bool frozen = false;
  void reset() {
    rows = tcap = fcount = prev_fcount = steps = 0;
    level_end.clear();
    frozen = false;
  }
};

// Device buffers of the search, kept by the context between searches so
// that repeated searches (the batch's overflow path) allocate nothing and
// never call hipFree, which would synchronise the whole device.
struct Workspace {
  DevBuf tables[2], masks[2], targets;
  std::vector<uint64_t> host_masks[2];
  std::vector<uint16_t> host_targets;
  Side F, B;
  DevBuf cand, valid, slot, keep, pos, small, scan_tmp, flags, probes;
  size_t scan_bytes = 0;
  // Subset tries over the forward and the backward frontier (meet checks,
  // and the forward one for the backward-deepening phase).
  DeviceTrie ftrie, btrie;
  // Backward-deepening phase: one list per depth below the frontier (sets
  // complemented, parent index in the list above, letter) and the hash
  // table that deduplicates a part's preimages.
  struct Level {
    DevBuf sets, parent, letter;
  };
  std::vector<std::unique_ptr<Level>> dfs;
  DevBuf dfs_table;
  // Pinned landing zone for the few scalars read back per step (a copy to
  // pageable memory would go through a staging buffer).
  uint64_t* host = nullptr;
  Workspace() { SW_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&host), 64, hipHostMallocDefault)); }
  ~Workspace() {
    if (host) (void)hipHostFree(host);
  }
  Workspace(const Workspace&) = delete;
  Workspace& operator=(const Workspace&) = delete;
};

Workspace& workspace(sw_ctx* ctx) {
  if (!ctx->exact_ws) {
    ctx->exact_ws = new Workspace();
    ctx->exact_ws_free = [](void* p) { delete static_cast<Workspace*>(p); };
  }
  return *static_cast<Workspace*>(ctx->exact_ws);
}

class Search {
 public:
  Search(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta, const sw_exact_config& cfg,
      const ExactSeed* seed = nullptr)
      : seed_(seed), ctx_(ctx), st_(ctx->stream), n_(n), k_(k), W_((n + 63) / 64), slices_((n + 7) / 8), delta_(delta),
        cfg_(cfg), ws_(workspace(ctx)), tables_(ws_.tables), F_(ws_.F), B_(ws_.B), cand_(ws_.cand),
        valid_(ws_.valid), slot_(ws_.slot), keep_(ws_.keep), pos_(ws_.pos), small_(ws_.small),
        scan_tmp_(ws_.scan_tmp), flags_(ws_.flags), probes_(ws_.probes), scan_bytes_(ws_.scan_bytes) {
    F_.reset();
    B_.reset();
    // Default window [5e10, 1e11] pairs: smaller passes prune too little to
    // repay themselves and larger ones cost more than they save on most
    // automata (n = 300, k = 2 measured).  An explicit budget b means every
    // pass of at most b pairs.
    budget_pairs_ = cfg.dominance_pairs ? cfg.dominance_pairs : 100000000000ull;
    min_pairs_ = cfg.dominance_pairs ? 0 : 50000000000ull;
    // Nor, by default, with fewer than 10 steps left below the upper bound:
    // the sets a late pass drops would be expanded only a few more times
    // (n = 300: seed 1, bound 41, passes at lengths 34-36, 25.7 -> 22.0 ms;
    // passes with 14 or more steps left pay on the hard seeds).
    min_left_ = cfg.dominance_pairs ? 0 : 10;
    // And, by default, not for automata of at most 192 states unless a side
    // holds 2 GiB of sets: on the batch's classes (n <= 160, k = 2 and 3,
    // Eppstein bounds) the passes cost more than they save (60 k = 3 searches
    // 310 -> 295 ms without them); the size test keeps pruning where memory
    // would run out without it.
    small_sets_ = !cfg.dominance_pairs && W_ <= 3;
    forward_cap_ = cfg.forward_cap ? cfg.forward_cap : (1ull << 22);
    mem_cap_ = (cfg.max_memory_mb ? cfg.max_memory_mb : 65536ull) << 20;
    dfs_min_list_ = cfg.dfs_min_list_size ? cfg.dfs_min_list_size : 8192;
    // With the deepening on, the levels kept whole leave a quarter of the
    // budget to it, so that its parts are not starved.
    mitm_cap_ = cfg.dfs ? mem_cap_ - mem_cap_ / 4 : mem_cap_;
    goal_mode_ = cfg.goal_check;
    side_rule_ = cfg.side_rule;
  }

  ExactOutcome run(uint64_t lower, uint64_t upper) {
    upper_ = upper;
    // The smaller frontier steps unless the plan is asked for: on 39 searches
    // (n = 300, 200 and 150 random automata from their Beam bounds) the
    // planned order took 1346 ms against 916 ms, and fixed biases of the
    // balance (0.5x to 2x) moved the total by at most 12 %, so the balanced
    // order is the default.
    if (side_rule_ == 0) side_rule_ = 1;
    auto t0 = std::chrono::steady_clock::now();
    ExactOutcome out;
    sw_exact_result& r = out.r;
    std::memset(&r, 0, sizeof(r));
    if (n_ == 1) {
      r.found = 1;
      r.length = 0;
      return out;
    }
    upload_tables();
    init_forward();
    init_backward();
    uint64_t t = 0;
    bool found = false;
    uint64_t wit = 0;
    // The goal check of length t is read only after the step of length t + 1
    // has been enqueued up to its new-set count (one wait for both); a meet
    // found then ends the search at t and that step is dropped before its
    // commit.
    auto read_goal = [&](uint64_t len) {
      const auto g0 = std::chrono::steady_clock::now();
      const uint64_t fc = F_.fcount, bc = B_.fcount;
      const bool met = goal_read(&wit);
      ms_goal_ += ms_since(g0);
      if (cfg_.verbose) {
        std::fprintf(stderr, "[sw exact] len %llu goal |F| %llu x |B| %llu\n", (unsigned long long)len,
            (unsigned long long)fc, (unsigned long long)bc);
        if (W_ <= 5) {
          std::fprintf(stderr, "[sw exact]   forward sizes %s\n", size_quantiles(F_, false).c_str());
          std::fprintf(stderr, "[sw exact]   backward sizes %s\n", size_quantiles(B_, true).c_str());
        }
      }
      return met;
    };
    // The supplied upper bound is the length of a known word: search below it.
    while (true) {
      if (t + 1 >= upper || F_.fcount == 0 || B_.fcount == 0) {  // (no sets: every continuation pruned)
        if (goal_pending_ && read_goal(t)) found = true;
        break;
      }
      bool fwd;
      if (F_.frozen) fwd = false;
      else if (side_rule_ == 2) fwd = plan_forward(t, upper);
      else fwd = F_.fcount <= B_.fcount;
      ++t;
      // A large step is not enqueued on a guess: when the goal check meets,
      // its work would be thrown away (a search that ends by a meet ends one
      // step into the guess), and next to its cost the wait it saves is small.
      if (goal_pending_ && (fwd ? F_ : B_).fcount * k_ > kGuessCandidates && read_goal(t - 1)) {
        --t;
        found = true;
        break;
      }
      // A step of the other side when the chosen one would pass the budget:
      // every level that still fits shortens the deepening that follows.
      bool stepped = false;
      for (int attempt = 0; attempt < 2 && !stepped; ++attempt) {
        const bool f = attempt == 0 ? fwd : !fwd;
        if (f && F_.frozen) continue;
        try {
          step_launch(f ? F_ : B_, !f);
          fwd = f;
          stepped = true;
        } catch (const BudgetExceeded&) {
        }
      }
      if (goal_pending_ && read_goal(t - 1)) {
        --t;
        found = true;
        break;
      }
      if (!stepped) {
        --t;  // lengths up to t were checked
        budget_end_ = true;
        break;
      }
      step_finish(fwd ? F_ : B_, !fwd, t);
      const auto g0 = std::chrono::steady_clock::now();
      goal_launch(fwd);
      ms_goal_ += ms_since(g0);
    }
    bool oom = false;
    uint64_t length = 0;
    std::vector<uint32_t> word;
    int64_t root = -1;
    if (found) {
      length = t;
      word = reconstruct(wit);
      root = word_root_;
    } else if (budget_end_ && t + 1 < upper && F_.fcount && B_.fcount) {
      if (cfg_.dfs) {
        // Deepen the backward frontier in parts (dfs.cpp:63-91): found,
        // the bound's word optimal, or the budget exhausted again.
        const auto d0 = std::chrono::steady_clock::now();
        oom = !deepen(t, upper);
        ms_dfs_ = ms_since(d0);
        r.dfs_entered = 1;
        if (dfs_best_) {
          word = dfs_word_;
          root = dfs_root_;
          length = dfs_best_;
          found = !oom;
        } else if (!oom) {
          length = upper;
        }
      } else {
        oom = true;
      }
    }
    r.steps_forward = F_.steps;
    r.steps_backward = B_.steps;
    r.sets_forward = F_.rows;
    r.sets_backward = B_.rows;
    r.pair_tests = pair_tests_;
    r.hash_probes = read_probes();
    r.ms_expand = ms_expand_;
    r.ms_dominance = ms_dom_;
    r.ms_goal = ms_goal_;
    r.dfs_calls = dfs_calls_;
    r.dfs_sets = dfs_sets_;
    r.trie_builds = trie_builds_;
    r.trie_queries = trie_queries_;
    r.ms_dfs = ms_dfs_;
    if (found) {
      r.found = 1;
      r.length = length;
      out.word = word;
      out.root = root;
      r.word_len = out.word.size();
    } else if (oom) {
      // Lengths up to t are excluded; a word the deepening found before the
      // budget ran out is returned as a tighter upper bound.
      r.found = 0;
      r.out_of_memory = 1;
      r.length = std::max<uint64_t>(lower, t + 1);
      out.word = word;
      out.root = root;
      r.word_len = out.word.size();
    } else {
      // No word shorter than the upper bound: the bound is the length.
      r.found = 1;
      r.bound_word = 1;
      r.length = upper;
    }
    r.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return out;
  }

  // The memory budget is not an error: a step that would pass it ends the
  // search with the lengths checked so far as a lower bound, as the
  // reference's meet in the middle does (meet_in_the_middle.cpp:98-102,
  // exact.cpp:75-96: "mlsw >= x").
  struct BudgetExceeded {};
  [[noreturn]] void set_memory_error() { throw BudgetExceeded{}; }

  // Forward levels only (distinct images of Q by depth, dominance-pruned
  // frontiers), copied to `out`; level_end[d] = rows after depth d.
  void forward_levels(uint32_t depth, uint64_t cap_rows, DevBuf& out, std::vector<uint64_t>& level_end) {
    upload_tables();
    init_forward();
    for (uint32_t d = 1; d <= depth; ++d) {
      if (!F_.fcount || F_.rows + F_.fcount * k_ > cap_rows) break;
      try {
        step(F_, false, d);
      } catch (const BudgetExceeded&) {
        break;
      }
    }
    out.ensure(F_.rows * row_bytes() + 8);
    SW_CHECK_HIP(hipMemcpyAsync(out.p, F_.arena.p, F_.rows * row_bytes(), hipMemcpyDeviceToDevice, st_));
    SW_CHECK_HIP(hipStreamSynchronize(st_));
    level_end = F_.level_end;
  }

 private:
  const ExactSeed* seed_;  // forward level 0 other than {Q} (a Reduce stage's list)
  int64_t word_root_ = -1, dfs_root_ = -1, last_root_ = -1;  // forward level-0 row of a word
  sw_ctx* ctx_;
  hipStream_t st_;
  uint32_t n_, k_, W_, slices_;
  const uint32_t* delta_;
  sw_exact_config cfg_;
  uint64_t budget_pairs_, min_pairs_, min_left_, forward_cap_, mem_cap_;
  bool small_sets_ = false;
  static constexpr uint64_t kSmallSetsDominanceBytes = 2ull << 30;
  bool goal_armed_ = false;  // the last commit reset the goal flag and witness
  bool budget_end_ = false;  // the search ended at the memory budget
  bool dfs_trie_ = false;    // the deepening's meet checks walk the forward trie
  uint64_t dfs_min_list_ = 8192;
  uint64_t mitm_cap_ = 0;  // budget of the levels kept whole
  int goal_mode_ = 0;
  int side_rule_ = 0;
  // Which frontier each trie holds: the side's step count when built (~0: none).
  uint64_t ftrie_steps_ = ~0ull, btrie_steps_ = ~0ull;
  uint64_t trie_builds_ = 0, trie_queries_ = 0;
  // Backward deepening: parts of at most part_ sets, words of at most
  // max_depth_ letters still wanted, the best word found.
  uint64_t part_ = 0, max_depth_ = 0, dfs_best_ = 0, dfs_calls_ = 0, dfs_sets_ = 0;
  std::vector<uint32_t> dfs_word_;
  double ms_dfs_ = 0;
  uint64_t upper_ = ~0ull;     // the search's upper bound (none: forward levels only)

  uint64_t pair_tests_ = 0;
  double ms_expand_ = 0, ms_dom_ = 0, ms_goal_ = 0;  // host time between synchronizations
  bool goal_pending_ = false;  // a goal check launched and not read yet
  // Largest step (candidates) enqueued before the goal check before it is read.
  static constexpr uint64_t kGuessCandidates = 1ull << 16;
  // The step whose expansion is enqueued (step_launch) and not yet taken.
  struct {
    uint64_t m = 0, cnt = 0;
    std::chrono::steady_clock::time_point t0;
  } pend_;
  Workspace& ws_;
  DevBuf (&tables_)[2];
  Side &F_, &B_;
  DevBuf &cand_, &valid_, &slot_, &keep_, &pos_, &small_, &scan_tmp_, &flags_, &probes_;
  size_t& scan_bytes_;

  uint64_t row_bytes() const { return uint64_t(W_) * 8; }

  // The search's lists meet rarely: a stream set usually lies in no tile set
  // and is dropped after its first few states, rarest in the tile first; it
  // is tested only against the tile sets holding its rarest state (anchor
  // buckets) or, where those do not apply, sorted by size, only against tiles
  // with a set at least as large.
  // Cardinalities of a frontier (backward: of the preimages, i.e. the rows
  // complemented): min, deciles 1/5/9, max.  Verbose log only.
  std::string size_quantiles(Side& S, bool complement) {
    DevBuf h;
    h.ensure(size_t(n_ + 1) * 4);
    SW_CHECK_HIP(launch_fill_u32(h.as<uint32_t>(), 0u, n_ + 1, st_));
    SW_CHECK_HIP(launch_size_hist(int(W_), S.front.as<uint64_t>(), S.fcount, n_, complement ? 1 : 0,
        h.as<uint32_t>(), st_));
    std::vector<uint32_t> hist(n_ + 1);
    SW_CHECK_HIP(hipMemcpyAsync(hist.data(), h.p, hist.size() * 4, hipMemcpyDeviceToHost, st_));
    SW_CHECK_HIP(hipStreamSynchronize(st_));
    const double qs[] = {0.0, 0.1, 0.5, 0.9, 1.0};
    std::string out;
    for (double q : qs) {
      const uint64_t want = std::min<uint64_t>(S.fcount - 1, uint64_t(q * double(S.fcount)));
      uint64_t acc = 0;
      uint32_t c = 0;
      for (; c <= n_; ++c) {
        acc += hist[c];
        if (acc > want) break;
      }
      out += std::to_string(c) + " ";
    }
    return out;
  }

  void search_walk(VsubParams& p) const {
    p.ix_early_exit = 1;
    p.order_states = 1;
    p.sort_by_size = 1;
    p.bucket_anchor = 1;
    // With other contexts' searches in flight (the batch's overflow workers,
    // the command line's -j), launches of half the workgroups pack better
    // beside theirs: the batch's 1045 overflow searches eight at a time took
    // 1.09 s instead of 1.13; alone, the full width is faster (n = 300 seed
    // 1: 16.0 ms against 16.9).
    if (p.target_wgs == kSearchWgs && user_contexts() > 1) p.target_wgs = kSearchWgs / 2;
  }
  static constexpr uint32_t kSearchWgs = 1024;
  static double ms_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
  }

  void upload_tables() {
    for (int dir = 0; dir < 2; ++dir)
      build_tables_device(n_, k_, delta_, dir, ws_.host_masks[dir], ws_.masks[dir], tables_[dir], st_);
    upload_targets(n_, k_, delta_, ws_.host_targets, ws_.targets, st_);
    small_.ensure(64);
    probes_.ensure(8);
    SW_CHECK_HIP(hipMemsetAsync(probes_.p, 0, 8, st_));
    SW_CHECK_HIP(hipStreamSynchronize(st_));
  }

  uint64_t read_probes() {
    SW_CHECK_HIP(hipMemcpyAsync(ws_.host, probes_.p, 8, hipMemcpyDeviceToHost, st_));
    SW_CHECK_HIP(hipStreamSynchronize(st_));
    return ws_.host[0];
  }

  uint32_t read_u32(const void* dev) {
    SW_CHECK_HIP(hipMemcpyAsync(ws_.host, dev, 4, hipMemcpyDeviceToHost, st_));
    SW_CHECK_HIP(hipStreamSynchronize(st_));
    return uint32_t(ws_.host[0]);
  }

  // A scan's total written by the scan's last kernel straight into pinned
  // host memory (no copy launch): the host reads it once the stream is idle.
  uint32_t* host_total() { return reinterpret_cast<uint32_t*>(ws_.host + 6); }
  uint32_t take_host_total() {
    SW_CHECK_HIP(hipStreamSynchronize(st_));
    return *reinterpret_cast<volatile uint32_t*>(host_total());
  }

  void ensure_table(Side& S, uint64_t incoming) {
    const uint64_t need = (S.rows + incoming) * 2 + 1024;
    if (S.tcap >= need) return;
    uint64_t cap = 1024;
    while (cap < need) cap <<= 1;
    // Slot indices must stay below kEmpty (0xffffffff), which marks a
    // candidate that claimed no slot.
    if (cap > (1ull << 31)) set_memory_error();
    S.table.ensure(cap * 4);
    S.tcap = cap;
    SW_CHECK_HIP(launch_fill_u32(S.table.as<uint32_t>(), kEmpty, cap, st_));
    SW_CHECK_HIP(launch_hash_rebuild(int(W_), S.arena.as<uint64_t>(), S.rows, S.table.as<uint32_t>(), cap - 1, st_));
  }

  void ensure_rows(Side& S, uint64_t rows) {
    if (rows * row_bytes() > mem_cap_ / 4) set_memory_error();
    if (rows >= 0x7fffffffull) set_memory_error();  // rows are 31-bit
    S.arena.ensure(rows * row_bytes(), st_, S.rows * row_bytes());
    S.parent.ensure(rows * 4, st_, S.rows * 4);
    S.letter.ensure(rows, st_, S.rows);
  }

  void ensure_scan(uint64_t count) {
    size_t bytes = 0;
    SW_CHECK_HIP(keep_scan_temp_bytes(count, &bytes));
    if (bytes > scan_bytes_) {
      scan_tmp_.release();
      scan_tmp_.ensure(bytes);
      scan_bytes_ = bytes;
    }
  }

  // Seed a side with host rows (level 0).
  void seed(Side& S, const std::vector<uint64_t>& rows) {
    const uint64_t cnt = rows.size() / W_;
    ensure_rows(S, cnt);
    SW_CHECK_HIP(hipMemcpyAsync(S.arena.p, rows.data(), rows.size() * 8, hipMemcpyHostToDevice, st_));
    SW_CHECK_HIP(hipMemsetAsync(S.parent.p, 0xff, cnt * 4, st_));
    SW_CHECK_HIP(hipMemsetAsync(S.letter.p, 0xff, cnt, st_));
    S.rows = cnt;
    S.level_end.push_back(cnt);
    ensure_table(S, 0);
    S.front.ensure(cnt * row_bytes());
    S.front_idx.ensure(cnt * 4);
    SW_CHECK_HIP(launch_select(int(W_), S.arena.as<uint64_t>(), 0, nullptr, nullptr, cnt, S.front.as<uint64_t>(),
        S.front_idx.as<uint32_t>(), st_));
    S.fcount = cnt;
    S.prev_fcount = cnt;
  }

  void init_forward() {
    if (seed_) {
      // Distinct device rows as level 0 (the caller deduplicated them).
      const uint64_t cnt = seed_->count;
      ensure_rows(F_, cnt);
      SW_CHECK_HIP(hipMemcpyAsync(F_.arena.p, seed_->rows, cnt * row_bytes(), hipMemcpyDeviceToDevice, st_));
      SW_CHECK_HIP(hipMemsetAsync(F_.parent.p, 0xff, cnt * 4, st_));
      SW_CHECK_HIP(hipMemsetAsync(F_.letter.p, 0xff, cnt, st_));
      F_.rows = cnt;
      F_.level_end.push_back(cnt);
      ensure_table(F_, 0);
      F_.front.ensure(cnt * row_bytes());
      F_.front_idx.ensure(cnt * 4);
      SW_CHECK_HIP(launch_select(int(W_), F_.arena.as<uint64_t>(), 0, nullptr, nullptr, cnt, F_.front.as<uint64_t>(),
          F_.front_idx.as<uint32_t>(), st_));
      F_.fcount = cnt;
      F_.prev_fcount = cnt;
      return;
    }
    std::vector<uint64_t> q(W_, 0);
    for (uint32_t s = 0; s < n_; ++s) q[s >> 6] |= 1ull << (s & 63);
    seed(F_, q);
  }

  // Sink-component states with two preimages under some letter
  // (exact.cpp:148-164 of the reference); stored complemented.
  void init_backward() {
    const std::vector<char> sink = sink_states(n_, k_, delta_);
    std::vector<uint32_t> indeg(size_t(n_) * k_, 0);
    for (uint32_t q = 0; q < n_; ++q)
      for (uint32_t a = 0; a < k_; ++a) indeg[size_t(delta_[size_t(q) * k_ + a]) * k_ + a]++;
    std::vector<uint64_t> rows;
    for (uint32_t q = 0; q < n_; ++q) {
      if (!sink[q]) continue;
      bool ok = false;
      for (uint32_t a = 0; a < k_; ++a) ok |= indeg[size_t(q) * k_ + a] > 1;
      if (!ok) continue;
      for (uint32_t w = 0; w < W_; ++w) {
        uint64_t x = ~0ull;
        if (n_ < (w + 1) * 64) x = (n_ > w * 64) ? ((1ull << (n_ - w * 64)) - 1) : 0;
        if ((q >> 6) == w) x &= ~(1ull << (q & 63));
        rows.push_back(x);
      }
    }
    seed(B_, rows);
  }

  // One expansion step of side S (backward: complemented sets).
  void step(Side& S, bool backward, uint64_t t) {
    step_launch(S, backward);
    step_finish(S, backward, t);
  }
  // The step up to its new-set count: budget tests (BudgetExceeded before
  // anything is enqueued), expansion, hash insertion and the scan whose
  // total lands in pinned memory.  Nothing of the side changes that a
  // search ending here would need: the commit is step_finish's.
  void step_launch(Side& S, bool backward) {
    const auto t0 = std::chrono::steady_clock::now();
    const uint64_t m = S.fcount;
    const uint64_t cnt = m * k_;
    // Every budget test comes before anything changes, so that a step that
    // would pass the budget leaves both sides as they were (the backward
    // deepening starts from them).  Indices of rows and candidates are 31-bit.
    if (cnt >= 0x7fffffffull || S.rows + cnt >= 0x7fffffffull) set_memory_error();
    if ((S.rows + cnt) * row_bytes() > mitm_cap_ / 3) set_memory_error();
    if (used_bytes() + cnt * (3 * row_bytes() + 37) > mitm_cap_) set_memory_error();
    cand_.ensure(cnt * row_bytes());
    valid_.ensure(cnt * 4);
    slot_.ensure(cnt * 4);
    keep_.ensure(cnt * 4 + 4);
    pos_.ensure(cnt * 4 + 4);
    ensure_scan(cnt);
    const int W = int(W_);
    SW_CHECK_HIP(launch_expand_sets(W, S.front.as<uint64_t>(), m, tables_[backward ? 1 : 0].as<uint64_t>(), slices_,
        ws_.targets.as<uint16_t>(), backward ? 1 : 0, k_, nullptr, n_, backward ? 1 : 0, cand_.as<uint64_t>(),
        valid_.as<uint32_t>(), st_));
    ensure_table(S, cnt);
    SW_CHECK_HIP(launch_hash_insert(W, cand_.as<uint64_t>(), valid_.as<uint32_t>(), cnt, S.table.as<uint32_t>(),
        S.tcap - 1, S.arena.as<uint64_t>(), slot_.as<uint32_t>(), probes_.as<unsigned long long>(), st_));
    SW_CHECK_HIP(launch_keep_scan(1, slot_.as<uint32_t>(), S.table.as<uint32_t>(), nullptr, cnt, keep_.as<uint32_t>(),
        pos_.as<uint32_t>(), host_total(), scan_tmp_.p, &scan_bytes_, st_));
    pend_.m = m;
    pend_.cnt = cnt;
    pend_.t0 = t0;
  }
  void step_finish(Side& S, bool backward, uint64_t t) {
    const auto t0 = pend_.t0;
    const uint64_t m = pend_.m, cnt = pend_.cnt;
    const int W = int(W_);
    uint32_t* total = host_total();
    const uint64_t nnew = take_host_total();
    ensure_rows(S, S.rows + nnew);
    const uint64_t lo = S.rows;
    const uint64_t pairs = (lo + nnew) * nnew;
    const bool dominance = nnew > 0 && pairs <= budget_pairs_ && pairs >= min_pairs_ && upper_ >= t + 1 + min_left_ &&
        (!small_sets_ || (lo + nnew) * row_bytes() >= kSmallSetsDominanceBytes);
    // Without a dominance pass the commit also writes the next frontier (the
    // new level itself) and arms the goal check.  front_idx holds the
    // parents' rows, which the commit reads: the new indices go to the other
    // buffer, and front_idx is only resized after the commit.
    S.front.ensure(nnew * row_bytes() + 8);
    if (!dominance) S.front_idx_next.ensure(nnew * 4 + 4);
    int* gfound = small_.as<int>() + 2;
    unsigned long long* gwit = reinterpret_cast<unsigned long long*>(small_.as<char>() + 16);
    SW_CHECK_HIP(launch_commit(W, cand_.as<uint64_t>(), keep_.as<uint32_t>(), pos_.as<uint32_t>(),
        slot_.as<uint32_t>(), cnt, m, S.front_idx.as<uint32_t>(), 0, nullptr, lo, S.arena.as<uint64_t>(),
        S.parent.as<uint32_t>(), S.letter.as<uint8_t>(), S.table.as<uint32_t>(), st_,
        dominance ? nullptr : S.front.as<uint64_t>(), dominance ? nullptr : S.front_idx_next.as<uint32_t>(),
        dominance ? nullptr : gfound, dominance ? nullptr : gwit));
    if (!dominance) {
      std::swap(S.front_idx.p, S.front_idx_next.p);
      std::swap(S.front_idx.cap, S.front_idx_next.cap);
    } else {
      S.front_idx.ensure(nnew * 4 + 4);  // a reallocation waits for the commit
    }
    goal_armed_ = !dominance;
    S.rows = lo + nnew;
    S.level_end.push_back(S.rows);
    S.steps++;
    S.prev_fcount = S.fcount;
    if (dominance) {
      // Host time up to here (the new-level size was read back, so the
      // expansion and dedup are done) counts as expansion; the commit still
      // in flight counts as dominance.
      const auto t1 = std::chrono::steady_clock::now();
      ms_expand_ += std::chrono::duration<double, std::milli>(t1 - t0).count();
      // Tile: the new level; stream: every row of the side (the new level
      // included; a row is never its own witness).
      const size_t fb = vsub_flag_bytes(W, nnew);
      flags_.ensure(fb);
      SW_CHECK_HIP(hipMemsetAsync(flags_.p, 0, fb, st_));
      VsubParams p;
      p.n = n_;
      p.flags = flags_.p;
      if (!backward) {
        // A new image is dropped when some kept image lies inside it: the
        // kept images stream (they are small sets, the kernel walks their
        // states), the new level is the tile, flagged per tile set.
        p.stream = S.arena.as<uint64_t>();
        p.s_count = S.rows;
        p.tile = S.arena.as<uint64_t>() + lo * W_;
        p.t_count = nnew;
        p.mode = VSUB_FLAGS;
        p.diag = int64_t(lo);
      } else {
        // A new preimage is dropped when it lies inside a kept one.  Rows are
        // stored complemented, so both lists are complemented on load; the
        // new level streams (small sets once complemented), flagged per
        // stream set, against every row of the side as the tile.
        p.stream = S.arena.as<uint64_t>() + lo * W_;
        p.s_count = nnew;
        p.tile = S.arena.as<uint64_t>();
        p.t_count = S.rows;
        p.complement_stream = 1;
        p.complement_tile = 1;
        p.mode = VSUB_SFLAGS;
        p.diag = -int64_t(lo);
      }
      p.distinct_sets = 1;  // the side's rows are deduplicated
      apply_tuning(ctx_, W, p);
      search_walk(p);
      run_subset(ctx_, W, p);
      pair_tests_ += S.rows * nnew;
      ensure_scan(nnew);
      SW_CHECK_HIP(launch_keep_scan(2, nullptr, nullptr, flags_.p, nnew, keep_.as<uint32_t>(), pos_.as<uint32_t>(),
          total, scan_tmp_.p, &scan_bytes_, st_));
      SW_CHECK_HIP(launch_select(W, S.arena.as<uint64_t>(), lo, keep_.as<uint32_t>(), pos_.as<uint32_t>(), nnew,
          S.front.as<uint64_t>(), S.front_idx.as<uint32_t>(), st_));
      S.fcount = take_host_total();
      ms_dom_ += ms_since(t1);
    } else {
      S.fcount = nnew;
      ms_expand_ += ms_since(t0);
    }
    if (!backward && S.fcount >= forward_cap_) S.frozen = true;
    if (cfg_.verbose)
      std::fprintf(stderr, "[sw exact] len %llu %s parents %llu candidates %llu new %llu kept %llu rows %llu  %.3f ms%s\n",
          (unsigned long long)t, backward ? "B" : "F", (unsigned long long)m, (unsigned long long)cnt,
          (unsigned long long)nnew, (unsigned long long)S.fcount, (unsigned long long)S.rows, ms_since(t0),
          dominance ? "" : " (no dominance)");
  }

  // Which side steps, planned by the upper bound (the part of the
  // reference's cost model that reads the bound, remaining_iterations in
  // meet_in_the_middle.cpp:465-470, and picks the step).  The lengths left
  // below the bound are r = upper - 1 - t; each side's frontier is taken to
  // grow by its last step's ratio (kept sets over parents).  A step costs
  // its candidates (expansion, dedup, commit: kCostCand each) and the meet
  // check of the new level against the other frontier (kCostPair per pair);
  // the cheapest order of the r steps is found by a table over (forward
  // steps, backward steps), and its first move is taken.  Frontier sizes do
  // not depend on the order, so the table only decides how many steps each
  // side takes and which comes last: the final checks dominate, and a
  // forward side past forward_cap_ takes no more steps.  Small frontiers,
  // where every step costs about one launch latency, and searches without a
  // finite bound take the smaller frontier.
  static constexpr uint64_t kPlanSteps = 64;
  static constexpr uint64_t kPlanMinSets = 4096;
  static constexpr double kCostCand = 0.33e-9;  // s per candidate, measured at 2^24 candidates (n = 300)
  static constexpr double kCostPair = 1.1e-15;  // s per pair of a large meet check (n = 300, seed 1017)
  static double growth(const Side& S, uint32_t k) {
    if (!S.steps || !S.prev_fcount) return double(k);
    return std::min(double(k), std::max(0.25, double(S.fcount) / double(S.prev_fcount)));
  }
  bool plan_forward(uint64_t t, uint64_t upper) {
    const bool greedy = F_.fcount <= B_.fcount;
    if (upper <= t + 1) return greedy;
    const uint64_t r = upper - 1 - t;
    if (r > kPlanSteps || !F_.steps || !B_.steps || std::min(F_.fcount, B_.fcount) < kPlanMinSets) return greedy;
    const double gF = growth(F_, k_), gB = growth(B_, k_), K = double(k_);
    const double F0 = double(F_.fcount), B0 = double(B_.fcount), cap = double(forward_cap_);
    const double inf = std::numeric_limits<double>::infinity();
    // cost[i][j]: cheapest cost of the steps left after i forward and j
    // backward ones (i + j <= r); filled from i + j = r down.
    std::vector<double> cost((r + 1) * (r + 1), 0.0);
    auto at = [&](uint64_t i, uint64_t j) -> double& { return cost[i * (r + 1) + j]; };
    std::vector<double> fp(r + 1), bp(r + 1);
    fp[0] = F0;
    bp[0] = B0;
    for (uint64_t i = 1; i <= r; ++i) {
      fp[i] = fp[i - 1] * gF;
      bp[i] = bp[i - 1] * gB;
    }
    bool first_fwd = greedy;
    for (uint64_t s = r; s-- > 0;) {
      for (uint64_t i = 0; i <= s; ++i) {
        const uint64_t j = s - i;
        const double F = fp[i], B = bp[j];
        const double cf = F >= cap ? inf : kCostCand * K * F + kCostPair * F * gF * B + at(i + 1, j);
        const double cb = kCostCand * K * B + kCostPair * F * B * gB + at(i, j + 1);
        at(i, j) = std::min(cf, cb);
        if (s == 0) first_fwd = cf < cb || (cf == cb && greedy);
      }
    }
    return first_fwd;
  }

  // Meet check: some forward frontier set inside some backward set.
  // Large checks walk a subset trie over the frontier that did not just
  // step (the frozen forward side keeps its trie); small ones take the
  // vertical subset kernel (stream: forward sets; tile: the backward sets,
  // complemented back on load).
  static constexpr uint64_t kTriePairs = 1ull << 26;
  // Query lists up to this size are walked in list order (the cardinality
  // order costs four launches).
  static constexpr uint64_t kSortQueries = 1ull << 14;
  uint32_t trie_grid() const { return uint32_t(std::max(1, ctx_->cus)) * 3; }

  void ensure_ftrie() {
    if (ftrie_steps_ == F_.steps) return;
    ws_.ftrie.build(st_, n_, F_.front.as<uint64_t>(), F_.fcount, 0, nullptr);
    ftrie_steps_ = F_.steps;
    ++trie_builds_;
  }
  void ensure_btrie() {
    if (btrie_steps_ == B_.steps) return;
    ws_.btrie.build(st_, n_, B_.front.as<uint64_t>(), B_.fcount, 1, nullptr);
    btrie_steps_ = B_.steps;
    ++trie_builds_;
  }

  // The goal check in two halves: the launch (kernels, then the found flag
  // and witness copied into pinned memory) and the read, which waits for the
  // stream; the search enqueues the next step's expansion in between.
  bool goal_check(uint64_t* witness, bool stepped_forward) {
    goal_launch(stepped_forward);
    return goal_read(witness);
  }
  bool goal_read(uint64_t* witness) {
    if (!goal_pending_) return false;
    goal_pending_ = false;
    SW_CHECK_HIP(hipStreamSynchronize(st_));
    const volatile uint64_t* h = ws_.host + 2;
    *witness = h[1];
    return uint32_t(h[0]) != 0;
  }
  void goal_launch(bool stepped_forward) {
    goal_pending_ = false;
    if (!F_.fcount || !B_.fcount) return;
    int* found = small_.as<int>() + 2;
    unsigned long long* wit = reinterpret_cast<unsigned long long*>(small_.as<char>() + 16);
    if (!goal_armed_) SW_CHECK_HIP(launch_goal_reset(found, wit, st_));
    goal_armed_ = false;
    const uint64_t pairs = F_.fcount * B_.fcount;
    // The forward trie answers a backward set's query by walking the few
    // branches its states allow (backward sets are small), but building it
    // costs more than a kernel check of the same frontiers (n = 300 seed 1:
    // 12.1 -> 17.3 ms when it was rebuilt for runs of backward steps); so by
    // default it serves only where one build meets many queries: the frozen
    // forward side and the deepening phase.  A trie over the backward
    // frontier asked for supersets of forward sets visits most of its nodes
    // (forward sets are the larger ones), so only goal_check 3 uses it.
    const bool reuse = F_.frozen;
    const bool trie = goal_mode_ == 2 || goal_mode_ == 3 || (goal_mode_ == 0 && pairs >= kTriePairs && reuse);
    if (trie && (goal_mode_ != 3 || F_.frozen || !stepped_forward)) {
      // Forward trie, backward queries: witness (forward tag << 32) | query.
      ensure_ftrie();
      ws_.ftrie.query(st_, trie_grid(), 0, B_.front.as<uint64_t>(), B_.fcount, 1, B_.fcount > kSortQueries, 0,
          found, wit, nullptr);
      trie_queries_ += B_.fcount;
    } else if (trie) {
      // Backward trie (the complements complemented back), forward queries:
      // witness (query << 32) | backward tag.
      ensure_btrie();
      ws_.btrie.query(st_, trie_grid(), 1, F_.front.as<uint64_t>(), F_.fcount, 0, F_.fcount > kSortQueries, 1,
          found, wit, nullptr);
      trie_queries_ += F_.fcount;
    } else {
      VsubParams p;
      p.stream = F_.front.as<uint64_t>();
      p.s_count = F_.fcount;
      p.tile = B_.front.as<uint64_t>();
      p.t_count = B_.fcount;
      p.n = n_;
      p.complement_tile = 1;
      p.mode = VSUB_ANY;
      p.found = found;
      p.witness = wit;
      apply_tuning(ctx_, int(W_), p);
      search_walk(p);
      run_subset(ctx_, int(W_), p);
      pair_tests_ += pairs;
    }
    // found (4 bytes at +8) and witness (8 bytes at +16) in one copy
    SW_CHECK_HIP(hipMemcpyAsync(ws_.host + 2, small_.as<char>() + 8, 16, hipMemcpyDeviceToHost, st_));
    goal_pending_ = true;
  }

  // Device bytes this search holds, from its counts (buffers the context
  // kept from earlier, larger searches are not charged to it).
  uint64_t used_bytes() const {
    const uint64_t rb = row_bytes();
    uint64_t b = 2ull * k_ * slices_ * 256 * rb;  // slice tables of both directions
    for (const Side* sd : {&F_, &B_}) b += sd->rows * (rb + 5) + sd->tcap * 4 + sd->fcount * (rb + 8);
    if (ftrie_steps_ != ~0ull) b += ws_.ftrie.m * 3 * (rb + 4) + uint64_t(ws_.ftrie.node_count) * 16;
    if (btrie_steps_ != ~0ull) b += ws_.btrie.m * 3 * (rb + 4) + uint64_t(ws_.btrie.node_count) * 16;
    return b;
  }

  // Backward deepening below the frontier, lengths t + 1 .. max_depth_
  // (the reference's Dfs::run, dfs.cpp:63-91, 161-218): the frontier is cut
  // into parts; each part's preimages form the next list, which is checked
  // against the forward frontier and deepened in parts in turn.  Parts hold at
  // most part_ sets, sized so that one list per remaining depth fits the
  // budget (dfs.cpp:349-373).  False when the budget does not allow it.
  bool deepen(uint64_t t, uint64_t upper) {
    max_depth_ = upper - 1;
    dfs_trie_ = goal_mode_ == 2 || goal_mode_ == 3;
    try {
      if (dfs_trie_) ensure_ftrie();
      const uint64_t rb = row_bytes();
      const uint64_t used = used_bytes();
      const uint64_t avail = mem_cap_ > used ? mem_cap_ - used : 0;
      const uint64_t levels = max_depth_ - t;
      const uint64_t per = uint64_t(k_) * (rb + 5) * levels + uint64_t(k_) * (rb + 32);
      part_ = avail / per;
      if (part_ < dfs_min_list_) {
        if (cfg_.strict_memory_limit) return false;
        part_ = dfs_min_list_;
      }
      part_ = std::min<uint64_t>(part_, (1ull << 30) / k_);
      if (cfg_.verbose)
        std::fprintf(stderr, "[sw exact] deepening: lengths %llu..%llu, forward trie %llu sets, %llu backward sets, "
            "parts of %llu\n", (unsigned long long)(t + 1), (unsigned long long)max_depth_,
            (unsigned long long)F_.fcount, (unsigned long long)B_.fcount, (unsigned long long)part_);
      for (uint64_t b = 0; b < B_.fcount && t < max_depth_; b += part_)
        dfs_part(0, b, std::min(B_.fcount, b + part_), t);
    } catch (const BudgetExceeded&) {
      return false;
    } catch (const Error& e) {
      if (e.code != SW_ERR_NOMEM) throw;
      return false;  // the device ran out of memory
    }
    return true;
  }

  // Preimages of rows [b, e) of list d (d = 0: the backward frontier), words
  // of length lsw + 1, as list d + 1 (dfs.cpp:228-347): sets of at most one
  // state dropped, sets the backward side already holds dropped (produced
  // at no greater depth), duplicates among them dropped; then the meet check
  // and, below max_depth_, the parts of the new list.
  void dfs_part(uint32_t d, uint64_t b, uint64_t e, uint64_t lsw) {
    const int W = int(W_);
    const uint64_t* list = d == 0 ? B_.front.as<uint64_t>() : ws_.dfs[d - 1]->sets.as<uint64_t>();
    const uint64_t p = e - b, cnt = p * k_;
    cand_.ensure(cnt * row_bytes());
    valid_.ensure(cnt * 4);
    slot_.ensure(cnt * 4);
    keep_.ensure(cnt * 4 + 4);
    pos_.ensure(cnt * 4 + 4);
    ensure_scan(cnt);
    uint64_t tcap = 1024;
    while (tcap < 2 * cnt + 1024) tcap <<= 1;
    ws_.dfs_table.ensure(tcap * 4);
    uint32_t* table = ws_.dfs_table.as<uint32_t>();
    SW_CHECK_HIP(launch_expand_sets(W, list + b * W_, p, tables_[1].as<uint64_t>(), slices_,
        ws_.targets.as<uint16_t>(), 1, k_, nullptr, n_, 1, cand_.as<uint64_t>(), valid_.as<uint32_t>(), st_));
    SW_CHECK_HIP(launch_hash_lookup(W, cand_.as<uint64_t>(), valid_.as<uint32_t>(), cnt, B_.table.as<uint32_t>(),
        B_.tcap - 1, B_.arena.as<uint64_t>(), st_));
    SW_CHECK_HIP(launch_fill_u32(table, kEmpty, tcap, st_));
    SW_CHECK_HIP(launch_hash_insert(W, cand_.as<uint64_t>(), valid_.as<uint32_t>(), cnt, table, tcap - 1,
        B_.arena.as<uint64_t>(), slot_.as<uint32_t>(), probes_.as<unsigned long long>(), st_));
    uint32_t* total = small_.as<uint32_t>();
    SW_CHECK_HIP(launch_keep_scan(1, slot_.as<uint32_t>(), table, nullptr, cnt, keep_.as<uint32_t>(),
        pos_.as<uint32_t>(), total, scan_tmp_.p, &scan_bytes_, st_));
    const uint64_t c = read_u32(total);
    ++dfs_calls_;
    if (!c) return;
    while (ws_.dfs.size() <= d) ws_.dfs.emplace_back(new Workspace::Level());
    Workspace::Level& lv = *ws_.dfs[d];
    lv.sets.ensure(c * row_bytes());
    lv.parent.ensure(c * 4);
    lv.letter.ensure(c);
    SW_CHECK_HIP(launch_commit(W, cand_.as<uint64_t>(), keep_.as<uint32_t>(), pos_.as<uint32_t>(),
        slot_.as<uint32_t>(), cnt, p, nullptr, b, nullptr, 0, lv.sets.as<uint64_t>(), lv.parent.as<uint32_t>(),
        lv.letter.as<uint8_t>(), table, st_));
    dfs_sets_ += c;
    int* found = small_.as<int>() + 2;
    unsigned long long* wit = reinterpret_cast<unsigned long long*>(small_.as<char>() + 16);
    SW_CHECK_HIP(launch_goal_reset(found, wit, st_));
    if (dfs_trie_) {
      ws_.ftrie.query(st_, trie_grid(), 0, lv.sets.as<uint64_t>(), c, 1, c > kSortQueries, 0, found, wit, nullptr);
      trie_queries_ += c;
    } else {
      // Forward frontier (stream) against the new list (tile, complemented
      // back on load): witness (forward row << 32) | list row, as the trie's.
      VsubParams vp;
      vp.stream = F_.front.as<uint64_t>();
      vp.s_count = F_.fcount;
      vp.tile = lv.sets.as<uint64_t>();
      vp.t_count = c;
      vp.n = n_;
      vp.complement_tile = 1;
      vp.mode = VSUB_ANY;
      vp.found = found;
      vp.witness = wit;
      apply_tuning(ctx_, W, vp);
      search_walk(vp);
      run_subset(ctx_, W, vp);
      pair_tests_ += F_.fcount * c;
    }
    if (cfg_.verbose && (dfs_calls_ & 1023) == 0)
      std::fprintf(stderr, "[sw exact] deepening: %llu parts, %llu sets, list %u, length %llu\n",
          (unsigned long long)dfs_calls_, (unsigned long long)dfs_sets_, d + 1, (unsigned long long)(lsw + 1));
    SW_CHECK_HIP(hipMemcpyAsync(ws_.host, small_.as<char>() + 8, 16, hipMemcpyDeviceToHost, st_));
    SW_CHECK_HIP(hipStreamSynchronize(st_));
    if (uint32_t(ws_.host[0])) {
      // A word of length lsw + 1: only shorter ones are wanted now.
      dfs_word_ = dfs_reconstruct(d + 1, ws_.host[1]);
      dfs_root_ = word_root_;
      dfs_best_ = lsw + 1;
      max_depth_ = lsw;
      if (cfg_.verbose)
        std::fprintf(stderr, "[sw exact] deepening: word of length %llu (list %u)\n", (unsigned long long)(lsw + 1),
            d + 1);
      return;
    }
    if (lsw + 1 >= max_depth_) return;
    for (uint64_t q = 0; q < c; q += part_) {
      dfs_part(d + 1, q, std::min(c, q + part_), lsw + 1);
      if (lsw + 1 >= max_depth_) return;
    }
  }

  // The word of a meet found in list `level`: forward letters to the forward
  // set, then each list's letter up to the frontier, then the frontier set's
  // backward letters to its singleton.
  std::vector<uint32_t> dfs_reconstruct(uint32_t level, uint64_t witness) {
    const uint32_t f = uint32_t(witness >> 32);
    uint32_t idx = uint32_t(witness & 0xffffffffu);
    std::vector<uint32_t> v;
    for (uint32_t l = level; l >= 1; --l) {
      Workspace::Level& lv = *ws_.dfs[l - 1];
      uint32_t par = 0;
      uint8_t let = 0;
      SW_CHECK_HIP(hipMemcpyAsync(&par, lv.parent.as<uint32_t>() + idx, 4, hipMemcpyDeviceToHost, st_));
      SW_CHECK_HIP(hipMemcpyAsync(&let, lv.letter.as<uint8_t>() + idx, 1, hipMemcpyDeviceToHost, st_));
      SW_CHECK_HIP(hipStreamSynchronize(st_));
      v.push_back(let);
      idx = par;
    }
    std::vector<uint32_t> tail = path(B_, idx);
    v.insert(v.end(), tail.begin(), tail.end());
    std::vector<uint32_t> u = path(F_, f);
    word_root_ = last_root_;
    std::reverse(u.begin(), u.end());
    u.insert(u.end(), v.begin(), v.end());
    return u;
  }

  std::vector<uint32_t> path(Side& S, uint32_t frontier_pos) {
    uint32_t row = 0;
    SW_CHECK_HIP(hipMemcpyAsync(&row, S.front_idx.as<uint32_t>() + frontier_pos, 4, hipMemcpyDeviceToHost, st_));
    std::vector<uint32_t> parent(S.rows);
    std::vector<uint8_t> letter(S.rows);
    SW_CHECK_HIP(hipMemcpyAsync(parent.data(), S.parent.p, S.rows * 4, hipMemcpyDeviceToHost, st_));
    SW_CHECK_HIP(hipMemcpyAsync(letter.data(), S.letter.p, S.rows, hipMemcpyDeviceToHost, st_));
    SW_CHECK_HIP(hipStreamSynchronize(st_));
    std::vector<uint32_t> w;
    const uint64_t level0 = S.level_end.front();
    while (row >= level0) {
      w.push_back(letter[row]);
      row = parent[row];
    }
    last_root_ = int64_t(row);
    return w;  // letters from the set back to level 0
  }

  std::vector<uint32_t> reconstruct(uint64_t witness) {
    const uint32_t s = uint32_t(witness >> 32), t = uint32_t(witness & 0xffffffffu);
    std::vector<uint32_t> u = path(F_, s);
    word_root_ = last_root_;
    std::reverse(u.begin(), u.end());  // forward letters from level 0
    std::vector<uint32_t> v = path(B_, t);  // backward letters towards the singleton
    u.insert(u.end(), v.begin(), v.end());
    return u;
  }
};

}  // namespace

// The sink component: the strongly connected component closed first by a
// depth-first search from state 0 (utils/connectivity.cpp:105-160 of the
// reference).  For a synchronizing automaton it is the unique sink component.
std::vector<char> sink_states(uint32_t n, uint32_t k, const uint32_t* delta) {
  std::vector<uint32_t> idx(n, 0), low(n, 0), stack;
  std::vector<char> on(n, 0), sink(n, 0);
  struct Frame {
    uint32_t v, a;
  };
  uint32_t counter = 1;
  std::vector<Frame> call{{0, 0}};
  idx[0] = low[0] = counter++;
  stack.push_back(0);
  on[0] = 1;
  while (!call.empty()) {
    Frame& f = call.back();
    if (f.a < k) {
      const uint32_t w = delta[size_t(f.v) * k + f.a];
      f.a++;
      if (!idx[w]) {
        idx[w] = low[w] = counter++;
        stack.push_back(w);
        on[w] = 1;
        call.push_back({w, 0});
      } else if (on[w]) {
        low[f.v] = std::min(low[f.v], idx[w]);
      }
      continue;
    }
    const uint32_t v = f.v;
    call.pop_back();
    if (!call.empty()) low[call.back().v] = std::min(low[call.back().v], low[v]);
    if (low[v] == idx[v]) {
      uint32_t w;
      do {
        w = stack.back();
        stack.pop_back();
        on[w] = 0;
        sink[w] = 1;
      } while (w != v);
      return sink;  // the first component closed has no edge leaving it
    }
  }
  return sink;
}

std::vector<uint64_t> build_tables(uint32_t n, uint32_t k, const uint32_t* delta, int dir) {
  const uint32_t W = (n + 63) / 64, S = (n + 7) / 8;
  std::vector<uint64_t> t(size_t(k) * S * 256 * W, 0);
  // Preimage lists per letter.
  std::vector<std::vector<uint32_t>> pre;
  if (dir == 1) {
    pre.assign(size_t(n) * k, {});
    for (uint32_t q = 0; q < n; ++q)
      for (uint32_t a = 0; a < k; ++a) pre[size_t(delta[size_t(q) * k + a]) * k + a].push_back(q);
  }
  for (uint32_t a = 0; a < k; ++a)
    for (uint32_t s = 0; s < S; ++s)
      for (uint32_t b = 1; b < 256; ++b) {
        uint64_t* row = &t[((size_t(a) * S + s) * 256 + b) * W];
        for (uint32_t j = 0; j < 8; ++j) {
          const uint32_t q = s * 8 + j;
          if (q >= n || !((b >> j) & 1)) continue;
          if (dir == 0) {
            const uint32_t x = delta[size_t(q) * k + a];
            row[x >> 6] |= 1ull << (x & 63);
          } else {
            for (uint32_t p : pre[size_t(q) * k + a]) row[p >> 6] |= 1ull << (p & 63);
          }
        }
      }
  return t;
}

void build_tables_device(uint32_t n, uint32_t k, const uint32_t* delta, int dir, std::vector<uint64_t>& host,
    DevBuf& masks, DevBuf& out, hipStream_t st) {
  const uint32_t W = (n + 63) / 64, S = (n + 7) / 8;
  host.assign(size_t(k) * n * W, 0);
  for (uint32_t q = 0; q < n; ++q)
    for (uint32_t a = 0; a < k; ++a) {
      const uint32_t x = delta[size_t(q) * k + a];
      if (dir == 0) host[(size_t(a) * n + q) * W + (x >> 6)] |= 1ull << (x & 63);  // image of q
      else host[(size_t(a) * n + x) * W + (q >> 6)] |= 1ull << (q & 63);          // q is a preimage of x
    }
  masks.ensure(host.size() * 8);
  out.ensure(size_t(k) * S * 256 * W * 8);
  SW_CHECK_HIP(hipMemcpyAsync(masks.p, host.data(), host.size() * 8, hipMemcpyHostToDevice, st));
  SW_CHECK_HIP(launch_tables(int(W), masks.as<uint64_t>(), n, k, out.as<uint64_t>(), st));
}

void upload_targets(uint32_t n, uint32_t k, const uint32_t* delta, std::vector<uint16_t>& host, DevBuf& out,
    hipStream_t st) {
  host.resize(size_t(k) * n);
  for (uint32_t q = 0; q < n; ++q)
    for (uint32_t a = 0; a < k; ++a) host[size_t(a) * n + q] = uint16_t(delta[size_t(q) * k + a]);
  out.ensure(host.size() * 2 + 8);
  SW_CHECK_HIP(hipMemcpyAsync(out.p, host.data(), host.size() * 2, hipMemcpyHostToDevice, st));
}

void forward_levels(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta, uint32_t depth, uint64_t cap_rows,
    DevBuf& out, std::vector<uint64_t>& level_end) {
  use(ctx);
  sw_exact_config cfg{};
  cfg.forward_cap = ~0ull;
  Search s(ctx, n, k, delta, cfg);
  s.forward_levels(depth, cap_rows, out, level_end);
}

ExactOutcome exact_search(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta, uint64_t lower, uint64_t upper,
    const sw_exact_config& cfg, bool allow_workgroup, const ExactSeed* seed) {
  use(ctx);
  if (seed) {
    SW_REQUIRE(seed->rows && seed->count && seed->count < 0x7fffffffull, "empty or oversized forward seed");
    allow_workgroup = false;  // the one-workgroup search starts from Q
  }
  // Small automata: the whole search in one workgroup when it fits (no
  // per-step launches; a few MB of scratch, inside any budget a plan sets);
  // the other settings below only concern the device-wide search.
  if (allow_workgroup && n <= 64 && !cfg.verbose && !cfg.goal_check && !cfg.dominance_pairs && !cfg.forward_cap &&
      (!cfg.max_memory_mb || cfg.max_memory_mb >= 64)) {
    ExactOutcome o;
    if (exact_search_workgroup(ctx, n, k, delta, upper, o)) {
      o.r.length = std::max(o.r.length, lower);
      return o;
    }
  }
  Search s(ctx, n, k, delta, cfg, seed);
  return s.run(lower, upper);
}

}  // namespace sw
