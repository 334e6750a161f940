// The plan's Reduce stage on the device (the reference's Reduce<N,K>,
// synchrolib/algorithm/reduce/reduce.hpp:31-103).
//
// A short forward search from {Q}: each level is the distinct images of the
// last one under every letter, less every image that contains a set of an
// earlier level (reduce.hpp:202; the implicit trie sorts and deduplicates
// both lists first, subsets_implicit_trie.hpp:29-33) and every image that
// properly contains another image of the level (reduce.hpp:203).  It stops when a level holds a singleton (the length is
// found), when the next step would reach the upper bound, or when a level
// holds more than list_size_threshold sets (reduce.hpp:108-122).  In the last
// case the states reachable from the level's sets (closed under the
// transitions, reduce.hpp:165-188) may be fewer than N: the automaton
// restricted to them, with the level renumbered, is what the Exact stage
// searches from that level on, and its length gets the steps added back
// (exact.cpp:83-87, 131-140 of the reference).  The reference rebuilds its
// JIT for the smaller N; here the search simply runs with fewer 64-bit words
// per set (n = 300 random automata reduce to 228-252 states: 5 words -> 4).
//
// Every level is computed on the device with the search's own kernels: the
// slice-table expansion, the dedup hash (the first of equal images is kept),
// two passes of the vertical subset kernel in FLAGS mode into one bitmap (the
// earlier levels against the new one; the new one against itself, pairs of
// equal cardinality skipped), keep scans and selects.  Only the level sizes,
// the kept candidates' indices (for the word) and a singleton flag are read
// back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "sw_host.hpp"
#include "sw_kernels.hpp"
#include "sw_reduce.hpp"

namespace sw {

std::vector<uint32_t> ReduceState::prefix(uint32_t e) const {
  // Level l (1-based) holds kept candidates c = a * m + i of the m sets of
  // level l - 1: set i under letter a.
  std::vector<uint32_t> w;
  for (size_t l = level_cand.size(); l >= 1; --l) {
    const uint32_t c = level_cand[l - 1][e];
    const uint64_t m = level_m[l - 1];
    w.push_back(uint32_t(c / m));
    e = uint32_t(c % m);
  }
  std::reverse(w.begin(), w.end());
  return w;
}

void reduce_stage(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* d, uint64_t upper, uint64_t threshold,
    ReduceState& rs) {
  use(ctx);
  hipStream_t st = ctx->stream;
  const uint32_t W = (n + 63) / 64, slices = (n + 7) / 8;
  const uint64_t rb = uint64_t(W) * 8;
  ReduceBufs& b = ctx->reduce;
  SW_REQUIRE(!rs.ran, "Reduce state used twice");
  rs.ran = true;
  std::vector<uint64_t> hmasks;
  build_tables_device(n, k, d, 0, hmasks, b.masks, b.tables, st);
  b.small.ensure(128);
  uint64_t* hs = b.host();

  // Level 0: {Q}.
  std::vector<uint64_t> q(W, 0);
  for (uint32_t s = 0; s < n; ++s) q[s >> 6] |= 1ull << (s & 63);
  b.list.ensure(rb);
  SW_CHECK_HIP(hipMemcpyAsync(b.list.p, q.data(), rb, hipMemcpyHostToDevice, st));
  uint64_t m = 1, vcount = 0, mlsw = 0;
  const uint64_t max_mlsw = upper - 1;
  int64_t single = -1;
  while (mlsw < max_mlsw) {
    if (m > threshold) break;
    if (m == 0) {
      // Every later level is empty as well: the search runs out at the bound.
      mlsw = max_mlsw;
      break;
    }
    ++mlsw;
    const uint64_t cnt = m * k;
    SW_REQUIRE(cnt < (1ull << 31) && vcount + cnt < (1ull << 31), "Reduce level beyond 2^31 sets");
    b.cand.ensure(cnt * rb);
    b.uniq.ensure(cnt * rb);
    b.next.ensure(cnt * rb);
    b.idx.ensure(cnt * 4 + 4);
    b.uidx.ensure(cnt * 4 + 4);
    b.keep.ensure(cnt * 4 + 4);
    b.pos.ensure(cnt * 4 + 4);
    b.slot.ensure(cnt * 4 + 4);
    uint64_t slots = 1024;
    while (slots < 2 * cnt) slots <<= 1;
    b.table.ensure(slots * 4);
    size_t tb = 0;
    SW_CHECK_HIP(keep_scan_temp_bytes(cnt, &tb));
    if (tb > b.scan_bytes) {
      b.scan.release();
      b.scan.ensure(tb);
      b.scan_bytes = tb;
    }
    SW_CHECK_HIP(launch_expand_sets(int(W), b.list.as<uint64_t>(), m, b.tables.as<uint64_t>(), slices, nullptr, 0, k,
        nullptr, n, 0, b.cand.as<uint64_t>(), nullptr, st));
    // Distinct images, the first of equal ones kept: uniq[j] = cand[uidx[j]].
    uint32_t* total = b.small.as<uint32_t>();
    SW_CHECK_HIP(launch_fill_u32(b.table.as<uint32_t>(), kEmpty, slots, st));
    SW_CHECK_HIP(launch_hash_insert(int(W), b.cand.as<uint64_t>(), nullptr, cnt, b.table.as<uint32_t>(), slots - 1,
        b.cand.as<uint64_t>(), b.slot.as<uint32_t>(), nullptr, st));
    SW_CHECK_HIP(launch_keep_scan(1, b.slot.as<uint32_t>(), b.table.as<uint32_t>(), nullptr, cnt, b.keep.as<uint32_t>(),
        b.pos.as<uint32_t>(), total, b.scan.p, &b.scan_bytes, st));
    SW_CHECK_HIP(launch_select(int(W), b.cand.as<uint64_t>(), 0, b.keep.as<uint32_t>(), b.pos.as<uint32_t>(), cnt,
        b.uniq.as<uint64_t>(), b.uidx.as<uint32_t>(), st));
    SW_CHECK_HIP(hipMemcpyAsync(hs, total, 4, hipMemcpyDeviceToHost, st));
    SW_CHECK_HIP(hipStreamSynchronize(st));
    const uint64_t cu = uint32_t(hs[0]);
    const size_t fb = vsub_flag_bytes(int(W), cu);
    b.flags.ensure(fb);
    SW_CHECK_HIP(hipMemsetAsync(b.flags.p, 0, fb, st));
    if (vcount) {
      // An image containing a set of an earlier level (equal included).
      VsubParams p;
      p.stream = b.visited.as<uint64_t>();
      p.s_count = vcount;
      p.tile = b.uniq.as<uint64_t>();
      p.t_count = cu;
      p.n = n;
      p.mode = VSUB_FLAGS;
      p.flags = b.flags.p;
      apply_tuning(ctx, int(W), p);
      run_subset(ctx, int(W), p);
    }
    {
      // An image properly containing another image of the level.
      VsubParams p;
      p.stream = b.uniq.as<uint64_t>();
      p.s_count = cu;
      p.tile = b.uniq.as<uint64_t>();
      p.t_count = cu;
      p.n = n;
      p.mode = VSUB_FLAGS;
      p.proper_by_size = 1;
      p.flags = b.flags.p;
      apply_tuning(ctx, int(W), p);
      run_subset(ctx, int(W), p);
    }
    SW_CHECK_HIP(launch_keep_scan(2, nullptr, nullptr, b.flags.p, cu, b.keep.as<uint32_t>(), b.pos.as<uint32_t>(),
        total, b.scan.p, &b.scan_bytes, st));
    SW_CHECK_HIP(launch_select(int(W), b.uniq.as<uint64_t>(), 0, b.keep.as<uint32_t>(), b.pos.as<uint32_t>(), cu,
        b.next.as<uint64_t>(), b.idx.as<uint32_t>(), st));
    unsigned long long* first = reinterpret_cast<unsigned long long*>(b.small.as<char>() + 8);
    SW_CHECK_HIP(hipMemsetAsync(first, 0xff, 8, st));
    SW_CHECK_HIP(hipMemcpyAsync(hs, total, 4, hipMemcpyDeviceToHost, st));
    SW_CHECK_HIP(hipStreamSynchronize(st));
    const uint64_t mnew = uint32_t(hs[0]);
    SW_CHECK_HIP(launch_first_singleton(b.next.as<uint64_t>(), mnew, W, first, st));
    std::vector<uint32_t> kept(mnew), uidx(cu);
    if (mnew) SW_CHECK_HIP(hipMemcpyAsync(kept.data(), b.idx.p, mnew * 4, hipMemcpyDeviceToHost, st));
    if (cu) SW_CHECK_HIP(hipMemcpyAsync(uidx.data(), b.uidx.p, cu * 4, hipMemcpyDeviceToHost, st));
    SW_CHECK_HIP(hipMemcpyAsync(hs + 1, first, 8, hipMemcpyDeviceToHost, st));
    // The level joins the earlier ones (the visited list).
    b.visited.ensure((vcount + mnew) * rb + 8, st, vcount * rb);
    if (mnew)
      SW_CHECK_HIP(hipMemcpyAsync(b.visited.as<char>() + vcount * rb, b.next.p, mnew * rb, hipMemcpyDeviceToDevice,
          st));
    SW_CHECK_HIP(hipStreamSynchronize(st));
    for (uint32_t& c : kept) c = uidx[c];  // index among the distinct images -> candidate index
    rs.level_cand.push_back(std::move(kept));
    rs.level_m.push_back(m);
    vcount += mnew;
    std::swap(b.list.p, b.next.p);
    std::swap(b.list.cap, b.next.cap);
    m = mnew;
    if (hs[1] != ~0ull) {
      single = int64_t(hs[1]);
      break;
    }
  }
  rs.steps = mlsw;
  if (single >= 0) {
    rs.decided = true;
    rs.length = mlsw;
    rs.word = rs.prefix(uint32_t(single));
    return;
  }
  if (mlsw == max_mlsw) {
    // No word shorter than the upper bound: its word is the answer.
    rs.decided = true;
    rs.length = mlsw + 1;
    return;
  }

  // States reachable from the level's sets (reduce.hpp:165-188).
  unsigned long long* un = reinterpret_cast<unsigned long long*>(b.small.as<char>() + 16);
  SW_CHECK_HIP(hipMemsetAsync(un, 0, 8 * W, st));
  SW_CHECK_HIP(launch_or_rows(b.list.as<uint64_t>(), m, W, un, st));
  std::vector<uint64_t> uw(W);
  SW_CHECK_HIP(hipMemcpyAsync(uw.data(), un, 8 * W, hipMemcpyDeviceToHost, st));
  SW_CHECK_HIP(hipStreamSynchronize(st));
  std::vector<char> reach(n, 0);
  std::vector<uint32_t> stack;
  for (uint32_t s = 0; s < n; ++s)
    if ((uw[s >> 6] >> (s & 63)) & 1) {
      reach[s] = 1;
      stack.push_back(s);
    }
  while (!stack.empty()) {
    const uint32_t s = stack.back();
    stack.pop_back();
    for (uint32_t a = 0; a < k; ++a) {
      const uint32_t t = d[size_t(s) * k + a];
      if (!reach[t]) {
        reach[t] = 1;
        stack.push_back(t);
      }
    }
  }
  uint32_t rn = 0;
  std::vector<uint16_t> map(n, 0xffff);
  for (uint32_t s = 0; s < n; ++s)
    if (reach[s]) map[s] = uint16_t(rn++);
  if (rn == n || rn == 0) return;  // nothing to reduce (reduce.hpp:131-133)

  // The reduced automaton (states renumbered in order) and the level on it.
  // The renumbering is one-to-one on the kept states, which hold every
  // state of the level, so the level's sets stay distinct.
  rs.reduced = true;
  rs.n_red = rn;
  rs.delta_red.assign(size_t(rn) * k, 0);
  for (uint32_t s = 0; s < n; ++s)
    if (reach[s])
      for (uint32_t a = 0; a < k; ++a) rs.delta_red[size_t(map[s]) * k + a] = map[d[size_t(s) * k + a]];
  const uint32_t W2 = (rn + 63) / 64;
  b.map.ensure(size_t(n) * 2 + 8);
  SW_CHECK_HIP(hipMemcpyAsync(b.map.p, map.data(), size_t(n) * 2, hipMemcpyHostToDevice, st));
  rs.list.ensure(m * uint64_t(W2) * 8 + 8);
  SW_CHECK_HIP(launch_remap_sets(b.list.as<uint64_t>(), m, W, b.map.as<uint16_t>(), n, rs.list.as<uint64_t>(), W2,
      st));
  SW_CHECK_HIP(hipStreamSynchronize(st));
  rs.list_count = m;
}

void distinct_rows(sw_ctx* ctx, uint32_t W, const uint64_t* host_rows, uint64_t count, DevBuf& out,
    uint64_t& out_count, std::vector<uint32_t>& src) {
  use(ctx);
  hipStream_t st = ctx->stream;
  ReduceBufs& b = ctx->reduce;
  const uint64_t rb = uint64_t(W) * 8;
  SW_REQUIRE(count && count < (1ull << 31), "list of 1 to 2^31 - 1 sets");
  b.cand.ensure(count * rb);
  SW_CHECK_HIP(hipMemcpyAsync(b.cand.p, host_rows, count * rb, hipMemcpyHostToDevice, st));
  uint64_t slots = 1024;
  while (slots < 2 * count) slots <<= 1;
  b.table.ensure(slots * 4);
  b.slot.ensure(count * 4 + 4);
  b.keep.ensure(count * 4 + 4);
  b.pos.ensure(count * 4 + 4);
  b.idx.ensure(count * 4 + 4);
  b.small.ensure(128);
  size_t tb = 0;
  SW_CHECK_HIP(keep_scan_temp_bytes(count, &tb));
  if (tb > b.scan_bytes) {
    b.scan.release();
    b.scan.ensure(tb);
    b.scan_bytes = tb;
  }
  out.ensure(count * rb + 8);
  uint32_t* total = b.small.as<uint32_t>();
  SW_CHECK_HIP(launch_fill_u32(b.table.as<uint32_t>(), kEmpty, slots, st));
  SW_CHECK_HIP(launch_hash_insert(int(W), b.cand.as<uint64_t>(), nullptr, count, b.table.as<uint32_t>(), slots - 1,
      b.cand.as<uint64_t>(), b.slot.as<uint32_t>(), nullptr, st));
  SW_CHECK_HIP(launch_keep_scan(1, b.slot.as<uint32_t>(), b.table.as<uint32_t>(), nullptr, count,
      b.keep.as<uint32_t>(), b.pos.as<uint32_t>(), total, b.scan.p, &b.scan_bytes, st));
  SW_CHECK_HIP(launch_select(int(W), b.cand.as<uint64_t>(), 0, b.keep.as<uint32_t>(), b.pos.as<uint32_t>(), count,
      out.as<uint64_t>(), b.idx.as<uint32_t>(), st));
  uint64_t* hs = b.host();
  SW_CHECK_HIP(hipMemcpyAsync(hs, total, 4, hipMemcpyDeviceToHost, st));
  SW_CHECK_HIP(hipStreamSynchronize(st));
  out_count = uint32_t(hs[0]);
  src.resize(out_count);
  SW_CHECK_HIP(hipMemcpyAsync(src.data(), b.idx.p, out_count * 4, hipMemcpyDeviceToHost, st));
  SW_CHECK_HIP(hipStreamSynchronize(st));
}

}  // namespace sw
