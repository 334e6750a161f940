// Choice between the two subset kernels and the stream-index pre-pass of the
// indexed one (k_subset_ix.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "sw_host.hpp"
#include "sw_kernels.hpp"

namespace sw {

void run_subset(sw_ctx* ctx, int W, VsubParams& p) {
  hipStream_t st = ctx->stream;
  const uint64_t tests = p.s_count * p.t_count;
  const uint64_t per_set = (uint64_t(p.n) + 3) & ~3ull;  // index entries per set, at most
  const bool indexed = ctx->ix_enabled && p.s_count && p.t_count && tests >= ctx->ix_min_pairs &&
      vsub_ix_fits(p.n, W, p.proper_by_size != 0) && (p.s_count + 1) * per_set < (1ull << 32) &&
      p.s_count < (1ull << 31);
  if (!indexed) {
    SW_CHECK_HIP(launch_vsub(W, p, st));
    return;
  }
  if (ctx->ix_walk == 1) {
    p.ix_early_exit = 0;
    p.order_states = 0;
    p.sort_by_size = 0;
    p.bucket_anchor = 0;
  } else if (ctx->ix_walk == 2) {
    p.ix_early_exit = 1;
    p.order_states = 1;
    p.sort_by_size = 1;
    p.bucket_anchor = 0;
  } else if (ctx->ix_walk == 3) {
    p.ix_early_exit = 0;
    p.order_states = 0;
    p.sort_by_size = 1;
    p.bucket_anchor = 0;
  } else if (ctx->ix_walk == 4) {
    p.ix_early_exit = 1;
    p.order_states = 1;
    p.sort_by_size = 1;
    p.bucket_anchor = 1;
  }
#ifndef SW_SORT_COUNTS
#define SW_SORT_COUNTS 1
#endif
// Launches of at least 2^SW_IX_SORT_LOG2 pairs sort both lists by size;
// anchor buckets from 2^SW_IX_ANCHOR_LOG2 pairs (64: never).  Measured on the
// batch's overflow searches (n <= 160, lists of 10^4..10^6 sets: 9 % less
// time than with both at 2^34, the bucket pass's round trip no longer paid
// below 2^38 pairs) and on the n = 300 searches (unchanged).
#ifndef SW_IX_SORT_LOG2
#define SW_IX_SORT_LOG2 30
#endif
#ifndef SW_IX_ANCHOR_LOG2
#define SW_IX_ANCHOR_LOG2 38
#endif
  // Counts walk every pair's states: sorted by size, the two sets of a pair
  // are of like size, so the shorter one reads fewer padding rows.
  if (SW_SORT_COUNTS && ctx->ix_walk == 0 && p.mode == VSUB_COUNT) p.sort_by_size = 1;
  const uint64_t s = p.s_count, pairs = (s + 1) / 2;
  // Sorting costs two passes over each list and a few launches; the
  // workgroups' early stop at their tile's largest set pays for it from the
  // indexed kernel's smallest launches on.
  const bool want_sort =
      p.sort_by_size && (ctx->ix_walk >= 2 || tests >= (1ull << SW_IX_SORT_LOG2)) && p.t_count < (1ull << 32);
  if (want_sort && p.diag != INT64_MIN && p.distinct_sets && vsub_ix_fits(p.n, W, true)) {
    p.diag = INT64_MIN;
    p.proper_by_size = 1;
  }
  const bool sorted = want_sort && p.diag == INT64_MIN;
  p.s_perm = p.t_perm = p.s_bound = p.t_bound = nullptr;
  p.work = nullptr;
  p.work_count = 0;
  bool ranked = false;  // rank / order of the states computed already
  uint32_t* f = nullptr;
  if (p.bucket_anchor && (ctx->ix_walk == 0 || ctx->ix_walk == 4) && p.diag == INT64_MIN &&
      (p.mode == VSUB_ANY || p.mode == VSUB_FLAGS || p.mode == VSUB_SFLAGS) &&
      (ctx->ix_walk == 4 || (SW_IX_ANCHOR_LOG2 < 64 && tests >= (1ull << (SW_IX_ANCHOR_LOG2 & 63)))) &&
      s < (1ull << 32) && p.t_count < (1ull << 32)) {
    // ---- anchor buckets
    const uint32_t n = p.n, n1 = n + 1;
    ctx->ix_order.ensure(3 * 512 * 4);
    f = ctx->ix_order.as<uint32_t>();  // freq, rank, order
    ctx->ix_sort.ensure(size_t(5) * n1 * 4);
    uint32_t* ha = ctx->ix_sort.as<uint32_t>();  // stream sets per anchor
    uint32_t* ca = ha + n1;                      // their cursor
    uint32_t* ct = ca + n1;                      // inverted-list cursor
    SW_CHECK_HIP(launch_fill_u32(f, 0u, 512, st));
    SW_CHECK_HIP(launch_fill_u32(ha, 0u, n1, st));
    SW_CHECK_HIP(launch_state_freq(W, p.tile, p.t_count, n, p.complement_tile, f, st));
    SW_CHECK_HIP(launch_state_rank(f, n, f + 512, f + 1024, st));
    ranked = true;
    ctx->ix_anchor.ensure(s * 4 + 16);
    SW_CHECK_HIP(launch_anchor(W, p.stream, s, n, p.complement_stream, f + 512, ctx->ix_anchor.as<uint32_t>(), ha, st));
    ctx->ix_counts.resize(2 * size_t(n1));
    SW_CHECK_HIP(hipMemcpyAsync(ctx->ix_counts.data(), ha, n1 * 4, hipMemcpyDeviceToHost, st));
    SW_CHECK_HIP(hipMemcpyAsync(ctx->ix_counts.data() + n1, f, n * 4, hipMemcpyDeviceToHost, st));
    SW_CHECK_HIP(hipStreamSynchronize(st));
    const uint32_t* hcount = ctx->ix_counts.data();  // stream sets per anchor (n: empty sets)
    const uint32_t* freq = hcount + n1;              // tile sets per state
    uint64_t lists = 0;
    for (uint32_t q = 0; q < n; ++q) lists += freq[q];
    if (hcount[n] == 0 && lists < (1ull << 32)) {
      SW_CHECK_HIP(launch_size_offsets(ha, n, ca, nullptr, st));
      ctx->ix_sperm.ensure(s * 4 + 16);
      SW_CHECK_HIP(launch_key_scatter(ctx->ix_anchor.as<uint32_t>(), s, n1, ca, ctx->ix_sperm.as<uint32_t>(), st));
      SW_CHECK_HIP(launch_size_offsets(f, n - 1, ct, nullptr, st));  // list q starts at the count below q
      ctx->ix_tlist.ensure(lists * 4 + 16);
      SW_CHECK_HIP(launch_invert(W, p.tile, p.t_count, n, p.complement_tile, ct, ctx->ix_tlist.as<uint32_t>(), st));
      // Work items: each anchor's stream sets (chunks of at most 2^17, pair
      // aligned) against its inverted list (tiles of 4096).
      std::vector<uint4> work;
      uint64_t s0 = 0, l0 = 0;
      const uint64_t chunk = uint64_t(1) << 17;
      for (uint32_t q = 0; q < n; ++q) {
        const uint64_t s1 = s0 + hcount[q], l1 = l0 + freq[q];
        if (hcount[q] && freq[q]) {
          for (uint64_t a = s0 & ~1ull; a < s1; a += chunk) {
            uint64_t b = std::min<uint64_t>(a + chunk, s1);
            b = std::min<uint64_t>((b + 1) & ~1ull, s);
            for (uint64_t t = l0; t < l1; t += 4096)
              work.push_back(make_uint4(uint32_t(t), uint32_t(std::min<uint64_t>(4096, l1 - t)), uint32_t(a),
                  uint32_t(b)));
          }
        }
        s0 = s1;
        l0 = l1;
      }
      ctx->ix_work.ensure(work.size() * sizeof(uint4) + 16);
      if (!work.empty())
        SW_CHECK_HIP(hipMemcpyAsync(ctx->ix_work.p, work.data(), work.size() * sizeof(uint4), hipMemcpyHostToDevice,
            st));
      SW_CHECK_HIP(hipStreamSynchronize(st));  // `work` leaves scope
      p.s_perm = ctx->ix_sperm.as<uint32_t>();
      p.t_perm = ctx->ix_tlist.as<uint32_t>();
      p.work = ctx->ix_work.as<uint4>();
      p.work_count = work.size();
      if (work.empty()) return;  // no stream set's anchor is in any tile set: no pair
    }
  }
  if (!p.work && sorted) {
    const uint32_t n1 = p.n + 1;
    ctx->ix_sort.ensure(size_t(5) * n1 * 4);
    uint32_t* hs = ctx->ix_sort.as<uint32_t>();  // stream histogram, then tile cursor
    uint32_t* ht = hs + n1;                      // tile histogram
    uint32_t* cs = ht + n1;                      // stream cursor
    uint32_t* bs = cs + n1;                      // stream sets of at most c states
    uint32_t* bt = bs + n1;                      // tile sets of at most c states
    ctx->ix_sperm.ensure(s * 4 + 16);
    ctx->ix_tperm.ensure(p.t_count * 4 + 16);
    SW_CHECK_HIP(launch_fill_u32(hs, 0u, 2 * n1, st));
    SW_CHECK_HIP(launch_size_hist(W, p.stream, s, p.n, p.complement_stream, hs, st));
    SW_CHECK_HIP(launch_size_hist(W, p.tile, p.t_count, p.n, p.complement_tile, ht, st));
    SW_CHECK_HIP(launch_size_offsets(hs, p.n, cs, bs, st));
    SW_CHECK_HIP(launch_size_offsets(ht, p.n, hs, bt, st));  // tile cursor over the stream histogram
    SW_CHECK_HIP(launch_size_scatter(W, p.stream, s, p.n, p.complement_stream, cs, ctx->ix_sperm.as<uint32_t>(), st));
    SW_CHECK_HIP(launch_size_scatter(W, p.tile, p.t_count, p.n, p.complement_tile, hs, ctx->ix_tperm.as<uint32_t>(),
        st));
    p.s_perm = ctx->ix_sperm.as<uint32_t>();
    p.t_perm = ctx->ix_tperm.as<uint32_t>();
    p.s_bound = bs;
    p.t_bound = bt;
  }
  ctx->ix_padded.ensure(pairs * 4 + 16);  // entries per pair
  ctx->ix_size.ensure(s * 4 + 16);
  ctx->ix_start.ensure((pairs + 1) * 4 + 16);
  if (!ctx->ix_host) SW_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&ctx->ix_host), 16, hipHostMallocDefault));
  SW_CHECK_HIP(launch_ix_pair_count(W, p.stream, p.s_perm, s, p.n, p.complement_stream, ctx->ix_size.as<uint32_t>(),
      ctx->ix_padded.as<uint32_t>(), st));
  size_t need = 0;
  SW_CHECK_HIP(launch_scan(nullptr, nullptr, pairs, nullptr, nullptr, &need, st));
  if (need > ctx->ix_tmp_bytes) {
    ctx->ix_tmp.release();
    ctx->ix_tmp.ensure(need);
    ctx->ix_tmp_bytes = need;
  }
  uint32_t* start = ctx->ix_start.as<uint32_t>();
  SW_CHECK_HIP(launch_scan(ctx->ix_padded.as<uint32_t>(), start, pairs, start + pairs, ctx->ix_tmp.p,
      &ctx->ix_tmp_bytes, st));
  // The index length: a pair holds at most 8 ceil(n / 4) entries.  While that
  // bound fits in 1 GiB the list is sized by it and nothing is read back (a
  // round trip per launch is a good part of a small search's step); past it,
  // the exact length is read back.
  const uint64_t bound = pairs * 8 * ((uint64_t(p.n) + 3) / 4);
  if (bound * 4 <= (uint64_t(1) << 30)) {
    // Growing frees the old list: the stream is idle first.
    if (bound * 4 + 64 > ctx->ix_list.cap) SW_CHECK_HIP(hipStreamSynchronize(st));
    ctx->ix_list.ensure(bound * 4 + 64);
  } else {
    SW_CHECK_HIP(hipMemcpyAsync(ctx->ix_host, start + pairs, 4, hipMemcpyDeviceToHost, st));
    SW_CHECK_HIP(hipStreamSynchronize(st));
    const uint64_t total = ctx->ix_host[0];
    ctx->ix_list.ensure(total * 4 + 64);
  }
  const uint32_t* rank = nullptr;
  const uint32_t* order = nullptr;
  if (p.order_states || ranked) {
    if (!ranked) {
      ctx->ix_order.ensure(3 * 512 * 4);
      f = ctx->ix_order.as<uint32_t>();
      SW_CHECK_HIP(launch_fill_u32(f, 0u, 512, st));
      SW_CHECK_HIP(launch_state_freq(W, p.tile, p.t_count, p.n, p.complement_tile, f, st));
      SW_CHECK_HIP(launch_state_rank(f, p.n, f + 512, f + 1024, st));
    }
    rank = f + 512;
    order = f + 1024;
  }
  uint32_t* head = nullptr;
  if (p.ix_early_exit) {
    ctx->ix_head.ensure(pairs * size_t(ix_head_groups()) * 32 + 64);
    head = ctx->ix_head.as<uint32_t>();
  }
  SW_CHECK_HIP(launch_ix_pair_fill(W, p.stream, p.s_perm, s, p.n, p.complement_stream, start, rank, order,
      ctx->ix_list.as<uint32_t>(), head, st));
  p.ix_head = head;
  p.ix_start = start;
  p.ix_list = ctx->ix_list.as<uint32_t>();
  p.ix_size = ctx->ix_size.as<uint32_t>();
  SW_CHECK_HIP(launch_vsub_ix(W, p, st));
}

}  // namespace sw
