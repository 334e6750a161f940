// Power-set BFS on the device (the reference's Brute<N,K>::run,
// synchrolib/algorithm/brute/brute.hpp:28-122), for one automaton or a batch.
//
// Automata whose visited bitmap (2^n bits) and image tables fit LDS run one
// workgroup each, every BFS level inside one launch (k_brute_wg); larger ones
// (up to 32 states, at most kBruteLevelMaxK letters) run one launch per level
// over the whole device (k_brute_level) with the bitmap in device memory.
//
// The reference keeps distances in 8 bits (its INF is 255), so lengths of 255
// and more are not found there; these kernels keep no distances and find them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "sw_host.hpp"
#include "sw_kernels.hpp"

namespace sw {

namespace {

bool fits_lds(uint32_t n, uint32_t k) { return size_t(brute_lds_bytes(n, k)) <= kBruteMaxLds; }

// One automaton of up to 32 states, one launch per level.
int64_t brute_levels(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta_host) {
  SW_REQUIRE(k <= uint32_t(kBruteLevelMaxK), "brute: more than 8 letters for an automaton whose bitmap is not in LDS");
  hipStream_t st = ctx->stream;
  const size_t subsets = size_t(1) << n;
  DevBuf d, cur, nxt, map, state;
  d.ensure(size_t(n) * k * 4);
  cur.ensure(subsets * 4);
  nxt.ensure(subsets * 4);
  map.ensure(std::max<size_t>(subsets / 8, 4));
  state.ensure(16);
  SW_CHECK_HIP(hipMemcpyAsync(d.p, delta_host, size_t(n) * k * 4, hipMemcpyHostToDevice, st));
  SW_CHECK_HIP(launch_fill_u32(map.as<uint32_t>(), 0u, std::max<size_t>(subsets / 32, 1), st));
  const uint32_t full = n == 32 ? 0xffffffffu : ((1u << n) - 1u);
  // state: [0] subsets in cur, [1] new subsets, [2] singleton reached.
  uint32_t init[4] = {1u, 0u, 0u, 0u};
  const uint32_t bit = 1u << (full & 31u);
  SW_CHECK_HIP(hipMemcpyAsync(cur.p, &full, 4, hipMemcpyHostToDevice, st));
  SW_CHECK_HIP(hipMemcpyAsync(map.as<uint32_t>() + (full >> 5), &bit, 4, hipMemcpyHostToDevice, st));
  SW_CHECK_HIP(hipMemcpyAsync(state.p, init, 16, hipMemcpyHostToDevice, st));
  uint32_t fc = 1;
  uint32_t upd[2] = {0u, 0u};  // read by the copy below until the next level's synchronisation
  for (int64_t level = 0;; ++level) {
    uint32_t* s = state.as<uint32_t>();
    SW_CHECK_HIP(launch_brute_level(n, k, d.as<uint32_t>(), cur.as<uint32_t>(), s, fc, nxt.as<uint32_t>(),
        map.as<uint32_t>(), s + 1, st));
    uint32_t back[2] = {0, 0};
    SW_CHECK_HIP(hipMemcpyAsync(back, s + 1, 8, hipMemcpyDeviceToHost, st));
    SW_CHECK_HIP(hipStreamSynchronize(st));
    if (back[1]) return level + 1;
    if (back[0] == 0) return -1;
    fc = back[0];
    upd[0] = fc;
    upd[1] = 0u;
    SW_CHECK_HIP(hipMemcpyAsync(s, upd, 8, hipMemcpyHostToDevice, st));
    std::swap(cur.p, nxt.p);
    std::swap(cur.cap, nxt.cap);
  }
}

}  // namespace

void brute_batch(sw_ctx* ctx, uint32_t count, const uint32_t* ns, const uint32_t* ks, const uint64_t* delta_offsets,
    const uint32_t* deltas, int64_t* lengths) {
  hipStream_t st = ctx->stream;
  std::vector<uint32_t> small;
  for (uint32_t i = 0; i < count; ++i) {
    SW_REQUIRE(ns[i] >= 1 && ns[i] <= 32, "brute: n must be in [1, 32]");
    SW_REQUIRE(ks[i] >= 1 && ks[i] <= 255, "brute: k must be in [1, 255]");
    if (ns[i] == 1) lengths[i] = 0;
    else if (fits_lds(ns[i], ks[i])) small.push_back(i);
    else lengths[i] = brute_levels(ctx, ns[i], ks[i], deltas + delta_offsets[i]);
  }
  // The small ones in launches of at most ~1 GiB of frontier buffers.
  const size_t budget = size_t(1) << 28;  // subsets (4 bytes each)
  size_t at = 0;
  while (at < small.size()) {
    std::vector<BruteJob> jobs;
    std::vector<uint32_t> dpack;
    size_t frontier = 0, lds = 0;
    size_t end = at;
    while (end < small.size()) {
      const uint32_t i = small[end];
      const size_t need = size_t(2) << ns[i];
      if (!jobs.empty() && frontier + need > budget) break;
      jobs.push_back(BruteJob{ns[i], ks[i], dpack.size(), frontier});
      dpack.insert(dpack.end(), deltas + delta_offsets[i], deltas + delta_offsets[i] + size_t(ns[i]) * ks[i]);
      frontier += need;
      lds = std::max(lds, size_t(brute_lds_bytes(ns[i], ks[i])));
      ++end;
    }
    ctx->brute_jobs.ensure(jobs.size() * sizeof(BruteJob));
    ctx->brute_delta.ensure(dpack.size() * 4 + 4);
    ctx->brute_frontier.ensure(frontier * 4);
    ctx->brute_out.ensure(jobs.size() * 8);
    SW_CHECK_HIP(hipMemcpyAsync(ctx->brute_jobs.p, jobs.data(), jobs.size() * sizeof(BruteJob), hipMemcpyHostToDevice,
        st));
    SW_CHECK_HIP(hipMemcpyAsync(ctx->brute_delta.p, dpack.data(), dpack.size() * 4, hipMemcpyHostToDevice, st));
    SW_CHECK_HIP(launch_brute_wg(ctx->brute_jobs.as<BruteJob>(), uint32_t(jobs.size()), ctx->brute_delta.as<uint32_t>(),
        ctx->brute_frontier.as<uint32_t>(), ctx->brute_out.as<int64_t>(), lds, st));
    std::vector<int64_t> out(jobs.size());
    SW_CHECK_HIP(hipMemcpyAsync(out.data(), ctx->brute_out.p, jobs.size() * 8, hipMemcpyDeviceToHost, st));
    SW_CHECK_HIP(hipStreamSynchronize(st));
    for (size_t j = 0; j < jobs.size(); ++j) lengths[small[at + j]] = out[j];
    at = end;
  }
}

}  // namespace sw

extern "C" {

int sw_brute(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta, int64_t* length) {
  using namespace sw;
  try {
    SW_REQUIRE(ctx && delta && length, "NULL argument");
    use(ctx);
    const uint64_t off = 0;
    brute_batch(ctx, 1, &n, &k, &off, delta, length);
    return SW_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  }
}

int sw_brute_batch(sw_ctx* ctx, uint32_t count, const uint32_t* ns, const uint32_t* ks, const uint64_t* delta_offsets,
    const uint32_t* deltas, int64_t* lengths) {
  using namespace sw;
  try {
    SW_REQUIRE(ctx && (count == 0 || (ns && ks && delta_offsets && deltas && lengths)), "NULL argument");
    use(ctx);
    brute_batch(ctx, count, ns, ks, delta_offsets, deltas, lengths);
    return SW_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  }
}

}  // extern "C"
