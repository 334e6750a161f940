// Eppstein's greedy bound on the device for one automaton or a batch (the
// reference's Eppstein<N,K>::run, synchrolib/algorithm/eppstein/eppstein.hpp:
// 25-114): non-synchronizing detection by the pair BFS, and, when every pair
// merges, the greedy word and its length.  One workgroup per automaton
// (k_eppstein.hip); batches are cut so that their scratch stays under 2 GiB.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "sw_host.hpp"
#include "sw_kernels.hpp"

namespace sw {

void eppstein_batch(sw_ctx* ctx, uint32_t count, const uint32_t* ns, const uint32_t* ks, const uint64_t* delta_offsets,
    const uint32_t* deltas, const uint64_t* uppers, uint64_t word_cap, int32_t* non_synchro, int32_t* improved,
    uint64_t* bounds, uint64_t* word_lens, uint32_t* words) {
  hipStream_t st = ctx->stream;
  for (uint32_t i = 0; i < count; ++i) {
    SW_REQUIRE(ns[i] >= 1 && ns[i] <= kEppMaxN, "eppstein: n must be in [1, 512]");
    SW_REQUIRE(ks[i] >= 1 && ks[i] <= 255, "eppstein: k must be in [1, 255]");
  }
  const size_t budget = size_t(2) << 30;
  uint32_t at = 0;
  while (at < count) {
    std::vector<EppJob> jobs;
    std::vector<uint32_t> dpack;
    size_t scratch = 0;
    uint32_t end = at;
    while (end < count) {
      const size_t need = eppstein_scratch_bytes(ns[end], ks[end]);
      if (!jobs.empty() && scratch + need > budget) break;
      jobs.push_back(EppJob{ns[end], ks[end], dpack.size(), scratch, uint64_t(end - at) * word_cap, word_cap,
          uppers[end]});
      dpack.insert(dpack.end(), deltas + delta_offsets[end], deltas + delta_offsets[end] + size_t(ns[end]) * ks[end]);
      scratch += need;
      ++end;
    }
    const size_t m = jobs.size();
    ctx->epp_jobs.ensure(m * sizeof(EppJob));
    ctx->epp_delta.ensure(dpack.size() * 4 + 4);
    ctx->epp_scratch.ensure(scratch);
    ctx->epp_words.ensure(m * word_cap * 4 + 4);
    ctx->epp_out.ensure(m * sizeof(EppOut));
    SW_CHECK_HIP(hipMemcpyAsync(ctx->epp_jobs.p, jobs.data(), m * sizeof(EppJob), hipMemcpyHostToDevice, st));
    SW_CHECK_HIP(hipMemcpyAsync(ctx->epp_delta.p, dpack.data(), dpack.size() * 4, hipMemcpyHostToDevice, st));
    SW_CHECK_HIP(launch_eppstein_wg(ctx->epp_jobs.as<EppJob>(), uint32_t(m), ctx->epp_delta.as<uint32_t>(),
        ctx->epp_scratch.as<unsigned char>(), ctx->epp_words.as<uint32_t>(), ctx->epp_out.as<EppOut>(), st));
    std::vector<EppOut> out(m);
    SW_CHECK_HIP(hipMemcpyAsync(out.data(), ctx->epp_out.p, m * sizeof(EppOut), hipMemcpyDeviceToHost, st));
    if (words && word_cap)
      SW_CHECK_HIP(hipMemcpyAsync(words + uint64_t(at) * word_cap, ctx->epp_words.p, m * word_cap * 4,
          hipMemcpyDeviceToHost, st));
    SW_CHECK_HIP(hipStreamSynchronize(st));
    for (size_t j = 0; j < m; ++j) {
      const uint32_t i = at + uint32_t(j);
      if (non_synchro) non_synchro[i] = out[j].non_synchro;
      if (improved) improved[i] = out[j].improved;
      if (bounds) bounds[i] = out[j].bound;
      if (word_lens) word_lens[i] = out[j].word_len;
    }
    at = end;
  }
}

}  // namespace sw

extern "C" {

int sw_eppstein(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta, uint64_t upper, int32_t* non_synchro,
    uint64_t* bound, uint32_t* word, uint64_t word_cap, uint64_t* word_len) {
  using namespace sw;
  try {
    SW_REQUIRE(ctx && delta && non_synchro && bound, "NULL argument");
    use(ctx);
    const uint64_t off = 0;
    int32_t imp = 0;
    uint64_t wl = 0;
    std::vector<uint32_t> w(word_cap ? word_cap : 1);
    eppstein_batch(ctx, 1, &n, &k, &off, delta, &upper, word_cap, non_synchro, &imp, bound, &wl, w.data());
    if (word_len) *word_len = imp ? wl : 0;
    if (word && imp) std::copy(w.begin(), w.begin() + std::min<uint64_t>(wl, word_cap), word);
    return SW_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  }
}

int sw_eppstein_batch(sw_ctx* ctx, uint32_t count, const uint32_t* ns, const uint32_t* ks,
    const uint64_t* delta_offsets, const uint32_t* deltas, const uint64_t* uppers, uint32_t word_cap,
    int32_t* non_synchro, uint64_t* bounds, uint64_t* word_lens, uint32_t* words) {
  using namespace sw;
  try {
    SW_REQUIRE(ctx && (count == 0 || (ns && ks && delta_offsets && deltas && uppers && non_synchro && bounds)),
        "NULL argument");
    use(ctx);
    std::vector<int32_t> imp(count);
    eppstein_batch(ctx, count, ns, ks, delta_offsets, deltas, uppers, word_cap, non_synchro, imp.data(), bounds,
        word_lens, words);
    if (word_lens)
      for (uint32_t i = 0; i < count; ++i)
        if (!imp[i]) word_lens[i] = 0;
    return SW_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  }
}

}  // extern "C"
