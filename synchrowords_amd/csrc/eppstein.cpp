// Eppstein's greedy bound on the device for one automaton or a batch (the
// reference's Eppstein<N,K>::run, synchrolib/algorithm/eppstein/eppstein.hpp:
// 25-114): non-synchronizing detection by the pair BFS, and, when every pair
// merges, the greedy word and its length.  One workgroup per automaton
// (k_eppstein.hip), or, for one automaton alone, the pair BFS over the whole
// device; batches are cut so that their scratch stays under 2 GiB.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "sw_host.hpp"
#include "sw_kernels.hpp"

namespace sw {

// Below 128 states one workgroup takes a lone automaton: the Cerny automaton
// C_64's pair BFS has thousands of levels a few pairs wide, where the
// whole-device path pays its launches per level (57 ms for C_64).
#ifndef SW_EPP_WIDE_MIN_N
#define SW_EPP_WIDE_MIN_N 128
#endif

namespace {

// One automaton over the whole device: the preimage lists built here (sources
// ascending, as the workgroup kernel builds them), the pair BFS a level at a
// time (six launches that read the level from a device state, enqueued
// several levels ahead), then the greedy phase in one workgroup.  A lone
// automaton in one workgroup leaves all but one CU idle; its BFS levels are
// thousands of parents wide.
void eppstein_wide(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta, uint64_t upper, uint64_t word_cap,
    EppOut& res, uint32_t* words) {
  hipStream_t st = ctx->stream;
  const EppLayout lay = epp_layout(n, k);
  const uint32_t kn = k * n, all = n * (n - 1) / 2;
  std::vector<uint32_t> beg(kn + 1, 0);
  std::vector<uint16_t> edges(kn);
  for (uint32_t q = 0; q < n; ++q)
    for (uint32_t a = 0; a < k; ++a) beg[a * n + delta[size_t(q) * k + a] + 1]++;
  for (uint32_t i = 0; i < kn; ++i) beg[i + 1] += beg[i];
  {
    std::vector<uint32_t> fill(beg.begin(), beg.end() - 1);
    for (uint32_t q = 0; q < n; ++q)
      for (uint32_t a = 0; a < k; ++a) edges[fill[a * n + delta[size_t(q) * k + a]]++] = uint16_t(q);
  }
  ctx->epp_scratch.ensure(lay.total);
  unsigned char* base = ctx->epp_scratch.as<unsigned char>();
  SW_CHECK_HIP(hipMemcpyAsync(base, beg.data(), beg.size() * 4, hipMemcpyHostToDevice, st));
  SW_CHECK_HIP(hipMemcpyAsync(base + lay.edges, edges.data(), edges.size() * 2, hipMemcpyHostToDevice, st));
  SW_CHECK_HIP(launch_fill_u32(reinterpret_cast<uint32_t*>(base + lay.info), 0u, size_t(n) * n, st));
  SW_CHECK_HIP(launch_fill_u32(reinterpret_cast<uint32_t*>(base + lay.keys), ~0u, size_t(2) * n * n, st));
  const size_t most = size_t(all) + kn;
  ctx->epp_off.ensure(most * 4 + 4);
  ctx->epp_tmp.ensure(size_t(epp_scan_tiles(n, k)) * 4 + 4);
  ctx->epp_small.ensure(sizeof(EppLevel));
  EppLevel* state = ctx->epp_small.as<EppLevel>();
  const EppLevel init{0, 0, 0, 0, 0};
  SW_CHECK_HIP(hipMemcpyAsync(state, &init, sizeof(EppLevel), hipMemcpyHostToDevice, st));
  // Levels are enqueued kChunk at a time; the state is read after each chunk.
  // Every level finds at least one new pair, so there are at most `all`.
  constexpr uint32_t kChunk = 6;
  EppLevel host{};
  for (uint32_t level = 0;;) {
    for (uint32_t c = 0; c < kChunk; ++c, ++level)
      SW_CHECK_HIP(launch_epp_level(level, n, k, base, ctx->epp_off.as<uint32_t>(), ctx->epp_tmp.as<uint32_t>(),
          state, st));
    SW_CHECK_HIP(hipMemcpyAsync(&host, state, sizeof(EppLevel), hipMemcpyDeviceToHost, st));
    SW_CHECK_HIP(hipStreamSynchronize(st));
    if (host.done) break;
    SW_REQUIRE(level <= all + 1, "eppstein: pair BFS did not end");
  }
  const uint64_t found = host.found;
  if (found < all) {
    res = EppOut{1, 0, upper, 0};
    return;
  }
  const EppJob job{n, k, 0, 0, 0, word_cap, upper};
  ctx->epp_jobs.ensure(sizeof(EppJob));
  ctx->epp_delta.ensure(size_t(kn) * 4 + 4);
  ctx->epp_words.ensure(word_cap * 4 + 4);
  ctx->epp_out.ensure(sizeof(EppOut));
  SW_CHECK_HIP(hipMemcpyAsync(ctx->epp_jobs.p, &job, sizeof(EppJob), hipMemcpyHostToDevice, st));
  SW_CHECK_HIP(hipMemcpyAsync(ctx->epp_delta.p, delta, size_t(kn) * 4, hipMemcpyHostToDevice, st));
  SW_CHECK_HIP(launch_epp_greedy(ctx->epp_jobs.as<EppJob>(), n, ctx->epp_delta.as<uint32_t>(), base,
      ctx->epp_words.as<uint32_t>(), ctx->epp_out.as<EppOut>(), st));
  SW_CHECK_HIP(hipMemcpyAsync(&res, ctx->epp_out.p, sizeof(EppOut), hipMemcpyDeviceToHost, st));
  if (words && word_cap)
    SW_CHECK_HIP(hipMemcpyAsync(words, ctx->epp_words.p, word_cap * 4, hipMemcpyDeviceToHost, st));
  SW_CHECK_HIP(hipStreamSynchronize(st));
}

}  // namespace

void eppstein_batch(sw_ctx* ctx, uint32_t count, const uint32_t* ns, const uint32_t* ks, const uint64_t* delta_offsets,
    const uint32_t* deltas, const uint64_t* uppers, uint64_t word_cap, int32_t* non_synchro, int32_t* improved,
    uint64_t* bounds, uint64_t* word_lens, uint32_t* words) {
  hipStream_t st = ctx->stream;
  for (uint32_t i = 0; i < count; ++i) {
    SW_REQUIRE(ns[i] >= 1 && ns[i] <= kEppMaxN, "eppstein: n must be in [1, 512]");
    SW_REQUIRE(ks[i] >= 1 && ks[i] <= 255, "eppstein: k must be in [1, 255]");
  }
  if (count == 1 && ns[0] >= SW_EPP_WIDE_MIN_N) {
    EppOut r{};
    eppstein_wide(ctx, ns[0], ks[0], deltas + delta_offsets[0], uppers[0], word_cap, r, words);
    if (non_synchro) non_synchro[0] = r.non_synchro;
    if (improved) improved[0] = r.improved;
    if (bounds) bounds[0] = r.bound;
    if (word_lens) word_lens[0] = r.word_len;
    return;
  }
  const size_t budget = size_t(2) << 30;
  uint32_t at = 0;
  while (at < count) {
    std::vector<EppJob> jobs;
    std::vector<uint32_t> dpack;
    size_t scratch = 0;
    uint32_t list_lds = 0, max_n = 0;
    uint32_t end = at;
    while (end < count) {
      const size_t need = eppstein_scratch_bytes(ns[end], ks[end]);
      if (!jobs.empty() && scratch + need > budget) break;
      jobs.push_back(EppJob{ns[end], ks[end], dpack.size(), scratch, uint64_t(end - at) * word_cap, word_cap,
          uppers[end]});
      dpack.insert(dpack.end(), deltas + delta_offsets[end], deltas + delta_offsets[end] + size_t(ns[end]) * ks[end]);
      scratch += need;
      list_lds = std::max(list_lds, eppstein_list_lds(ns[end], ks[end]));
      max_n = std::max(max_n, ns[end]);
      ++end;
    }
    const size_t m = jobs.size();
    ctx->epp_jobs.ensure(m * sizeof(EppJob));
    ctx->epp_delta.ensure(dpack.size() * 4 + 4);
    ctx->epp_scratch.ensure(scratch);
    ctx->epp_words.ensure(m * word_cap * 4 + 4);
    ctx->epp_out.ensure(m * sizeof(EppOut));
    // Copies through the context's pinned staging: jobs and transitions in,
    // results and words out.
    const size_t jb = m * sizeof(EppJob), db = dpack.size() * 4, ob = m * sizeof(EppOut);
    const size_t wb = words && word_cap ? m * word_cap * 4 : 0;
    unsigned char* stage = static_cast<unsigned char*>(stage_bytes(ctx, jb + db + ob + wb));
    std::memcpy(stage, jobs.data(), jb);
    std::memcpy(stage + jb, dpack.data(), db);
    SW_CHECK_HIP(hipMemcpyAsync(ctx->epp_jobs.p, stage, jb, hipMemcpyHostToDevice, st));
    SW_CHECK_HIP(hipMemcpyAsync(ctx->epp_delta.p, stage + jb, db, hipMemcpyHostToDevice, st));
    SW_CHECK_HIP(launch_eppstein_wg(ctx->epp_jobs.as<EppJob>(), uint32_t(m), ctx->epp_delta.as<uint32_t>(),
        ctx->epp_scratch.as<unsigned char>(), ctx->epp_words.as<uint32_t>(), ctx->epp_out.as<EppOut>(), list_lds,
        eppstein_letter_lds(max_n, list_lds), max_n <= 64, st));
    SW_CHECK_HIP(hipMemcpyAsync(stage + jb + db, ctx->epp_out.p, ob, hipMemcpyDeviceToHost, st));
    if (wb) SW_CHECK_HIP(hipMemcpyAsync(stage + jb + db + ob, ctx->epp_words.p, wb, hipMemcpyDeviceToHost, st));
    SW_CHECK_HIP(hipStreamSynchronize(st));
    std::vector<EppOut> out(m);
    std::memcpy(out.data(), stage + jb + db, ob);
    if (wb) std::memcpy(words + uint64_t(at) * word_cap, stage + jb + db + ob, wb);
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
