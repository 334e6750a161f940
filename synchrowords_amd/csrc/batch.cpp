// Batched exact search, host side: one persistent kernel launch over a
// cost-ordered instance queue (k_batch.hip), then the whole-device search
// (exact.cpp) for the instances that outgrew a workgroup.
//
// The reference solves a file of automata one at a time, re-entering its JIT
// per (N, K) (app/main.cpp:49-78, app/jit.hpp:25-56); here one call takes the
// whole batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "sw_batch.hpp"
#include "sw_host.hpp"
#include "sw_kernels.hpp"

namespace sw {

// One automaton of at most 64 states (one word per set) through the batch
// kernel's search in a single workgroup: the whole meet in the middle runs
// inside one launch with its tables in LDS, where the device-wide search
// pays about ten launches and two read-backs per step (the Cerny automaton
// C_64 takes 3968 steps with frontiers of one or two sets).  False when the
// instance outgrows the workgroup (then the device-wide search runs).
bool exact_search_workgroup(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* d, uint64_t upper,
    ExactOutcome& out) {
  if (n < 2 || n > 64 || k > kBatchMaxK || upper < 2 || upper > 4096) return false;
  const auto t0 = std::chrono::steady_clock::now();
  BatchInstance b{};
  b.n = n;
  b.k = k;
  b.delta_off = 0;
  b.upper = uint32_t(upper);
  const std::vector<char> sink = sink_states(n, k, d);
  std::vector<uint32_t> indeg(size_t(n) * k, 0);
  for (uint32_t q = 0; q < n; ++q)
    for (uint32_t a = 0; a < k; ++a) indeg[size_t(d[size_t(q) * k + a]) * k + a]++;
  for (uint32_t q = 0; q < n; ++q) {
    if (!sink[q]) continue;
    bool ok = false;
    for (uint32_t a = 0; a < k; ++a) ok |= indeg[size_t(q) * k + a] > 1;
    if (ok) b.b0[0] |= 1ull << q;
  }
  std::vector<uint8_t> bytes(size_t(n) * k);
  for (size_t j = 0; j < bytes.size(); ++j) bytes[j] = uint8_t(d[j]);
  BatchLimits L;
  L.cap_rows = 1u << 16;
  L.cap_cand = 3u * (1u << 15);
  uint32_t hs = 1024;
  while (hs < 2 * (L.cap_rows + L.cap_cand)) hs <<= 1;
  L.hash_slots = hs;
  L.max_pairs = 1ull << 32;
  const uint32_t wcap = uint32_t(std::max<uint64_t>(upper, 64));
  const size_t stride = batch_scratch_bytes(1, L);
  hipStream_t s = ctx->stream;
  ctx->wg_in.ensure(sizeof(BatchInstance) + 64 + bytes.size() + 64);
  ctx->wg_scratch.ensure(stride);
  ctx->wg_out.ensure(sizeof(BatchResult) + 64 + wcap);
  uint8_t* din = ctx->wg_in.as<uint8_t>();
  BatchInstance* dinst = reinterpret_cast<BatchInstance*>(din);
  uint32_t* dorder = reinterpret_cast<uint32_t*>(din + sizeof(BatchInstance));      // [0] = 0: the one instance
  uint32_t* dhead = dorder + 4;                                                    // queue head
  uint8_t* dbytes = din + sizeof(BatchInstance) + 64;
  BatchResult* dres = ctx->wg_out.as<BatchResult>();
  uint8_t* dwords = ctx->wg_out.as<uint8_t>() + sizeof(BatchResult) + 64;
  // One copy in (instance, queue words, transitions: wg_in's layout) and one
  // out (result, then the word), through pinned staging.
  const size_t in_bytes = sizeof(BatchInstance) + 64 + bytes.size();
  const size_t out_bytes = sizeof(BatchResult) + 64 + wcap;
  uint8_t* stage = static_cast<uint8_t*>(stage_bytes(ctx, in_bytes + out_bytes));
  std::memset(stage, 0, sizeof(BatchInstance) + 64);
  std::memcpy(stage, &b, sizeof(b));
  std::memcpy(stage + sizeof(BatchInstance) + 64, bytes.data(), bytes.size());
  SW_CHECK_HIP(hipMemcpyAsync(din, stage, in_bytes, hipMemcpyHostToDevice, s));
  SW_CHECK_HIP(launch_fill_u32(reinterpret_cast<uint32_t*>(dres), 0u, sizeof(BatchResult) / 4, s));
  SW_CHECK_HIP(launch_batch(1, dinst, dbytes, dorder, 1, dhead, dres, dwords, wcap, ctx->wg_scratch.as<uint8_t>(),
      stride, L, 1, s, 1024));
  uint8_t* back = stage + in_bytes;
  SW_CHECK_HIP(hipMemcpyAsync(back, dres, out_bytes, hipMemcpyDeviceToHost, s));
  SW_CHECK_HIP(hipStreamSynchronize(s));
  BatchResult r;
  std::memcpy(&r, back, sizeof(r));
  if (r.status != kBatchFound && r.status != kBatchBound) return false;
  std::memset(&out.r, 0, sizeof(out.r));
  out.word.clear();
  out.r.found = 1;
  out.r.length = r.length;
  out.r.steps_forward = r.steps;
  if (r.status == kBatchFound) {
    SW_REQUIRE(r.word_len <= wcap, "workgroup search: word past its buffer");
    const uint8_t* w = back + sizeof(BatchResult) + 64;
    out.word.assign(w, w + r.word_len);
    out.r.word_len = out.word.size();
  } else {
    out.r.bound_word = 1;
  }
  out.r.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return true;
}

}  // namespace sw

extern "C" int sw_batch_solve(sw_ctx* ctx, uint32_t count, const uint32_t* ns, const uint32_t* ks,
    const uint64_t* delta_offsets, const uint32_t* deltas, const sw_batch_config* cfg_in, int32_t* non_synchro,
    uint64_t* lengths, uint32_t* word_lens, uint32_t* words, uint32_t word_cap, sw_batch_stats* stats) {
  using namespace sw;
  try {
    SW_REQUIRE(ctx && ns && ks && delta_offsets && deltas && non_synchro && lengths, "NULL argument");
    sw_batch_config cfg{};
    if (cfg_in) cfg = *cfg_in;
    use(ctx);
    auto t0 = std::chrono::steady_clock::now();
    sw_batch_stats st{};

    // ---- preparation: validation, synchronizability and upper bounds, sink singletons
    struct Prep {
      bool ns = false, device = false;
      uint32_t wmax = 0;
      uint64_t b0[kBatchMaxW] = {0, 0, 0};
    };
    std::vector<Prep> prep(count);
    uint32_t wmax = 1;
    for (uint32_t i = 0; i < count; ++i) {
      SW_REQUIRE(ns[i] >= 1 && ks[i] >= 1 && ks[i] <= 255, "instance with n < 1 or k outside [1, 255]");
      if (ns[i] > 512) throw Error{SW_ERR_UNSUPPORTED, "n above 512 states is not compiled"};
      const uint32_t* d = deltas + delta_offsets[i];
      for (size_t j = 0; j < size_t(ns[i]) * ks[i]; ++j) SW_REQUIRE(d[j] < ns[i], "transition out of range");
    }
    // Non-synchronizing automata (some pair never merges) and Eppstein's
    // greedy bound with its word, for every instance in one launch
    // (k_eppstein.hip).  A bound whose word fits is the search's upper bound:
    // when nothing shorter exists, that word is the answer.
    constexpr uint32_t kEppWordCap = 1024;
    std::vector<int32_t> e_ns(count), e_imp(count);
    std::vector<uint64_t> e_up(count), e_bound(count), e_wl(count);
    std::vector<uint32_t> e_words(size_t(count) * kEppWordCap);
    for (uint32_t i = 0; i < count; ++i) e_up[i] = uint64_t(ns[i] - 1) * (ns[i] - 1) + 1;
    eppstein_batch(ctx, count, ns, ks, delta_offsets, deltas, e_up.data(), kEppWordCap, e_ns.data(), e_imp.data(),
        e_bound.data(), e_wl.data(), e_words.data());
    std::vector<uint64_t> upper(count);
    std::vector<char> upper_has_word(count, 0);
    for (uint32_t i = 0; i < count; ++i) {
      upper[i] = e_up[i];
      if (!e_ns[i] && e_imp[i] && e_wl[i] <= kEppWordCap && e_bound[i] < upper[i]) {
        upper[i] = e_bound[i];
        upper_has_word[i] = 1;
      }
    }
    auto upper_word = [&](uint32_t i) {
      return std::vector<uint32_t>(e_words.begin() + size_t(i) * kEppWordCap,
          e_words.begin() + size_t(i) * kEppWordCap + e_wl[i]);
    };
    {
      std::atomic<uint32_t> next{0};
      const unsigned nt = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
      std::vector<std::thread> pool;
      for (unsigned t = 0; t < nt; ++t)
        pool.emplace_back([&] {
          for (uint32_t i; (i = next.fetch_add(1)) < count;) {
            const uint32_t n = ns[i], k = ks[i];
            const uint32_t* d = deltas + delta_offsets[i];
            Prep& p = prep[i];
            p.ns = e_ns[i] != 0;
            if (p.ns || n == 1) continue;
            p.device = n <= kBatchMaxW * 64 && k <= kBatchMaxK;
            if (!p.device) continue;
            std::vector<char> sink = sink_states(n, k, d);
            std::vector<uint32_t> indeg(size_t(n) * k, 0);
            for (uint32_t q = 0; q < n; ++q)
              for (uint32_t a = 0; a < k; ++a) indeg[size_t(d[size_t(q) * k + a]) * k + a]++;
            for (uint32_t q = 0; q < n; ++q) {
              if (!sink[q]) continue;
              bool ok = false;
              for (uint32_t a = 0; a < k; ++a) ok |= indeg[size_t(q) * k + a] > 1;
              if (ok) p.b0[q / 64] |= 1ull << (q % 64);
            }
            p.wmax = (n + 63) / 64;
          }
        });
      for (auto& th : pool) th.join();
    }
    std::vector<BatchInstance> inst;
    std::vector<uint32_t> map;  // kernel slot -> instance
    std::vector<uint8_t> bytes;
    for (uint32_t i = 0; i < count; ++i) {
      non_synchro[i] = prep[i].ns;
      lengths[i] = 0;
      if (word_lens) word_lens[i] = 0;
      if (prep[i].ns) {
        st.non_synchro++;
        continue;
      }
      if (ns[i] == 1) {
        st.found++;
        continue;
      }
      if (!prep[i].device) continue;
      BatchInstance b{};
      b.n = ns[i];
      b.k = ks[i];
      b.delta_off = uint32_t(bytes.size());
      b.upper = uint32_t(std::min<uint64_t>(upper[i], 0xffffffffu));
      std::memcpy(b.b0, prep[i].b0, sizeof(b.b0));
      const uint32_t* d = deltas + delta_offsets[i];
      for (size_t j = 0; j < size_t(ns[i]) * ks[i]; ++j) bytes.push_back(uint8_t(d[j]));
      SW_REQUIRE(ns[i] <= 256, "byte transitions need n <= 256");
      wmax = std::max(wmax, prep[i].wmax);
      inst.push_back(b);
      map.push_back(i);
    }
    // Expensive first: more letters, then more states.
    std::vector<uint32_t> order(inst.size());
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
      if (inst[a].k != inst[b].k) return inst[a].k > inst[b].k;
      return inst[a].n > inst[b].n;
    });
    auto t1 = std::chrono::steady_clock::now();
    st.ms_host_prep = std::chrono::duration<double, std::milli>(t1 - t0).count();

    // ---- the persistent launch: 256-thread workgroups, several per CU.
    // Instances that outgrow a workgroup leave for the whole-device search
    // below.  A second launch of 1024-thread workgroups with 2^20 rows per
    // side for them was measured slower (5.2 s for the 1099 such instances
    // of the 4096-instance batch, against 1.5 s on the whole device).
    std::vector<BatchResult> res(inst.size());
    const uint32_t wc = std::max<uint32_t>(word_cap ? word_cap : 0, 2);
    const uint32_t kernel_wcap = std::min<uint32_t>(std::max<uint32_t>(wc, 64), 4096);
    std::vector<uint8_t> wbuf(size_t(inst.size()) * kernel_wcap);
    if (!inst.empty()) {
      BatchLimits L{};
      L.cap_rows = cfg.cap_rows ? cfg.cap_rows : (1u << 16);
      L.cap_cand = cfg.cap_cand ? cfg.cap_cand : 3u * (1u << 15);
      L.max_pairs = cfg.max_pairs ? cfg.max_pairs : (1ull << 32);
      uint32_t hs = 1024;
      while (hs < 2 * (L.cap_rows + L.cap_cand)) hs <<= 1;
      L.hash_slots = hs;
      const size_t stride = batch_scratch_bytes(int(wmax), L);
      // Workgroups resident at once: four per CU, three when the widest
      // instances' LDS (tables and goal tile, about 45 KB at n > 128) allows
      // no more.
      uint32_t grid = cfg.grid ? cfg.grid : uint32_t(ctx->cus) * (wmax >= 3 ? 3 : 4);
      grid = std::min<uint32_t>(grid, uint32_t(inst.size()));
      DevBuf d_inst, d_bytes, d_order, d_head, d_res, d_words, d_scratch;
      d_inst.ensure(inst.size() * sizeof(BatchInstance));
      d_bytes.ensure(bytes.size() + 8);
      d_order.ensure(order.size() * 4);
      d_head.ensure(8);
      d_res.ensure(res.size() * sizeof(BatchResult));
      d_words.ensure(wbuf.size() + 8);
      d_scratch.ensure(stride * grid);
      hipStream_t s = ctx->stream;
      SW_CHECK_HIP(hipMemcpyAsync(d_inst.p, inst.data(), inst.size() * sizeof(BatchInstance), hipMemcpyHostToDevice, s));
      SW_CHECK_HIP(hipMemcpyAsync(d_bytes.p, bytes.data(), bytes.size(), hipMemcpyHostToDevice, s));
      SW_CHECK_HIP(hipMemcpyAsync(d_order.p, order.data(), order.size() * 4, hipMemcpyHostToDevice, s));
      SW_CHECK_HIP(hipMemsetAsync(d_head.p, 0, 8, s));
      SW_CHECK_HIP(hipMemcpyAsync(d_res.p, res.data(), res.size() * sizeof(BatchResult), hipMemcpyHostToDevice, s));
      auto k0 = std::chrono::steady_clock::now();
      SW_CHECK_HIP(launch_batch(int(wmax), d_inst.as<BatchInstance>(), d_bytes.as<uint8_t>(), d_order.as<uint32_t>(),
          uint32_t(order.size()), d_head.as<uint32_t>(), d_res.as<BatchResult>(), d_words.as<uint8_t>(), kernel_wcap,
          d_scratch.as<uint8_t>(), stride, L, grid, s));
      SW_CHECK_HIP(hipStreamSynchronize(s));
      st.ms_kernel = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - k0).count();
      SW_CHECK_HIP(hipMemcpyAsync(res.data(), d_res.p, res.size() * sizeof(BatchResult), hipMemcpyDeviceToHost, s));
      SW_CHECK_HIP(hipMemcpyAsync(wbuf.data(), d_words.p, wbuf.size(), hipMemcpyDeviceToHost, s));
      SW_CHECK_HIP(hipStreamSynchronize(s));
      st.grid = grid;
      st.scratch_bytes = stride * grid;
    }

    // ---- results; overflowed (and out-of-range) instances on the whole device
    auto t2 = std::chrono::steady_clock::now();
    std::vector<char> done(count, 0);
    for (size_t j = 0; j < inst.size(); ++j) {
      const uint32_t i = map[j];
      const BatchResult& r = res[j];
      if (r.status == kBatchFound || r.status == kBatchBound) {
        lengths[i] = r.length;
        if (r.status == kBatchFound) {
          st.found++;
          if (words && word_lens && r.word_len <= word_cap) {
            word_lens[i] = r.word_len;
            for (uint32_t c = 0; c < r.word_len; ++c) words[size_t(i) * word_cap + c] = wbuf[j * kernel_wcap + c];
          }
        } else {
          st.bound++;
          // Nothing shorter than the upper bound: its word, when it has one.
          if (upper_has_word[i] && words && word_lens && e_wl[i] <= word_cap) {
            word_lens[i] = uint32_t(e_wl[i]);
            const std::vector<uint32_t> w = upper_word(i);
            std::copy(w.begin(), w.end(), words + size_t(i) * word_cap);
          }
        }
        done[i] = 1;
      } else if (r.status == kBatchOverflow) {
        switch (r.detail) {
          case kOverflowCand: st.over_cand++; break;
          case kOverflowRows: st.over_rows++; break;
          case kOverflowWord: st.over_word++; break;
          case kOverflowHash: st.over_hash++; break;
          case kOverflowPairs: st.over_pairs++; break;
          default: break;
        }
      }
    }
    // Overflowed instances: several host threads, each with its own context
    // (own stream, own reusable search buffers), so their many small
    // launches overlap on the device.
    std::vector<uint32_t> rest;
    for (uint32_t i = 0; i < count; ++i)
      if (!done[i] && !prep[i].ns && ns[i] != 1) rest.push_back(i);
    std::sort(rest.begin(), rest.end(), [&](uint32_t a, uint32_t b) {
      if (ks[a] != ks[b]) return ks[a] > ks[b];
      return ns[a] > ns[b];
    });
    st.overflowed = uint32_t(rest.size());
    if (!rest.empty()) {
      const uint32_t workers = std::min<uint32_t>(cfg.overflow_workers ? cfg.overflow_workers : 8,
          uint32_t(rest.size()));
      std::atomic<uint32_t> next{0};
      std::vector<std::string> errors(workers);
      std::vector<int> codes(workers, SW_OK);
      std::vector<double> busy_ms(workers, 0.0);
      std::vector<std::thread> pool;
      for (uint32_t t = 0; t < workers; ++t)
        pool.emplace_back([&, t] {
          sw_ctx* wctx = nullptr;
          try {
            if (t == 0) {
              wctx = ctx;
            } else {
              if (sw_ctx_create(ctx->device, &wctx) != SW_OK) throw Error{SW_ERR_HIP, sw_last_error()};
              wctx->tune = ctx->tune;
            }
            for (uint32_t j; (j = next.fetch_add(1)) < rest.size();) {
              const uint32_t i = rest[j];
              const uint32_t* d = deltas + delta_offsets[i];
              sw_exact_config ec{};
              ec.dfs = 1;
              const auto s0 = std::chrono::steady_clock::now();
              auto r = exact_search(wctx, ns[i], ks[i], d, 0, upper[i], ec, false);
              busy_ms[t] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - s0).count();
              if (!r.r.found) throw Error{SW_ERR_NOMEM, "exact search ended without a length (memory budget)"};
              lengths[i] = r.r.length;
              if (r.r.bound_word && upper_has_word[i]) r.word = upper_word(i);
              if (words && word_lens && r.word.size() <= word_cap) {
                word_lens[i] = uint32_t(r.word.size());
                std::copy(r.word.begin(), r.word.end(), words + size_t(i) * word_cap);
              }
            }
          } catch (const Error& e) {
            errors[t] = e.msg;
            codes[t] = e.code;
          } catch (const std::exception& e) {
            errors[t] = e.what();
            codes[t] = SW_ERR_INVALID;
          }
          if (wctx && wctx != ctx) sw_ctx_destroy(wctx);
        });
      for (auto& th : pool) th.join();
      for (double b : busy_ms) st.ms_overflow_search_sum += b;
      // The first worker's failure by severity: a memory budget end
      // (SW_ERR_NOMEM) is reported as such, not as a device error.
      int worst = -1;
      for (uint32_t t = 0; t < workers; ++t) {
        if (codes[t] == SW_OK) continue;
        if (worst < 0 || (codes[t] == SW_ERR_NOMEM && codes[worst] != SW_ERR_NOMEM)) worst = int(t);
      }
      if (worst >= 0) throw Error{codes[worst], "batch overflow path: " + errors[worst]};
      st.found += uint32_t(rest.size());
    }
    st.ms_overflow = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t2).count();
    st.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = st;
    return SW_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::exception& e) {
    set_error(e.what());
    return SW_ERR_INVALID;
  }
}
