// sw_frontier_step: one breadth-first step of the search on its own, the
// reference's bfs_step with a visited list
// (synchrolib/algorithm/exact/meet_in_the_middle.cpp:186-229; preimages as in
// invbfs_step, :245-263): every frontier set under every letter, duplicates
// dropped
// (sort_keep_unique there, an open-addressing hash set here) and sets of the
// visited list dropped (a merge against the sorted visited list there, a
// lookup in the same hash table here).
//
// The kernels are the device search's own (k_frontier.hip): the expansion,
// k_hash_insert over a table pre-filled with the visited rows
// (k_hash_rebuild), the keep scan and k_commit.  Each is timed with HIP
// events on the context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>

#include "sw_host.hpp"
#include "sw_kernels.hpp"

namespace sw {
namespace {

struct FrontierWs {
  DevBuf tables, masks, targets, in_visited, in_frontier, cand, slot, keep, pos, table, tmp, small, out, parent, letter;
  std::vector<uint64_t> host_masks;
  std::vector<uint16_t> host_targets;
  size_t tmp_bytes = 0;
  hipEvent_t ev[6] = {};
  uint64_t* host = nullptr;
  FrontierWs() {
    for (auto& e : ev) SW_CHECK_HIP(hipEventCreate(&e));
    SW_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&host), 64, hipHostMallocDefault));
  }
  ~FrontierWs() {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (host) (void)hipHostFree(host);
  }
};

// One workspace per context, freed with it.
FrontierWs& workspace(sw_ctx* ctx) {
  if (!ctx->frontier_ws) {
    ctx->frontier_ws = new FrontierWs();
    ctx->frontier_ws_free = [](void* p) { delete static_cast<FrontierWs*>(p); };
  }
  return *static_cast<FrontierWs*>(ctx->frontier_ws);
}

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  SW_CHECK_HIP(hipEventElapsedTime(&ms, a, b));
  return ms;
}

}  // namespace
}  // namespace sw

extern "C" int sw_frontier_step(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta, int dir,
    const uint64_t* visited, uint64_t visited_cnt, const uint64_t* frontier, uint64_t cnt, uint64_t* out,
    uint32_t* out_parent, uint8_t* out_letter, uint64_t* out_count, int where, sw_frontier_stats* stats) {
  using namespace sw;
  try {
    SW_REQUIRE(ctx && delta && out_count, "NULL argument");
    SW_REQUIRE((visited || !visited_cnt) && (frontier || !cnt) && (out || !cnt), "NULL argument");
    SW_REQUIRE(dir == 0 || dir == 1, "dir must be 0 or 1");
    SW_REQUIRE(k >= 1 && k <= 255, "k must be in [1, 255]");
    SW_REQUIRE(n >= 1, "n must be at least 1");
    if (n > 512) throw Error{SW_ERR_UNSUPPORTED, "n above 512 states is not compiled"};
    SW_REQUIRE(where == SW_HOST || where == SW_DEVICE, "where must be SW_HOST or SW_DEVICE");
    for (size_t i = 0; i < size_t(n) * k; ++i) SW_REQUIRE(delta[i] < n, "transition out of range");
    const uint64_t total = cnt * k;
    // Candidate and row indices are 31-bit (the high bit of a slot marks a
    // candidate of this step); parents are 32-bit.
    SW_REQUIRE(total < (1ull << 31) && visited_cnt < (1ull << 31), "more than 2^31 candidates or visited sets");
    use(ctx);
    FrontierWs& ws = workspace(ctx);
    hipStream_t st = ctx->stream;
    const uint32_t W = (n + 63) / 64, S = (n + 7) / 8;
    const uint64_t rb = uint64_t(W) * 8;

    build_tables_device(n, k, delta, dir, ws.host_masks, ws.masks, ws.tables, st);
    upload_targets(n, k, delta, ws.host_targets, ws.targets, st);
    const uint64_t* dv = visited;
    const uint64_t* df = frontier;
    if (where == SW_HOST) {
      ws.in_visited.ensure(visited_cnt * rb + 8);
      ws.in_frontier.ensure(cnt * rb + 8);
      if (visited_cnt)
        SW_CHECK_HIP(hipMemcpyAsync(ws.in_visited.p, visited, visited_cnt * rb, hipMemcpyHostToDevice, st));
      if (cnt) SW_CHECK_HIP(hipMemcpyAsync(ws.in_frontier.p, frontier, cnt * rb, hipMemcpyHostToDevice, st));
      dv = ws.in_visited.as<uint64_t>();
      df = ws.in_frontier.as<uint64_t>();
    }
    // Hash table at most half full: visited rows and candidates.
    uint64_t cap = 1024;
    while (cap < 2 * (visited_cnt + total)) cap <<= 1;
    if (cap > (1ull << 31)) throw Error{SW_ERR_NOMEM, "hash table above 2^31 slots"};
    ws.table.ensure(cap * 4);
    ws.cand.ensure(total * rb + 8);
    ws.slot.ensure(total * 4 + 4);
    ws.keep.ensure(total * 4 + 4);
    ws.pos.ensure(total * 4 + 4);
    ws.small.ensure(64);
    size_t tb = 0;
    SW_CHECK_HIP(keep_scan_temp_bytes(total, &tb));
    if (tb > ws.tmp_bytes) {
      ws.tmp.release();
      ws.tmp.ensure(tb);
      ws.tmp_bytes = tb;
    }
    uint64_t* dout = out;
    uint32_t* dpar = out_parent;
    uint8_t* dlet = out_letter;
    if (where == SW_HOST) {
      ws.out.ensure(total * rb + 8);
      dout = ws.out.as<uint64_t>();
      if (out_parent) {
        ws.parent.ensure(total * 4 + 4);
        dpar = ws.parent.as<uint32_t>();
      }
      if (out_letter) {
        ws.letter.ensure(total + 4);
        dlet = ws.letter.as<uint8_t>();
      }
    }
    unsigned long long* probes = reinterpret_cast<unsigned long long*>(ws.small.as<char>());
    uint32_t* newc = reinterpret_cast<uint32_t*>(ws.small.as<char>() + 8);
    SW_CHECK_HIP(launch_fill_u32(reinterpret_cast<uint32_t*>(probes), 0, 4, st));
    SW_CHECK_HIP(launch_fill_u32(ws.table.as<uint32_t>(), kEmpty, cap, st));
    SW_CHECK_HIP(launch_hash_rebuild(int(W), dv, visited_cnt, ws.table.as<uint32_t>(), cap - 1, st));
    SW_CHECK_HIP(hipEventRecord(ws.ev[0], st));
    SW_CHECK_HIP(launch_expand_sets(int(W), df, cnt, ws.tables.as<uint64_t>(), S, ws.targets.as<uint16_t>(), dir, k,
        nullptr, n, 0, ws.cand.as<uint64_t>(), nullptr, st));
    SW_CHECK_HIP(hipEventRecord(ws.ev[1], st));
    SW_CHECK_HIP(launch_hash_insert(int(W), ws.cand.as<uint64_t>(), nullptr, total, ws.table.as<uint32_t>(), cap - 1,
        dv, ws.slot.as<uint32_t>(), probes, st));
    SW_CHECK_HIP(hipEventRecord(ws.ev[2], st));
    SW_CHECK_HIP(launch_keep_scan(1, ws.slot.as<uint32_t>(), ws.table.as<uint32_t>(), nullptr, total,
        ws.keep.as<uint32_t>(), ws.pos.as<uint32_t>(), newc, ws.tmp.p, &ws.tmp_bytes, st));
    SW_CHECK_HIP(hipEventRecord(ws.ev[3], st));
    SW_CHECK_HIP(launch_commit(int(W), ws.cand.as<uint64_t>(), ws.keep.as<uint32_t>(), ws.pos.as<uint32_t>(),
        ws.slot.as<uint32_t>(), total, cnt, nullptr, 0, nullptr, 0, dout, dpar, dlet, nullptr, st));
    SW_CHECK_HIP(hipEventRecord(ws.ev[4], st));
    SW_CHECK_HIP(hipMemcpyAsync(ws.host, ws.small.p, 16, hipMemcpyDeviceToHost, st));
    SW_CHECK_HIP(hipStreamSynchronize(st));
    const uint64_t nnew = total ? uint32_t(ws.host[1]) : 0;
    if (where == SW_HOST && nnew) {
      SW_CHECK_HIP(hipMemcpyAsync(out, dout, nnew * rb, hipMemcpyDeviceToHost, st));
      if (out_parent) SW_CHECK_HIP(hipMemcpyAsync(out_parent, dpar, nnew * 4, hipMemcpyDeviceToHost, st));
      if (out_letter) SW_CHECK_HIP(hipMemcpyAsync(out_letter, dlet, nnew, hipMemcpyDeviceToHost, st));
      SW_CHECK_HIP(hipStreamSynchronize(st));
    }
    *out_count = nnew;
    if (stats) {
      stats->candidates = total;
      stats->new_sets = nnew;
      stats->probes = ws.host[0];
      stats->table_slots = cap;
      stats->ms_expand = elapsed(ws.ev[0], ws.ev[1]);
      stats->ms_insert = elapsed(ws.ev[1], ws.ev[2]);
      stats->ms_scan = elapsed(ws.ev[2], ws.ev[3]);
      stats->ms_commit = elapsed(ws.ev[3], ws.ev[4]);
    }
    return SW_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host out of memory");
    return SW_ERR_NOMEM;
  } catch (const std::exception& e) {
    set_error(e.what());
    return SW_ERR_INVALID;
  }
}
