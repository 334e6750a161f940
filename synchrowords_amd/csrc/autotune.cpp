// Autotuner of the subset kernel's launch shape.
//
// For a (W, stream size, tile size) class it times every candidate shape on
// random lists of the given densities resident on the device, and records the
// fastest in the context's table; later calls of the same class (exact
// search, batch labelling, the seams) launch with it.  For the bit-walking
// kernel the shape is the mask width (64/32), the workgroup size
// (256/512/1024 threads), the workgroups aimed at (256..4096) and the
// per-word early exit; for a class that takes the indexed kernel, the
// workgroups aimed at and the early exit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>

#include "sw_host.hpp"
#include "sw_kernels.hpp"

namespace sw {

namespace {
uint32_t size_class(uint64_t c) {
  uint32_t k = 0;
  while (c > 1) {
    c >>= 1;
    ++k;
  }
  return k;
}
uint32_t key(int W, uint64_t s, uint64_t t) { return (uint32_t(W) << 16) | (size_class(s) << 8) | size_class(t); }
}  // namespace

void apply_tuning(sw_ctx* ctx, int W, VsubParams& p) {
  if (!ctx || ctx->tune.empty()) return;
  const uint32_t k = key(W, p.s_count, p.t_count);
  auto it = ctx->tune.find(k);
  if (it == ctx->tune.end()) {
    // Nearest class of the same W.
    int best = 1 << 30;
    for (auto jt = ctx->tune.begin(); jt != ctx->tune.end(); ++jt) {
      if ((jt->first >> 16) != uint32_t(W)) continue;
      const int ds = int((jt->first >> 8) & 255) - int((k >> 8) & 255);
      const int dt = int(jt->first & 255) - int(k & 255);
      const int d = ds * ds + dt * dt;
      if (d < best) {
        best = d;
        it = jt;
      }
    }
    if (it == ctx->tune.end()) return;
  }
  p.threads = it->second.threads;
  p.target_wgs = it->second.target_wgs;
  p.early_exit = int(it->second.early_exit);
  p.ix_early_exit = int(it->second.early_exit);  // the class's kernel is the one tuned
  p.mask_bits = int(it->second.mask_bits);
}

}  // namespace sw

extern "C" {

// Tune the subset kernel for stream/tile sizes and densities; returns the
// chosen shape.  Lists are generated on the device (seeded) and freed after.
int sw_autotune_subset(sw_ctx* ctx, uint32_t n, uint64_t s_count, uint64_t t_count, float p_stream, float p_tile,
    uint32_t reps, uint32_t* threads, uint32_t* target_wgs, uint32_t* early_exit, float* ms) {
  using namespace sw;
  try {
    SW_REQUIRE(ctx && s_count && t_count, "NULL argument or empty lists");
    SW_REQUIRE(n >= 1 && n <= 512, "n in [1, 512]");
    use(ctx);
    const int W = int((n + 63) / 64);
    hipStream_t st = ctx->stream;
    DevBuf S, T, out;
    S.ensure(s_count * W * 8);
    T.ensure(t_count * W * 8);
    out.ensure(64);
    SW_CHECK_HIP(launch_random_sets(W, S.as<uint64_t>(), s_count, n, p_stream, 0x5eed, st));
    SW_CHECK_HIP(launch_random_sets(W, T.as<uint64_t>(), t_count, n, p_tile, 0xfeed, st));
    hipEvent_t e0, e1;
    SW_CHECK_HIP(hipEventCreate(&e0));
    SW_CHECK_HIP(hipEventCreate(&e1));
    sw_ctx::Tune best{1024, 1024, 1, 64, 1e30f};
    const bool indexed = ctx->ix_enabled && s_count * t_count >= ctx->ix_min_pairs && vsub_ix_fits(n, W, false);
    if (indexed) {
      // Launches of this class take the indexed kernel (1024 threads, 64-bit
      // masks): the number of workgroups aimed at and the early exit are
      // free.  The index is built by each warm call and reused by the timed
      // ones.
      for (uint32_t ee = 0; ee < 2; ++ee)
      for (uint32_t wgs : {256u, 512u, 1024u, 2048u, 4096u}) {
        VsubParams p;
        p.stream = S.as<uint64_t>();
        p.s_count = s_count;
        p.tile = T.as<uint64_t>();
        p.t_count = t_count;
        p.n = n;
        p.mode = VSUB_COUNT;
        p.count_out = out.as<unsigned long long>();
        p.target_wgs = wgs;
        p.ix_early_exit = int(ee);
        run_subset(ctx, W, p);  // warm; builds the index
        SW_CHECK_HIP(hipEventRecord(e0, st));
        for (uint32_t r = 0; r < (reps ? reps : 3); ++r) SW_CHECK_HIP(launch_vsub_ix(W, p, st));
        SW_CHECK_HIP(hipEventRecord(e1, st));
        SW_CHECK_HIP(hipEventSynchronize(e1));
        float t = 0;
        SW_CHECK_HIP(hipEventElapsedTime(&t, e0, e1));
        t /= float(reps ? reps : 3);
        if (t < best.ms) best = {1024, wgs, ee, 64, t};
      }
    }
    const uint32_t th[] = {256, 512, 1024};
    const uint32_t wg[] = {256, 512, 1024, 2048, 4096};
    const uint32_t mbs[] = {64, 32};
    for (uint32_t mb : mbs)
    for (uint32_t threads_i : th)
      if (!indexed)
      for (uint32_t wgs : wg)
        for (uint32_t ee = 0; ee < 2; ++ee) {
          if (mb == 64 && vsub_mask_bits(n, W, 64, false) != 64) continue;
          VsubParams p;
          p.stream = S.as<uint64_t>();
          p.s_count = s_count;
          p.tile = T.as<uint64_t>();
          p.t_count = t_count;
          p.n = n;
          p.mode = VSUB_COUNT;
          p.count_out = out.as<unsigned long long>();
          p.threads = threads_i;
          p.target_wgs = wgs;
          p.early_exit = int(ee);
          p.mask_bits = int(mb);
          SW_CHECK_HIP(launch_vsub(W, p, st));  // warm
          SW_CHECK_HIP(hipEventRecord(e0, st));
          for (uint32_t r = 0; r < (reps ? reps : 3); ++r) SW_CHECK_HIP(launch_vsub(W, p, st));
          SW_CHECK_HIP(hipEventRecord(e1, st));
          SW_CHECK_HIP(hipEventSynchronize(e1));
          float t = 0;
          SW_CHECK_HIP(hipEventElapsedTime(&t, e0, e1));
          t /= float(reps ? reps : 3);
          if (t < best.ms) best = {threads_i, wgs, ee, mb, t};
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->tune[key(W, s_count, t_count)] = best;
    if (threads) *threads = best.threads;
    if (target_wgs) *target_wgs = best.target_wgs;
    if (early_exit) *early_exit = best.early_exit | (best.mask_bits << 8);
    if (ms) *ms = best.ms;
    return SW_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  }
}

// Record a launch shape for a size class by hand (one an earlier autotune
// reported), so a later process can reuse it without timing again.
int sw_set_subset_shape(sw_ctx* ctx, uint32_t n, uint64_t s_count, uint64_t t_count, uint32_t threads,
    uint32_t target_wgs, uint32_t early_exit) {
  using namespace sw;
  try {
    SW_REQUIRE(ctx && s_count && t_count, "NULL argument or empty lists");
    SW_REQUIRE(n >= 1 && n <= 512, "n in [1, 512]");
    SW_REQUIRE(threads == 256 || threads == 512 || threads == 1024, "threads must be 256, 512 or 1024");
    SW_REQUIRE(target_wgs >= 1, "target_wgs must be positive");
    const int W = int((n + 63) / 64);
    const uint32_t mb = early_exit >> 8;
    SW_REQUIRE(mb == 32 || (mb == 64 && vsub_mask_bits(n, W, 64, false) == 64),
        "mask bits must be 32, or 64 when the tile fits LDS (n <= 318)");
    ctx->tune[key(W, s_count, t_count)] = sw_ctx::Tune{threads, target_wgs, early_exit & 1, mb, 0.f};
    return SW_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  }
}

// Which subset kernel later launches use: the indexed one (k_vsub_ix) for
// launches of at least min_pairs pairs when `indexed`, else always the
// bit-walking one (k_vsub).
int sw_set_subset_kernel(sw_ctx* ctx, int indexed, uint64_t min_pairs) {
  using namespace sw;
  try {
    SW_REQUIRE(ctx, "NULL argument");
    ctx->ix_enabled = indexed ? 1 : 0;
    ctx->ix_min_pairs = min_pairs ? min_pairs : (1ull << 30);
    return SW_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  }
}

// How the indexed kernel walks the states of a stream set: as each launch
// asks (0: the exact search asks for 2, counts for sorted lists, other
// launches take the tuned or default shape), in set order without early exit
// (1), rarest-in-tile first with early exit over size-sorted lists (2),
// size-sorted lists in set order without early exit (3), or rarest first with
// early exit over anchor buckets where the mode allows, else as 2 (4).
int sw_set_subset_walk(sw_ctx* ctx, int walk) {
  using namespace sw;
  try {
    SW_REQUIRE(ctx, "NULL argument");
    SW_REQUIRE(walk >= 0 && walk <= 4, "walk must be in [0, 4]");
    ctx->ix_walk = walk;
    return SW_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  }
}

}  // extern "C"
