// Choice between the two subset kernels and the stream-index pre-pass of the
// indexed one (k_subset_ix.hip).
#include <hip/hip_runtime.h>

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
  const uint64_t s = p.s_count, pairs = (s + 1) / 2;
  ctx->ix_padded.ensure(pairs * 4 + 16);  // entries per pair
  ctx->ix_size.ensure(s * 4 + 16);
  ctx->ix_start.ensure((pairs + 1) * 4 + 16);
  if (!ctx->ix_host) SW_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&ctx->ix_host), 16, hipHostMallocDefault));
  SW_CHECK_HIP(launch_ix_pair_count(W, p.stream, s, p.n, p.complement_stream, ctx->ix_size.as<uint32_t>(),
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
  SW_CHECK_HIP(hipMemcpyAsync(ctx->ix_host, start + pairs, 4, hipMemcpyDeviceToHost, st));
  SW_CHECK_HIP(hipStreamSynchronize(st));
  const uint64_t total = ctx->ix_host[0];
  ctx->ix_list.ensure(total * 4 + 64);
  SW_CHECK_HIP(launch_ix_pair_fill(W, p.stream, s, p.n, p.complement_stream, start, ctx->ix_list.as<uint32_t>(), st));
  p.ix_start = start;
  p.ix_list = ctx->ix_list.as<uint32_t>();
  p.ix_size = ctx->ix_size.as<uint32_t>();
  SW_CHECK_HIP(launch_vsub_ix(W, p, st));
}

}  // namespace sw
