// C ABI: housekeeping and the kernel-level entry points.
#include <hip/hip_runtime.h>

#include <atomic>
#include <climits>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "sw_gen.hpp"
#include "sw_host.hpp"
#include "sw_kernels.hpp"
#include "sw_reduce.hpp"

namespace sw {

static thread_local std::string g_error;

void set_error(const std::string& msg) { g_error = msg; }

void use(sw_ctx* ctx) { SW_CHECK_HIP(hipSetDevice(ctx->device)); }

}  // namespace sw

using sw::DevBuf;
using sw::Error;

#define SW_API_BEGIN try {
#define SW_API_END                      \
  }                                     \
  catch (const Error& e) {              \
    sw::set_error(e.msg);               \
    return e.code;                      \
  }                                     \
  catch (const std::bad_alloc&) {       \
    sw::set_error("host out of memory");\
    return SW_ERR_NOMEM;                \
  }                                     \
  catch (const std::exception& e) {     \
    sw::set_error(e.what());            \
    return SW_ERR_INVALID;              \
  }                                     \
  return SW_OK;

namespace {

uint32_t words_of(uint32_t n) { return (n + 63) / 64; }

void check_n(uint32_t n) {
  SW_REQUIRE(n >= 1, "n must be at least 1");
  if (n > 512) throw Error{SW_ERR_UNSUPPORTED, "n above 512 states is not compiled"};
}

// Device view of a caller list: the pointer itself (SW_DEVICE) or a copy in
// scratch buffer `buf` (SW_HOST).
const uint64_t* to_device(sw_ctx* ctx, DevBuf& buf, const uint64_t* p, uint64_t words, int where) {
  if (where == SW_DEVICE) return p;
  SW_REQUIRE(where == SW_HOST, "where must be SW_HOST or SW_DEVICE");
  buf.ensure(words * 8 + 8);
  if (words) SW_CHECK_HIP(hipMemcpyAsync(buf.p, p, words * 8, hipMemcpyHostToDevice, ctx->stream));
  return buf.as<uint64_t>();
}

}  // namespace

namespace {
std::atomic<int> g_user_contexts{0};
}  // namespace

namespace sw {
void* stage_bytes(sw_ctx* ctx, size_t bytes) {
  if (bytes > ctx->stage_cap) {
    size_t cap = ctx->stage_cap ? ctx->stage_cap : 16384;
    while (cap < bytes) cap *= 2;
    if (ctx->stage) SW_CHECK_HIP(hipHostFree(ctx->stage));
    ctx->stage = nullptr;
    ctx->stage_cap = 0;
    SW_CHECK_HIP(hipHostMalloc(&ctx->stage, cap, hipHostMallocDefault));
    ctx->stage_cap = cap;
  }
  return ctx->stage;
}
int user_contexts() { return g_user_contexts.load(); }
void mark_helper_context(sw_ctx* ctx) {
  if (!ctx->helper_ctx) {
    ctx->helper_ctx = true;
    g_user_contexts.fetch_sub(1);
  }
}
}  // namespace sw

extern "C" {

const char* sw_last_error(void) { return sw::g_error.c_str(); }

int sw_version(void) { return 10000; }

void sw_gen_random_dfa(uint32_t n, uint32_t k, uint64_t seed, uint32_t* out) {
  if (n && k && out) sw::random_dfa(n, k, seed, out);
}

int sw_ctx_create(int device, sw_ctx** out) {
  SW_API_BEGIN
  SW_REQUIRE(out, "out is NULL");
  int count = 0;
  SW_CHECK_HIP(hipGetDeviceCount(&count));
  SW_REQUIRE(device >= 0 && device < count, "no such HIP device");
  SW_CHECK_HIP(hipSetDevice(device));
  auto ctx = std::make_unique<sw_ctx>();
  ctx->device = device;
  SW_CHECK_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  hipDeviceProp_t prop;
  SW_CHECK_HIP(hipGetDeviceProperties(&prop, device));
  ctx->name = prop.gcnArchName;
  ctx->cus = prop.multiProcessorCount;
  ctx->mem = prop.totalGlobalMem;
  sw::noalloc_init(device);  // the no-allocation entry point's scratch address
  *out = ctx.release();
  g_user_contexts.fetch_add(1);
  SW_API_END
}

void sw_ctx_destroy(sw_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->helper && ctx->helper_free) ctx->helper_free(ctx->helper);  // joins its thread, ends its context
  ctx->helper = nullptr;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->a.release();
  ctx->b.release();
  ctx->c.release();
  ctx->d.release();
  ctx->e.release();
  if (ctx->exact_ws && ctx->exact_ws_free) ctx->exact_ws_free(ctx->exact_ws);
  ctx->exact_ws = nullptr;
  if (ctx->beam_ws && ctx->beam_ws_free) ctx->beam_ws_free(ctx->beam_ws);
  ctx->beam_ws = nullptr;
  if (ctx->frontier_ws && ctx->frontier_ws_free) ctx->frontier_ws_free(ctx->frontier_ws);
  ctx->frontier_ws = nullptr;
  if (ctx->ix_host) (void)hipHostFree(ctx->ix_host);
  ctx->ix_host = nullptr;
  if (ctx->stage) (void)hipHostFree(ctx->stage);
  ctx->stage = nullptr;
  if (ctx->stop_cell) (void)hipHostFree(ctx->stop_cell);
  ctx->stop_cell = nullptr;
  if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
  if (ctx->ev_stop) (void)hipEventDestroy(ctx->ev_stop);
  (void)hipStreamDestroy(ctx->stream);
  if (!ctx->helper_ctx) g_user_contexts.fetch_sub(1);
  delete ctx;
}

int sw_device_info(sw_ctx* ctx, char* name, size_t name_cap, int* cus, uint64_t* mem_bytes) {
  SW_API_BEGIN
  SW_REQUIRE(ctx, "ctx is NULL");
  if (name && name_cap) {
    std::snprintf(name, name_cap, "%s", ctx->name.c_str());
  }
  if (cus) *cus = ctx->cus;
  if (mem_bytes) *mem_bytes = ctx->mem;
  SW_API_END
}

int sw_dev_alloc(sw_ctx* ctx, size_t bytes, void** out) {
  SW_API_BEGIN
  SW_REQUIRE(ctx && out, "NULL argument");
  sw::use(ctx);
  SW_CHECK_HIP(hipMalloc(out, bytes ? bytes : 1));
  SW_API_END
}

int sw_dev_free(sw_ctx* ctx, void* p) {
  SW_API_BEGIN
  SW_REQUIRE(ctx, "ctx is NULL");
  sw::use(ctx);
  if (p) SW_CHECK_HIP(hipFree(p));
  SW_API_END
}

int sw_memcpy_h2d(sw_ctx* ctx, void* dst, const void* src, size_t bytes) {
  SW_API_BEGIN
  SW_REQUIRE(ctx, "ctx is NULL");
  sw::use(ctx);
  SW_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  SW_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  SW_API_END
}

int sw_memcpy_d2h(sw_ctx* ctx, void* dst, const void* src, size_t bytes) {
  SW_API_BEGIN
  SW_REQUIRE(ctx, "ctx is NULL");
  sw::use(ctx);
  SW_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  SW_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  SW_API_END
}

int sw_sync(sw_ctx* ctx) {
  SW_API_BEGIN
  SW_REQUIRE(ctx, "ctx is NULL");
  sw::use(ctx);
  SW_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  SW_API_END
}

int sw_dev_copy(sw_ctx* ctx, void* dst, const void* src, size_t bytes) {
  SW_API_BEGIN
  SW_REQUIRE(ctx && dst && src && bytes % 16 == 0, "NULL argument or size not a multiple of 16");
  sw::use(ctx);
  SW_CHECK_HIP(sw::launch_copy(src, dst, bytes, ctx->stream));
  SW_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  SW_API_END
}

int sw_calibrate_lds(sw_ctx* ctx, uint32_t bytes_per_lane, double* tb_per_s) {
  SW_API_BEGIN
  SW_REQUIRE(ctx && tb_per_s, "NULL argument");
  SW_REQUIRE(bytes_per_lane == 8 || bytes_per_lane == 16, "bytes_per_lane must be 8 or 16");
  sw::use(ctx);
  const uint32_t blocks = uint32_t(ctx->cus) * 2;  // 64 KB each: two workgroups per CU
  const uint64_t iters = 20000;
  ctx->d.ensure(size_t(blocks) * 4 + 8);
  hipEvent_t e0, e1;
  SW_CHECK_HIP(hipEventCreate(&e0));
  SW_CHECK_HIP(hipEventCreate(&e1));
  SW_CHECK_HIP(sw::launch_lds_read(int(bytes_per_lane), iters / 10, blocks, ctx->d.as<uint32_t>(), ctx->stream));
  SW_CHECK_HIP(hipEventRecord(e0, ctx->stream));
  SW_CHECK_HIP(sw::launch_lds_read(int(bytes_per_lane), iters, blocks, ctx->d.as<uint32_t>(), ctx->stream));
  SW_CHECK_HIP(hipEventRecord(e1, ctx->stream));
  SW_CHECK_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  SW_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  const double bytes = double(blocks) * 16.0 /* waves */ * double(iters) * 8.0 * 64.0 * double(bytes_per_lane);
  *tb_per_s = bytes / (double(ms) * 1e-3) / 1e12;
  SW_API_END
}

int sw_ctx_reset(sw_ctx* ctx) {
  SW_API_BEGIN
  SW_REQUIRE(ctx, "ctx is NULL");
  sw::use(ctx);
  SW_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  for (DevBuf* b : {&ctx->a, &ctx->b, &ctx->c, &ctx->d, &ctx->e, &ctx->ix_padded, &ctx->ix_size, &ctx->ix_start,
           &ctx->ix_list, &ctx->ix_tmp, &ctx->ix_order, &ctx->ix_sort, &ctx->ix_sperm, &ctx->ix_tperm,
           &ctx->ix_head, &ctx->ix_anchor, &ctx->ix_tlist, &ctx->ix_work, &ctx->wg_in, &ctx->wg_scratch,
           &ctx->wg_out, &ctx->brute_jobs, &ctx->brute_delta, &ctx->brute_frontier, &ctx->brute_out,
           &ctx->epp_jobs, &ctx->epp_delta, &ctx->epp_scratch, &ctx->epp_words, &ctx->epp_out, &ctx->epp_off,
           &ctx->epp_tmp, &ctx->epp_small})
    b->release();
  ctx->ix_tmp_bytes = 0;
  if (ctx->exact_ws && ctx->exact_ws_free) ctx->exact_ws_free(ctx->exact_ws);
  ctx->exact_ws = nullptr;
  if (ctx->beam_ws && ctx->beam_ws_free) ctx->beam_ws_free(ctx->beam_ws);
  ctx->beam_ws = nullptr;
  if (ctx->frontier_ws && ctx->frontier_ws_free) ctx->frontier_ws_free(ctx->frontier_ws);
  ctx->frontier_ws = nullptr;
  SW_API_END
}

int sw_stream_timer(sw_ctx* ctx, int op, float* ms) {
  SW_API_BEGIN
  SW_REQUIRE(ctx, "ctx is NULL");
  sw::use(ctx);
  if (!ctx->ev_start) {
    SW_CHECK_HIP(hipEventCreate(&ctx->ev_start));
    SW_CHECK_HIP(hipEventCreate(&ctx->ev_stop));
  }
  if (op == 0) {
    SW_CHECK_HIP(hipEventRecord(ctx->ev_start, ctx->stream));
  } else {
    SW_REQUIRE(ms, "ms is NULL");
    SW_CHECK_HIP(hipEventRecord(ctx->ev_stop, ctx->stream));
    SW_CHECK_HIP(hipEventSynchronize(ctx->ev_stop));
    SW_CHECK_HIP(hipEventElapsedTime(ms, ctx->ev_start, ctx->ev_stop));
  }
  SW_API_END
}

int sw_transition_apply(sw_ctx* ctx, uint32_t n, const uint64_t* table, const uint64_t* from, uint64_t* to,
    uint64_t cnt, int where) {
  SW_API_BEGIN
  SW_REQUIRE(ctx && table && (from || !cnt) && (to || !cnt), "NULL argument");
  check_n(n);
  sw::use(ctx);
  const uint32_t W = words_of(n), S = (n + 7) / 8;
  const size_t tw = size_t(S) * 256 * W;
  ctx->a.ensure(tw * 8);
  SW_CHECK_HIP(hipMemcpyAsync(ctx->a.p, table, tw * 8, hipMemcpyHostToDevice, ctx->stream));
  const uint64_t* din = to_device(ctx, ctx->b, from, cnt * W, where);
  uint64_t* dout = to;
  if (where == SW_HOST) {
    ctx->c.ensure(cnt * W * 8 + 8);
    dout = ctx->c.as<uint64_t>();
  }
  SW_CHECK_HIP(sw::launch_expand(int(W), din, cnt, ctx->a.as<uint64_t>(), S, 1, nullptr, n, 0, dout, nullptr,
      ctx->stream));
  if (where == SW_HOST && cnt)
    SW_CHECK_HIP(hipMemcpyAsync(to, dout, cnt * W * 8, hipMemcpyDeviceToHost, ctx->stream));
  SW_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  SW_API_END
}

int sw_expand(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta, int dir, const uint64_t* in, uint64_t cnt,
    uint64_t* out, int where) {
  SW_API_BEGIN
  SW_REQUIRE(ctx && delta && (in || !cnt) && (out || !cnt), "NULL argument");
  SW_REQUIRE(dir == 0 || dir == 1, "dir must be 0 or 1");
  SW_REQUIRE(k >= 1 && k <= 255, "k must be in [1, 255]");
  check_n(n);
  for (size_t i = 0; i < size_t(n) * k; ++i) SW_REQUIRE(delta[i] < n, "transition out of range");
  sw::use(ctx);
  const uint32_t W = words_of(n), S = (n + 7) / 8;
  auto t = sw::build_tables(n, k, delta, dir);
  ctx->a.ensure(t.size() * 8);
  SW_CHECK_HIP(hipMemcpyAsync(ctx->a.p, t.data(), t.size() * 8, hipMemcpyHostToDevice, ctx->stream));
  const uint64_t* din = to_device(ctx, ctx->b, in, cnt * W, where);
  uint64_t* dout = out;
  if (where == SW_HOST) {
    ctx->c.ensure(cnt * k * W * 8 + 8);
    dout = ctx->c.as<uint64_t>();
  }
  SW_CHECK_HIP(sw::launch_expand(int(W), din, cnt, ctx->a.as<uint64_t>(), S, k, nullptr, n, 0, dout, nullptr,
      ctx->stream));
  if (where == SW_HOST && cnt)
    SW_CHECK_HIP(hipMemcpyAsync(out, dout, cnt * k * W * 8, hipMemcpyDeviceToHost, ctx->stream));
  SW_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  SW_API_END
}

}  // extern "C"

struct sw_leaf {
  sw_ctx* ctx;
  uint32_t n;
  uint32_t small_cnt, large_cap;
  DevBuf small, large, flags;
};

extern "C" {

int sw_leaf_create(sw_ctx* ctx, uint32_t n, const uint64_t* small, uint32_t small_cnt, uint32_t large_cnt,
    sw_leaf** out) {
  SW_API_BEGIN
  SW_REQUIRE(ctx && out && (small || !small_cnt), "NULL argument");
  check_n(n);
  sw::use(ctx);
  auto leaf = std::make_unique<sw_leaf>();
  leaf->ctx = ctx;
  leaf->n = n;
  leaf->small_cnt = small_cnt;
  leaf->large_cap = large_cnt;
  const uint32_t W = words_of(n);
  leaf->small.ensure(size_t(small_cnt) * W * 8 + 8);
  if (small_cnt)
    SW_CHECK_HIP(hipMemcpyAsync(leaf->small.p, small, size_t(small_cnt) * W * 8, hipMemcpyHostToDevice, ctx->stream));
  leaf->large.ensure(size_t(large_cnt) * W * 8 + 8);
  leaf->flags.ensure(sw::vsub_flag_bytes(int(W), large_cnt) + 8);
  SW_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  *out = leaf.release();
  SW_API_END
}

int sw_leaf_run(sw_leaf* leaf, int proper, uint32_t small_pos, uint32_t small_cnt, const uint64_t* large,
    uint32_t large_cnt, uint8_t* ret) {
  SW_API_BEGIN
  SW_REQUIRE(leaf && (large || !large_cnt) && (ret || !large_cnt), "NULL argument");
  SW_REQUIRE(uint64_t(small_pos) + small_cnt <= leaf->small_cnt, "small range out of bounds");
  SW_REQUIRE(large_cnt <= leaf->large_cap, "large_cnt above the allocated size");
  sw_ctx* ctx = leaf->ctx;
  sw::use(ctx);
  const uint32_t W = words_of(leaf->n);
  if (!large_cnt) return SW_OK;
  SW_CHECK_HIP(hipMemcpyAsync(leaf->large.p, large, size_t(large_cnt) * W * 8, hipMemcpyHostToDevice, ctx->stream));
  const size_t fb = sw::vsub_flag_bytes(int(W), large_cnt);
  SW_CHECK_HIP(hipMemsetAsync(leaf->flags.p, 0, fb, ctx->stream));
  sw::VsubParams p;
  p.stream = leaf->small.as<uint64_t>() + size_t(small_pos) * W;
  p.s_count = small_cnt;
  p.tile = leaf->large.as<uint64_t>();
  p.t_count = large_cnt;
  p.n = leaf->n;
  p.mode = sw::VSUB_FLAGS;
  p.proper_by_size = proper ? 1 : 0;
  p.flags = leaf->flags.p;
  sw::apply_tuning(ctx, int(W), p);
  sw::run_subset(ctx, int(W), p);
  std::unique_ptr<uint8_t[]> bits(new uint8_t[fb]);
  SW_CHECK_HIP(hipMemcpyAsync(bits.get(), leaf->flags.p, fb, hipMemcpyDeviceToHost, ctx->stream));
  SW_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  for (uint32_t j = 0; j < large_cnt; ++j) ret[j] = (bits[j >> 3] >> (j & 7)) & 1;
  SW_API_END
}

void sw_leaf_destroy(sw_leaf* leaf) {
  if (!leaf) return;
  (void)hipSetDevice(leaf->ctx->device);
  (void)hipStreamSynchronize(leaf->ctx->stream);
  delete leaf;
}

int sw_subset_check(sw_ctx* ctx, uint32_t n, const uint64_t* small, uint64_t small_cnt, const uint64_t* large,
    uint64_t large_cnt, uint32_t thresh, int where, uint8_t* results, uint64_t* pair_count, int32_t* any) {
  SW_API_BEGIN
  SW_REQUIRE(ctx && (small || !small_cnt) && (large || !large_cnt), "NULL argument");
  check_n(n);
  SW_REQUIRE(small_cnt < (1ull << 32) && large_cnt < (1ull << 32), "lists above 2^32 sets");
  sw::use(ctx);
  const uint32_t W = words_of(n);
  const uint64_t* ds = to_device(ctx, ctx->a, small, small_cnt * W, where);
  const uint64_t* dl = to_device(ctx, ctx->b, large, large_cnt * W, where);
  // Scratch: [0] pair count, [8] found flag, [16] witness.
  ctx->c.ensure(64);
  unsigned long long* dcount = ctx->c.as<unsigned long long>();
  int* dfound = reinterpret_cast<int*>(ctx->c.as<char>() + 8);
  unsigned long long* dwit = reinterpret_cast<unsigned long long*>(ctx->c.as<char>() + 16);
  SW_CHECK_HIP(hipMemsetAsync(ctx->c.p, 0, 16, ctx->stream));
  SW_CHECK_HIP(hipMemsetAsync(dwit, 0xff, 8, ctx->stream));
  sw::VsubParams p;
  p.stream = ds;
  p.s_count = small_cnt;
  p.tile = dl;
  p.t_count = large_cnt;
  p.n = n;
  uint64_t total = 0;
  int32_t found = 0;
  if (results) {
    ctx->d.ensure(small_cnt * 4 + 8);
    SW_CHECK_HIP(hipMemsetAsync(ctx->d.p, 0, small_cnt * 4, ctx->stream));
    p.mode = sw::VSUB_SCOUNT;
    p.scount = ctx->d.as<uint32_t>();
    sw::apply_tuning(ctx, int(W), p);
    sw::run_subset(ctx, int(W), p);
    std::unique_ptr<uint32_t[]> c(new uint32_t[small_cnt + 1]);
    if (small_cnt)
      SW_CHECK_HIP(hipMemcpyAsync(c.get(), ctx->d.p, small_cnt * 4, hipMemcpyDeviceToHost, ctx->stream));
    SW_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    for (uint64_t i = 0; i < small_cnt; ++i) {
      results[i] = c[i] >= thresh;
      total += c[i];
    }
    found = total > 0;
  } else if (pair_count) {
    p.mode = sw::VSUB_COUNT;
    p.count_out = dcount;
    sw::apply_tuning(ctx, int(W), p);
    sw::run_subset(ctx, int(W), p);
    unsigned long long c = 0;
    SW_CHECK_HIP(hipMemcpyAsync(&c, dcount, 8, hipMemcpyDeviceToHost, ctx->stream));
    SW_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    total = c;
    found = c > 0;
  } else {
    p.mode = sw::VSUB_ANY;
    p.found = dfound;
    p.witness = dwit;
    sw::apply_tuning(ctx, int(W), p);
    sw::run_subset(ctx, int(W), p);
    SW_CHECK_HIP(hipMemcpyAsync(&found, dfound, 4, hipMemcpyDeviceToHost, ctx->stream));
    SW_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  }
  if (pair_count) *pair_count = total;
  if (any) *any = found ? 1 : 0;
  SW_API_END
}

int sw_count_bits(sw_ctx* ctx, uint32_t n, const uint64_t* sets, uint64_t cnt, uint64_t* acc, int where) {
  SW_API_BEGIN
  SW_REQUIRE(ctx && acc && (sets || !cnt), "NULL argument");
  check_n(n);
  sw::use(ctx);
  const uint32_t W = words_of(n);
  const uint64_t* d = to_device(ctx, ctx->a, sets, cnt * W, where);
  ctx->c.ensure(size_t(n) * 8);
  SW_CHECK_HIP(hipMemsetAsync(ctx->c.p, 0, size_t(n) * 8, ctx->stream));
  SW_CHECK_HIP(sw::launch_count_bits(int(W), d, cnt, n, ctx->c.as<unsigned long long>(), ctx->stream));
  SW_CHECK_HIP(hipMemcpyAsync(acc, ctx->c.p, size_t(n) * 8, hipMemcpyDeviceToHost, ctx->stream));
  SW_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  SW_API_END
}

int sw_exact_solve(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta, uint64_t lower, uint64_t upper,
    const uint32_t* upper_word, uint64_t upper_word_len, const sw_exact_config* cfg, sw_exact_result* out,
    uint32_t* word, uint64_t word_cap) {
  SW_API_BEGIN
  SW_REQUIRE(ctx && delta && out, "NULL argument");
  SW_REQUIRE(k >= 1 && k <= 255, "k must be in [1, 255]");
  check_n(n);
  for (size_t i = 0; i < size_t(n) * k; ++i) SW_REQUIRE(delta[i] < n, "transition out of range");
  SW_REQUIRE(lower <= upper, "lower above upper");
  sw_exact_config c{};
  if (cfg) c = *cfg;
  if (c.dfs == 0 && !cfg) c.dfs = 1;
  auto res = sw::exact_search(ctx, n, k, delta, lower, upper, c);
  *out = res.r;
  if (res.r.bound_word) {
    // The known word of the upper bound is optimal.
    res.word.assign(upper_word ? upper_word : nullptr, upper_word ? upper_word + upper_word_len : nullptr);
    if (!upper_word || upper_word_len != upper) res.word.clear();
  }
  out->word_len = res.word.size();
  if (word && res.word.size() <= word_cap) std::copy(res.word.begin(), res.word.end(), word);
  SW_API_END
}

int sw_exact_solve_from_list(sw_ctx* ctx, uint32_t n, uint32_t k, const uint32_t* delta, const uint64_t* list,
    uint64_t list_count, uint64_t lower, uint64_t upper, const sw_exact_config* cfg, sw_exact_result* out,
    uint32_t* word, uint64_t word_cap, int64_t* list_index) {
  SW_API_BEGIN
  SW_REQUIRE(ctx && delta && list && out, "NULL argument");
  SW_REQUIRE(k >= 1 && k <= 255, "k must be in [1, 255]");
  check_n(n);
  for (size_t i = 0; i < size_t(n) * k; ++i) SW_REQUIRE(delta[i] < n, "transition out of range");
  SW_REQUIRE(lower <= upper, "lower above upper");
  SW_REQUIRE(list_count >= 1, "empty list");
  const uint32_t W = (n + 63) / 64;
  for (uint64_t i = 0; i < list_count; ++i) {
    uint32_t c = 0;
    for (uint32_t w = 0; w < W; ++w) {
      const uint64_t x = list[i * W + w];
      const uint32_t lo = w * 64;
      const uint64_t valid = n >= lo + 64 ? ~0ull : ((1ull << (n - lo)) - 1);  // n > lo for w < W
      SW_REQUIRE(!(x & ~valid), "list set with a state >= n");
      c += uint32_t(__builtin_popcountll(x));
    }
    SW_REQUIRE(c >= 2, "list sets must hold at least two states");
  }
  sw_exact_config cc{};
  if (cfg) cc = *cfg;
  if (cc.dfs == 0 && !cfg) cc.dfs = 1;
  sw::DevBuf rows;
  uint64_t distinct = 0;
  std::vector<uint32_t> src;
  sw::distinct_rows(ctx, W, list, list_count, rows, distinct, src);
  sw::ExactSeed seed;
  seed.rows = rows.as<uint64_t>();
  seed.count = distinct;
  auto res = sw::exact_search(ctx, n, k, delta, lower, upper, cc, true, &seed);
  *out = res.r;
  if (res.r.bound_word) res.word.clear();  // the bound's word, if any, is the caller's
  out->word_len = res.word.size();
  if (list_index) *list_index = (!res.word.empty() && res.root >= 0) ? int64_t(src[size_t(res.root)]) : -1;
  if (word && res.word.size() <= word_cap) std::copy(res.word.begin(), res.word.end(), word);
  SW_API_END
}

}  // extern "C"
