// sw_subset_check_noalloc: the list-vs-list checker for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).
//
// Same semantics as sw_subset_check (the reference's checker,
// subsets_checker_kernel.cu:8-33), on device pointers only, fully
// asynchronous on the context's stream: no hipMalloc, hipFree, hipMemset,
// host copy or synchronisation.  Its scratch — the per-row pair counts from
// which the per-row results and the total come, and the counters — lives in
// the file-scope array `scratch` below, one copy per GPU, held until the
// process ends.  The array's address on each device is looked up once, with
// that device current, when the first context on the device is created
// (noalloc_init, called by sw_ctx_create); host code uses only that address.
// Every call on one GPU shares the array, so calls that overlap on one GPU
// (on several contexts or streams) race; a call clears what must start at
// zero before its kernels run.
#include <hip/hip_runtime.h>

#include <mutex>

#include "sw_host.hpp"
#include "sw_kernels.hpp"

namespace sw {

// Per-row counts of up to kNoAllocRows small sets, after a 256-byte counter
// block: [0] pair count (u64), [8] found flag (i32), [16] witness (u64).
constexpr uint64_t kNoAllocRows = 1ull << 22;
constexpr uint64_t kNoAllocCounterBytes = 256;
constexpr uint64_t kNoAllocScratchBytes = kNoAllocCounterBytes + kNoAllocRows * 4;  // 16 MiB + 256 B

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_scratch[kMaxDevices] = {};

// Clear the counters (witness to all ones) and `rows` per-row counts.
__global__ void k_noalloc_clear(unsigned char* base, uint64_t rows) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i == 0) {
    *reinterpret_cast<unsigned long long*>(base) = 0ull;
    *reinterpret_cast<int*>(base + 8) = 0;
    *reinterpret_cast<unsigned long long*>(base + 16) = ~0ull;
  }
  uint32_t* cnt = reinterpret_cast<uint32_t*>(base + kNoAllocCounterBytes);
  for (uint64_t r = i; r < rows; r += uint64_t(gridDim.x) * blockDim.x) cnt[r] = 0;
}

// results[i] = cnt_i >= thresh; the counts are summed into the pair counter.
__global__ __launch_bounds__(256) void k_noalloc_rows(unsigned char* base, uint64_t rows, uint32_t thresh,
    uint8_t* results) {
  const uint32_t* cnt = reinterpret_cast<const uint32_t*>(base + kNoAllocCounterBytes);
  __shared__ unsigned long long part[4];
  unsigned long long sum = 0;
  for (uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; r < rows; r += uint64_t(gridDim.x) * blockDim.x) {
    const uint32_t c = cnt[r];
    if (results) results[r] = c >= thresh ? 1 : 0;
    sum += c;
  }
  for (int off = 32; off > 0; off >>= 1) sum += __shfl_down(sum, off, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long s = part[0] + part[1] + part[2] + part[3];
    if (s) atomicAdd(reinterpret_cast<unsigned long long*>(base), s);
  }
}

// The caller's outputs from the counters: *pair_count and *any.
__global__ void k_noalloc_out(const unsigned char* base, int any_from_flag, unsigned long long* pair_count,
    int32_t* any) {
  const unsigned long long c = *reinterpret_cast<const unsigned long long*>(base);
  const int f = *reinterpret_cast<const int*>(base + 8);
  if (pair_count) *pair_count = c;
  if (any) *any = any_from_flag ? (f != 0) : (c != 0);
}

}  // namespace

// Look up this device's copy of `scratch` (the device must be current).
// Once per device; later calls return at once.
void noalloc_init(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_scratch[device]) return;
  void* ptr = nullptr;
  SW_CHECK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  SW_REQUIRE(ptr != nullptr, "no device address for the no-allocation scratch");
  g_scratch[device] = ptr;
}

static unsigned char* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_mutex);
  return static_cast<unsigned char*>(g_scratch[device]);
}

}  // namespace sw

extern "C" int sw_subset_check_noalloc(sw_ctx* ctx, uint32_t n, const uint64_t* small, uint64_t small_cnt,
    const uint64_t* large, uint64_t large_cnt, uint32_t thresh, uint8_t* results, uint64_t* pair_count,
    int32_t* any) {
  try {
    SW_REQUIRE(ctx && (small || !small_cnt) && (large || !large_cnt), "NULL argument");
    SW_REQUIRE(n >= 1, "n must be >= 1");
    if (n > 512) throw sw::Error{SW_ERR_UNSUPPORTED, "n above 512 states is not compiled"};
    SW_REQUIRE(small_cnt < (1ull << 32) && large_cnt < (1ull << 32), "lists above 2^32 sets");
    if (results && small_cnt > sw::kNoAllocRows)
      throw sw::Error{SW_ERR_INVALID, "sw_subset_check_noalloc: per-row results are limited to 4194304 small sets "
                                      "(the static scratch holds 2^22 row counts)"};
    SW_CHECK_HIP(hipSetDevice(ctx->device));
    unsigned char* base = sw::noalloc_scratch(ctx->device);
    SW_REQUIRE(base != nullptr, "no-allocation scratch not set up on this device");
    const uint32_t W = (n + 63) / 64;
    hipStream_t st = ctx->stream;
    const uint64_t rows = results ? small_cnt : 0;
    const unsigned clear_blocks = unsigned(rows ? std::min<uint64_t>((rows + 255) / 256, 4096) : 1);
    hipLaunchKernelGGL(sw::k_noalloc_clear, dim3(clear_blocks), dim3(256), 0, st, base, rows);
    SW_CHECK_HIP(hipGetLastError());
    sw::VsubParams p;
    p.stream = small;
    p.s_count = small_cnt;
    p.tile = large;
    p.t_count = large_cnt;
    p.n = n;
    int any_from_flag = 0;
    if (results) {
      p.mode = sw::VSUB_SCOUNT;
      p.scount = reinterpret_cast<uint32_t*>(base + sw::kNoAllocCounterBytes);
    } else if (pair_count) {
      p.mode = sw::VSUB_COUNT;
      p.count_out = reinterpret_cast<unsigned long long*>(base);
    } else {
      p.mode = sw::VSUB_ANY;
      p.found = reinterpret_cast<int*>(base + 8);
      p.witness = reinterpret_cast<unsigned long long*>(base + 16);
      any_from_flag = 1;
    }
    sw::apply_tuning(ctx, int(W), p);
    // The bit-walking kernel: it needs no stream index, so nothing beyond the
    // counters and the row counts.
    SW_CHECK_HIP(sw::launch_vsub(int(W), p, st));
    if (results) {
      const unsigned blocks = unsigned(std::max<uint64_t>(1, std::min<uint64_t>((rows + 255) / 256, 4096)));
      hipLaunchKernelGGL(sw::k_noalloc_rows, dim3(blocks), dim3(256), 0, st, base, rows, thresh, results);
      SW_CHECK_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(sw::k_noalloc_out, dim3(1), dim3(1), 0, st, base, any_from_flag,
        reinterpret_cast<unsigned long long*>(pair_count), any);
    SW_CHECK_HIP(hipGetLastError());
    return SW_OK;
  } catch (const sw::Error& e) {
    sw::set_error(e.msg);
    return e.code;
  } catch (const std::exception& e) {
    sw::set_error(e.what());
    return SW_ERR_INVALID;
  }
}
