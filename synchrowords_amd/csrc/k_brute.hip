// Power-set BFS for small automata (the reference's Brute,
// synchrolib/algorithm/brute/brute.hpp:28-122): breadth-first search over the
// subsets of Q (n <= 32 states, one bit each in a 32-bit mask) from Q itself;
// the first level that reaches a singleton is the shortest reset length.
//
// The image of a subset under a letter is looked up per byte of the mask:
// table[a][b][v] = image of the states 8 b + j, j in v, so an image is four
// loads and three ORs.  Visited subsets are one bit each.
//
// * k_brute_wg: one workgroup per automaton, every level inside one launch,
//   bitmap in LDS (up to 2^20 subsets for two letters), for batches of small
//   automata and for one small automaton alike.
// * k_brute_level: one launch per level over the whole device, bitmap in
//   device memory, for one automaton of up to 32 states.
#include <hip/hip_runtime.h>

#include <atomic>

#include "sw_kernels.hpp"

namespace sw {

namespace {

__device__ inline uint32_t image_of(const uint32_t* __restrict__ tbl, uint32_t k, uint32_t a, uint32_t m) {
  const uint32_t* t = tbl + size_t(a) * 1024;
  return t[m & 255u] | t[256 + ((m >> 8) & 255u)] | t[512 + ((m >> 16) & 255u)] | t[768 + (m >> 24)];
}

// table[a][b][v] for the automaton's k letters; threads of one block.
__device__ inline void build_tables(uint32_t* tbl, uint32_t n, uint32_t k, const uint32_t* __restrict__ delta) {
  for (uint32_t t = threadIdx.x; t < k * 1024; t += blockDim.x) {
    const uint32_t a = t >> 10, b = (t >> 8) & 3u, v = t & 255u;
    uint32_t img = 0;
    for (uint32_t j = 0; j < 8; ++j) {
      const uint32_t q = 8 * b + j;
      if (((v >> j) & 1u) && q < n) img |= 1u << delta[size_t(q) * k + a];
    }
    tbl[t] = img;
  }
}

// Appends the lanes' new subsets to `out` with one LDS atomic per wave.
__device__ inline void append(bool add, uint32_t v, uint32_t* counter, uint32_t* out) {
  const uint64_t bal = __builtin_amdgcn_ballot_w64(add);
  if (!bal) return;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t below = __builtin_amdgcn_mbcnt_hi(uint32_t(bal >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(bal), 0u));
  const uint32_t leader = __builtin_ctzll(bal);
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, uint32_t(__popcll(bal)));
  base = __shfl(base, int(leader), 64);
  if (add) out[base + below] = v;
}

// One workgroup per job; every level in this launch.  LDS: k x 4 KB of
// tables, then the visited bitmap (2^n bits).  out[job] = length or -1.
__global__ __launch_bounds__(1024) void k_brute_wg(const BruteJob* __restrict__ jobs, const uint32_t* __restrict__ deltas,
    uint32_t* __restrict__ frontier, int64_t* __restrict__ out) {
  extern __shared__ uint32_t sm[];
  __shared__ uint32_t next[3];  // per level L, counter L % 3 (reset two levels ahead)
  __shared__ uint32_t found;
  const BruteJob job = jobs[blockIdx.x];
  const uint32_t n = job.n, k = job.k;
  if (n <= 1) {
    if (threadIdx.x == 0) out[blockIdx.x] = 0;
    return;
  }
  uint32_t* tbl = sm;
  uint32_t* map = sm + k * 1024;
  const uint32_t words = n >= 5 ? (1u << (n - 5)) : 1u;
  build_tables(tbl, n, k, deltas + job.delta_off);
  for (uint32_t t = threadIdx.x; t < words; t += blockDim.x) map[t] = 0;
  if (threadIdx.x < 3) next[threadIdx.x] = 0;
  if (threadIdx.x == 0) found = 0;
  __syncthreads();
  const uint32_t full = n == 32 ? 0xffffffffu : ((1u << n) - 1u);
  uint32_t* cur = frontier + job.frontier_off;
  uint32_t* nxt = cur + (size_t(1) << n);
  if (threadIdx.x == 0) {
    cur[0] = full;
    map[full >> 5] |= 1u << (full & 31u);
  }
  __syncthreads();
  uint32_t fc = 1;
  for (uint32_t level = 0;; ++level) {
    uint32_t* counter = &next[level % 3];
    if (threadIdx.x == 0) next[(level + 1) % 3] = 0;  // last read two levels ago
    for (uint32_t base = 0; base < fc; base += blockDim.x) {
      const uint32_t i = base + threadIdx.x;
      const uint32_t m = i < fc ? cur[i] : 0u;
      for (uint32_t a = 0; a < k; ++a) {
        bool add = false;
        uint32_t v = 0;
        if (i < fc) {
          v = image_of(tbl, k, a, m);
          const uint32_t bit = 1u << (v & 31u);
          add = !(atomicOr(&map[v >> 5], bit) & bit);
          if (add && __popc(v) == 1) found = 1;
        }
        append(add, v, counter, nxt);
      }
    }
    __syncthreads();
    const uint32_t nf = *counter;
    if (found) {
      if (threadIdx.x == 0) out[blockIdx.x] = int64_t(level) + 1;
      return;
    }
    if (nf == 0) {
      if (threadIdx.x == 0) out[blockIdx.x] = -1;
      return;
    }
    fc = nf;
    uint32_t* t = cur;
    cur = nxt;
    nxt = t;
  }
}

// One level of the BFS for one automaton over the whole device: images of
// cur[0 .. *fc) marked in the bitmap, new subsets appended to nxt.
// state[0] = new count, state[1] = a singleton was reached.
__global__ __launch_bounds__(256) void k_brute_level(uint32_t n, uint32_t k, const uint32_t* __restrict__ delta,
    const uint32_t* __restrict__ cur, const uint32_t* __restrict__ fc_p, uint32_t* __restrict__ nxt,
    uint32_t* __restrict__ map, uint32_t* __restrict__ state) {
  extern __shared__ uint32_t tbl[];
  __shared__ uint32_t wg_count, wg_base;
  build_tables(tbl, n, k, delta);
  if (threadIdx.x == 0) wg_count = 0;
  __syncthreads();
  const uint32_t fc = *fc_p;
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (uint64_t base = uint64_t(blockIdx.x) * blockDim.x; base < fc; base += stride) {
    const uint64_t i = base + threadIdx.x;
    const uint32_t m = i < fc ? cur[i] : 0u;
    // Count this pass's new subsets in the workgroup, reserve one range.
    uint32_t v[kBruteLevelMaxK];
    uint32_t newmask = 0;
#pragma unroll
    for (uint32_t a = 0; a < kBruteLevelMaxK; ++a) {  // constant indices: v stays in registers
      v[a] = 0;
      if (a < k && i < fc) {
        v[a] = image_of(tbl, k, a, m);
        const uint32_t bit = 1u << (v[a] & 31u);
        if (!(atomicOr(&map[v[a] >> 5], bit) & bit)) {
          newmask |= 1u << a;
          if (__popc(v[a]) == 1) state[1] = 1u;
        }
      }
    }
    const uint32_t mine = __popc(newmask);
    const uint32_t off = atomicAdd(&wg_count, mine);
    __syncthreads();
    if (threadIdx.x == 0) {
      wg_base = wg_count ? atomicAdd(&state[0], wg_count) : 0u;
      wg_count = 0;
    }
    __syncthreads();
    uint32_t o = wg_base + off;
#pragma unroll
    for (uint32_t a = 0; a < kBruteLevelMaxK; ++a)
      if ((newmask >> a) & 1u) nxt[o++] = v[a];
    __syncthreads();
  }
}

}  // namespace

int brute_lds_bytes(uint32_t n, uint32_t k) {
  const uint32_t words = n >= 5 ? (1u << (n - 5)) : 1u;
  return int(size_t(k) * 4096 + size_t(words) * 4);
}

hipError_t launch_brute_wg(const BruteJob* jobs, uint32_t count, const uint32_t* deltas, uint32_t* frontier,
    int64_t* out, size_t lds, hipStream_t st) {
  if (!count) return hipSuccess;
  if (lds > kBruteMaxLds) return hipErrorInvalidValue;
  static std::atomic<bool> attr_set{false};
  if (!attr_set.load(std::memory_order_acquire)) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_brute_wg),
        hipFuncAttributeMaxDynamicSharedMemorySize, kBruteMaxLds);
    if (e != hipSuccess) return e;
    attr_set.store(true, std::memory_order_release);
  }
  hipLaunchKernelGGL(k_brute_wg, dim3(count), dim3(1024), lds, st, jobs, deltas, frontier, out);
  return hipGetLastError();
}

hipError_t launch_brute_level(uint32_t n, uint32_t k, const uint32_t* delta, const uint32_t* cur, const uint32_t* fc,
    uint32_t fc_host, uint32_t* nxt, uint32_t* map, uint32_t* state, hipStream_t st) {
  if (k > uint32_t(kBruteLevelMaxK) || n > 32 || n < 2) return hipErrorInvalidValue;
  uint64_t blocks = (uint64_t(fc_host) + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(k_brute_level, dim3(unsigned(blocks)), dim3(256), size_t(k) * 4096, st, n, k, delta, cur, fc,
      nxt, map, state);
  return hipGetLastError();
}

}  // namespace sw
