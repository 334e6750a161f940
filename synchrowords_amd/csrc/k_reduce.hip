// Kernels of the Reduce stage (reduce.cpp): the union of a list's states,
// the first singleton of a list, and a list's sets renumbered onto the
// states of the reduced automaton.
//
// The lists of that stage are short (at most a few times the plan's
// list_size_threshold, N * K * 8 by default), so these are single-pass
// kernels with one thread per set.
#include "sw_device.hpp"
#include "sw_kernels.hpp"

namespace sw {

namespace {

// out[w] |= OR over the rows of word w: a wave reduces its 64 rows by a
// butterfly, lane 0 adds the result with one 64-bit atomic OR per word.
__global__ __launch_bounds__(256) void k_or_rows(const uint64_t* __restrict__ rows, uint64_t count, uint32_t W,
    unsigned long long* out) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  for (uint32_t w = 0; w < W; ++w) {
    uint64_t v = i < count ? rows[i * W + w] : 0ull;
    for (int off = 32; off > 0; off >>= 1) v |= __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicOr(out + w, (unsigned long long)v);
  }
}

// *first = min index i whose row holds exactly one state (unchanged if none).
__global__ __launch_bounds__(256) void k_first_singleton(const uint64_t* __restrict__ rows, uint64_t count,
    uint32_t W, unsigned long long* first) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  uint32_t c = 0;
  for (uint32_t w = 0; w < W; ++w) c += uint32_t(__popcll(rows[i * W + w]));
  if (c == 1) atomicMin(first, (unsigned long long)i);
}

// out row i (w_out words) = { map[q] : q in row i }, where map[q] is the
// new number of state q; every state of the rows has one (the kept states
// are closed under the transitions and contain the rows' union).
__global__ __launch_bounds__(256) void k_remap_sets(const uint64_t* __restrict__ in, uint64_t count, uint32_t w_in,
    const uint16_t* __restrict__ map, uint32_t n_in, uint64_t* __restrict__ out, uint32_t w_out) {
  __shared__ uint16_t smap[512];
  for (uint32_t q = threadIdx.x; q < n_in; q += blockDim.x) smap[q] = map[q];
  __syncthreads();
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  uint64_t r[kMaxW];
#pragma unroll
  for (int w = 0; w < kMaxW; ++w) r[w] = 0;
  for (uint32_t w = 0; w < w_in; ++w) {
    uint64_t bits = in[i * w_in + w];
    while (bits) {
      const uint32_t q = w * 64 + uint32_t(__builtin_ctzll(bits));
      bits &= bits - 1;
      const uint32_t t = smap[q];
#pragma unroll
      for (int v = 0; v < kMaxW; ++v) r[v] |= (int(t >> 6) == v) ? (1ull << (t & 63)) : 0ull;
    }
  }
#pragma unroll
  for (int w = 0; w < kMaxW; ++w)
    if (uint32_t(w) < w_out) out[i * w_out + w] = r[w];
}

inline unsigned blocks_of(uint64_t count) { return unsigned((count + 255) / 256); }

}  // namespace

hipError_t launch_or_rows(const uint64_t* rows, uint64_t count, uint32_t W, unsigned long long* out, hipStream_t st) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(k_or_rows, dim3(blocks_of(count)), dim3(256), 0, st, rows, count, W, out);
  return hipGetLastError();
}

hipError_t launch_first_singleton(const uint64_t* rows, uint64_t count, uint32_t W, unsigned long long* first,
    hipStream_t st) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(k_first_singleton, dim3(blocks_of(count)), dim3(256), 0, st, rows, count, W, first);
  return hipGetLastError();
}

hipError_t launch_remap_sets(const uint64_t* in, uint64_t count, uint32_t w_in, const uint16_t* map, uint32_t n_in,
    uint64_t* out, uint32_t w_out, hipStream_t st) {
  if (w_in > uint32_t(kMaxW) || w_out > uint32_t(kMaxW) || n_in > 512) return hipErrorInvalidValue;
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(k_remap_sets, dim3(blocks_of(count)), dim3(256), 0, st, in, count, w_in, map, n_in, out, w_out);
  return hipGetLastError();
}

}  // namespace sw
