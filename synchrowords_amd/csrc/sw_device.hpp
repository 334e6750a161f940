// Device-side helpers shared by the HIP kernels.
//
// State sets are bitsets of N states in W = ceil(N/64) 64-bit words, state q
// at bit q%64 of word q/64, bits >= N zero — the reference's Subset<N>
// layout (synchrolib/data_structures/subset.hpp:21-28, 190-196), so host
// arrays of the reference's Subset<N> can be passed through the C ABI as
// they are.  Lists are row-major: set i occupies words [i*W, i*W + W).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sw {

constexpr int kMaxW = 8;  // N <= 512 states

// Mask of the valid bits of word w for n states.
__host__ __device__ inline uint64_t word_mask(uint32_t n, int w) {
  const uint32_t lo = uint32_t(w) * 64u;
  if (n >= lo + 64u) return ~0ull;
  if (n <= lo) return 0ull;
  return (1ull << (n - lo)) - 1ull;
}

__device__ inline uint32_t rfl(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

// 64 x 64 bit transpose across the lanes of a wave: lane i holds row i (bit
// b = entry (i, b)); lane j gets column j (bit i = bit j of lane i's row).
// Six butterfly steps swap the off-diagonal blocks of size s = 32 ... 1
// between lanes i and i ^ s, about 50 instructions where 64 ballots take
// about 500.
__device__ inline uint64_t lane_transpose64(uint64_t x) {
  const uint32_t lane = __lane_id();
  constexpr uint64_t kLow[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
      0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int s = 32 >> k;
    const uint64_t lo = kLow[k];
    const uint64_t y = __shfl_xor(x, s, 64);
    x = (lane & uint32_t(s)) ? ((x & ~lo) | ((y & ~lo) >> s)) : ((x & lo) | ((y & lo) << s));
  }
  return x;
}
__device__ inline uint64_t rfl64(uint64_t x) {
  return (uint64_t(rfl(uint32_t(x >> 32))) << 32) | rfl(uint32_t(x));
}

// 64-bit mix of a set (splitmix64 finaliser over the words).
template <int W>
__device__ inline uint64_t set_hash(const uint64_t* s) {
  uint64_t h = 0x9E3779B97F4A7C15ull * (W + 1);
#pragma unroll
  for (int w = 0; w < W; ++w) {
    uint64_t z = h ^ s[w];
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    h = z ^ (z >> 31);
  }
  return h;
}

template <int W>
__device__ inline int set_popc(const uint64_t* s) {
  int c = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) c += __popcll(s[w]);
  return c;
}

template <int W>
__device__ inline bool set_eq(const uint64_t* a, const uint64_t* b) {
  bool e = true;
#pragma unroll
  for (int w = 0; w < W; ++w) e &= (a[w] == b[w]);
  return e;
}

}  // namespace sw
