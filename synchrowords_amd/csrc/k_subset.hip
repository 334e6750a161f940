// List-vs-list subset test on CDNA4: the "vertical" kernel.
//
// Question answered (the reference's leaf and checker kernels,
// cuda/subsets_implicit_trie_kernel.cu:12-91, cuda/subsets_checker_kernel.cu:8-33):
// for stream sets S_i and tile sets T_j, which pairs have S_i a subset of T_j?
//
// Layout.  A workgroup owns a tile of 64*MB tile sets (MB = 64 for N <= 256,
// 32 above).  The tile is transposed into LDS: row q, lane L holds an MB-bit
// mask whose bit b says "state q is in tile set L*MB + b".  A stream set S is
// wave-uniform (scalar loads); the wave walks only the states q of S and ANDs
// row q into a per-lane accumulator, so after |S| steps lane L knows, for all
// MB of its tile sets at once, which contain S.  One LDS row read per state of
// S tests 64*MB pairs; the stream set's bits are iterated on the scalar unit.
//
// Modes: COUNT (number of pairs), ANY (first pair found; every wave polls a
// flag and stops), FLAGS (per tile set: does some stream set lie inside it),
// SCOUNT (per stream set: how many tile sets contain it).
#include "sw_device.hpp"
#include "sw_kernels.hpp"

#include <atomic>
#include <climits>

namespace sw {

// Mask width of the tile: 64 tile sets per lane when the transposed tile
// ((n + 1) rows of 64 lanes x 8 bytes) fits the 160 KB of LDS (n <= 318; one
// workgroup per CU near the limit), else 32 (any n <= 512, half the LDS).
int vsub_mask_bits(uint32_t n, int W, int requested, bool proper_by_size) {
  const size_t lds64 = size_t(n + 1) * 64 * 8 + (proper_by_size ? size_t(64) * 64 * 2 : 0);
  const bool fits64 = W <= 5 && lds64 <= 160 * 1024;
  if (requested == 32) return 32;
  return fits64 ? 64 : 32;
}

size_t vsub_lds_bytes(uint32_t n, int W, int mask_bits, bool proper_by_size) {
  const int mb = vsub_mask_bits(n, W, mask_bits, proper_by_size);
  size_t b = size_t(n + 1) * 64 * (mb / 8);
  if (proper_by_size) b += size_t(64) * mb * sizeof(uint16_t);
  return b;
}

template <typename M>
__device__ inline int mpopc(M x) {
  return sizeof(M) == 8 ? __popcll(uint64_t(x)) : __popc(uint32_t(x));
}

// a & b & c in one gfx950 three-input bit operation per 32 bits
// (truth table 0x80: only the input 111 maps to 1).
__device__ inline uint32_t and3(uint32_t a, uint32_t b, uint32_t c) { return __builtin_amdgcn_bitop3_b32(a, b, c, 0x80); }
__device__ inline uint64_t and3(uint64_t a, uint64_t b, uint64_t c) {
  const uint32_t lo = and3(uint32_t(a), uint32_t(b), uint32_t(c));
  const uint32_t hi = and3(uint32_t(a >> 32), uint32_t(b >> 32), uint32_t(c >> 32));
  return (uint64_t(hi) << 32) | lo;
}

// Index of the lowest set bit of a wave-uniform word, which is then cleared:
// s_ff1 and s_bitset0 on the scalar unit (two instructions per bit).
__device__ inline uint32_t pop_low(uint64_t& bits) {
  const uint32_t q = __builtin_ctzll(bits);
  asm("s_bitset0_b64 %0, %1" : "+s"(bits) : "s"(q));
  return q;
}

// MODE is a template parameter: a runtime mode costs scalar tests and
// branches on every stream set.
template <int W, typename M, int MODE>
__global__ __launch_bounds__(1024) void k_vsub(VsubParams p) {
  constexpr int MB = sizeof(M) * 8;
  constexpr int TILE = 64 * MB;
  extern __shared__ __align__(16) unsigned char smem[];
  M* rows = reinterpret_cast<M*>(smem);  // [(n + 1) rows][64 lanes]
  const uint32_t n = p.n;
  uint16_t* tsize = reinterpret_cast<uint16_t*>(smem + size_t(n + 1) * 64 * sizeof(M));
  const int lane = threadIdx.x & 63;
  const int wave = rfl(threadIdx.x >> 6);
  const int nwaves = blockDim.x >> 6;
  const uint64_t tile_base = uint64_t(blockIdx.x) * TILE;
  const uint64_t tile_left = p.t_count - tile_base;  // > 0 by grid construction

  // ---- transpose the tile: group g = tile sets [g*64, g*64+64), one per lane.
  // Groups past the tile's last set are skipped: only lanes whose valid mask
  // is zero read them, and the accumulator only loses bits (a 136-set tile
  // took 75 us when all 64 groups were transposed).
  constexpr int NGROUPS = TILE / 64;
  const int ngroups = tile_left < uint64_t(TILE) ? int((tile_left + 63) / 64) : NGROUPS;
  for (int g = wave; g < ngroups; g += nwaves) {
    const uint64_t t = uint64_t(g) * 64 + lane;
    const bool valid = t < tile_left;
    int pc = 0;
    for (int w = 0; w < W; ++w) {
      uint64_t x = 0;
      if (valid) {
        x = p.tile[(tile_base + t) * W + w];
        if (p.complement_tile) x = ~x & word_mask(n, w);
      }
      pc += __popcll(x);
      const uint64_t kept = lane_transpose64(x);  // lane j: which of the 64 sets hold state w*64 + j
      const uint32_t q = uint32_t(w) * 64 + lane;
      if (q < n) {
        if (MB == 64) {
          rows[size_t(q) * 64 + g] = M(kept);
        } else {
          rows[size_t(q) * 64 + 2 * g] = M(kept);
          rows[size_t(q) * 64 + 2 * g + 1] = M(kept >> 32);
        }
      }
    }
    if (p.proper_by_size) tsize[t] = uint16_t(valid ? pc : 0xffff);
  }
  if (wave == 0) rows[size_t(n) * 64 + lane] = M(~M(0));  // padding row
  __syncthreads();

  // Lane L's tile sets that exist.
  M valid;
  {
    const int64_t c = int64_t(tile_left < uint64_t(TILE) ? tile_left : uint64_t(TILE)) - int64_t(lane) * MB;
    valid = c >= MB ? M(~M(0)) : (c <= 0 ? M(0) : M((M(1) << c) - 1));
  }

  const uint64_t s_begin = uint64_t(blockIdx.y) * p.s_chunk;
  uint64_t s_end = s_begin + p.s_chunk;
  if (s_end > p.s_count) s_end = p.s_count;

  uint32_t cnt = 0;
  M hits = 0;
  uint32_t iter = 0;
  // The next stream set is loaded while the current one is walked; the index
  // is clamped instead of tested, so the load has no branch.
  const uint64_t cmask = p.complement_stream ? ~0ull : 0ull;
  const uint64_t s_last = s_end > s_begin ? s_end - 1 : s_begin;
  uint64_t nx[W];
  {
    const uint64_t s0 = s_begin + wave < s_last ? s_begin + wave : s_last;
#pragma unroll
    for (int w = 0; w < W; ++w) nx[w] = p.stream[s0 * W + w];
  }
  for (uint64_t s = s_begin + wave; s < s_end; s += nwaves, ++iter) {
    uint64_t aw[W];
#pragma unroll
    for (int w = 0; w < W; ++w) aw[w] = (rfl64(nx[w]) ^ cmask) & word_mask(n, w);  // states >= n have no row
    {
      const uint64_t sn = s + nwaves < s_last ? s + nwaves : s_last;
#pragma unroll
      for (int w = 0; w < W; ++w) nx[w] = p.stream[sn * W + w];
    }
    if (MODE == VSUB_ANY && (iter & 15) == 0) {
      if (rfl(__hip_atomic_load(p.found, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) break;
    }
    M r = valid;
    bool go = true;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      if (!go) continue;
      uint64_t bits = aw[w];
      const M* rw = rows + size_t(w) * 64 * 64 + lane;  // row (w*64 + q) of this lane
      int c = __popcll(bits);
      for (; c >= 4; c -= 4) {
        const uint32_t q0 = pop_low(bits);
        const uint32_t q1 = pop_low(bits);
        const uint32_t q2 = pop_low(bits);
        const uint32_t q3 = pop_low(bits);
        const M m0 = rw[q0 * 64];
        const M m1 = rw[q1 * 64];
        const M m2 = rw[q2 * 64];
        const M m3 = rw[q3 * 64];
        r = and3(and3(r, m0, m1), m2, m3);
      }
      if (c >= 2) {
        const uint32_t q0 = pop_low(bits);
        const uint32_t q1 = pop_low(bits);
        r = and3(r, rw[q0 * 64], rw[q1 * 64]);
      }
      if (c & 1) r &= rw[__builtin_ctzll(bits) * 64];
      if (p.early_exit && w + 1 < W) go = __builtin_amdgcn_ballot_w64(r != 0) != 0;
    }
    if (p.diag != INT64_MIN) {
      // Skip the pair (s, t) with s - t == diag: the same element of one list.
      const int64_t t = int64_t(s) - p.diag - int64_t(tile_base);
      if (t >= 0 && t < TILE && lane == int(t / MB)) r &= ~(M(1) << (t % MB));
    }
    if (p.proper_by_size && r) {
      // A stream set inside a tile set of equal cardinality is that set.
      int c = 0;
#pragma unroll
      for (int w = 0; w < W; ++w) c += __popcll(aw[w]);
      M rr = r;
      while (rr) {
        const int b = sizeof(M) == 8 ? __builtin_ctzll(uint64_t(rr)) : __builtin_ctz(uint32_t(rr));
        rr &= rr - 1;
        if (tsize[lane * MB + b] == c) r &= ~(M(1) << b);
      }
    }
    if (MODE == VSUB_COUNT) {
      cnt += mpopc(r);
    } else if (MODE == VSUB_SCOUNT) {
      // Per stream set: pairs summed over the wave's lanes.
      if (__builtin_amdgcn_ballot_w64(r != 0)) {
        uint32_t c = mpopc(r);
        for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
        if (lane == 0) atomicAdd(&p.scount[s], c);
      }
    } else if (MODE == VSUB_FLAGS) {
      hits |= r;
    } else if (MODE == VSUB_SFLAGS) {
      if (__builtin_amdgcn_ballot_w64(r != 0) && lane == 0)
        atomicOr(reinterpret_cast<unsigned long long*>(p.flags) + (s >> 6), 1ull << (s & 63));
    } else if (r) {
      const int b = sizeof(M) == 8 ? __builtin_ctzll(uint64_t(r)) : __builtin_ctz(uint32_t(r));
      const unsigned long long wv = (s << 32) | (tile_base + uint64_t(lane) * MB + b);
      atomicMin(p.witness, wv);
      __hip_atomic_store(p.found, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }

  if (MODE == VSUB_COUNT) {
    uint64_t c = cnt;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if (lane == 0 && c) atomicAdd(p.count_out, (unsigned long long)c);
  } else if (MODE == VSUB_FLAGS) {
    if (hits) {
      M* f = reinterpret_cast<M*>(p.flags);
      atomicOr(&f[tile_base / MB + lane], hits);
    }
  }
}

template <int W, typename M, int MODE>
static hipError_t launch_vsub_wmm(const VsubParams& p, uint64_t gx, uint64_t gy, size_t lds, hipStream_t st) {
  // Idempotent, so concurrent first calls from several host threads are harmless.
  static std::atomic<bool> attr_set{false};
  if (!attr_set.load(std::memory_order_acquire)) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_vsub<W, M, MODE>),
        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e != hipSuccess) return e;
    attr_set.store(true, std::memory_order_release);
  }
  hipLaunchKernelGGL((k_vsub<W, M, MODE>), dim3(unsigned(gx), unsigned(gy)), dim3(p.threads), lds, st, p);
  return hipGetLastError();
}

template <int W, typename M>
static hipError_t launch_vsub_wm(const VsubParams& p0, hipStream_t st) {
  VsubParams p = p0;
  if (p.t_count == 0 || p.s_count == 0) return hipSuccess;
  constexpr int MB = sizeof(M) * 8;
  const size_t tile = size_t(64) * MB;
  const uint64_t gx = (p.t_count + tile - 1) / tile;
  if (p.threads != 256 && p.threads != 512 && p.threads != 1024) return hipErrorInvalidValue;
  if (p.s_chunk == 0) {
    // Aim for target_wgs workgroups, at least 64 stream sets per workgroup.
    const uint64_t target = p.target_wgs ? p.target_wgs : 1024;
    uint64_t gy = (target + gx - 1) / gx;
    uint64_t chunk = (p.s_count + gy - 1) / gy;
    if (chunk < 64) chunk = 64;
    p.s_chunk = chunk;
  }
  const uint64_t gy = (p.s_count + p.s_chunk - 1) / p.s_chunk;
  if (gx > 0x7fffffffull || gy > 65535ull) return hipErrorInvalidValue;
  const size_t lds = vsub_lds_bytes(p.n, W, MB, p.proper_by_size != 0);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  switch (p.mode) {
    case VSUB_COUNT: return launch_vsub_wmm<W, M, VSUB_COUNT>(p, gx, gy, lds, st);
    case VSUB_ANY: return launch_vsub_wmm<W, M, VSUB_ANY>(p, gx, gy, lds, st);
    case VSUB_FLAGS: return launch_vsub_wmm<W, M, VSUB_FLAGS>(p, gx, gy, lds, st);
    case VSUB_SCOUNT: return launch_vsub_wmm<W, M, VSUB_SCOUNT>(p, gx, gy, lds, st);
    case VSUB_SFLAGS: return launch_vsub_wmm<W, M, VSUB_SFLAGS>(p, gx, gy, lds, st);
  }
  return hipErrorInvalidValue;
}

template <int W>
static hipError_t launch_vsub_w(const VsubParams& p, hipStream_t st) {
  if (W <= 5 && vsub_mask_bits(p.n, W, p.mask_bits, p.proper_by_size != 0) == 64)
    return launch_vsub_wm<W, uint64_t>(p, st);
  return launch_vsub_wm<W, uint32_t>(p, st);
}

template <int W>
__global__ void k_random_sets(uint64_t* out, uint64_t count, uint32_t n, uint32_t thr, uint64_t seed) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  for (int w = 0; w < W; ++w) {
    uint64_t x = 0;
    for (int b = 0; b < 64; ++b) {
      const uint32_t q = w * 64 + b;
      uint64_t z = seed + (i * (W * 64) + q) * 0x9E3779B97F4A7C15ull;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      z ^= z >> 31;
      if (q < n && uint32_t(z >> 48) < thr) x |= 1ull << b;
    }
    out[i * W + w] = x;
  }
}

hipError_t launch_random_sets(int W, uint64_t* out, uint64_t count, uint32_t n, float p, uint64_t seed,
    hipStream_t st) {
  if (!count) return hipSuccess;
  const uint32_t thr = uint32_t(p * 65536.0f);
  const unsigned blocks = unsigned((count + 255) / 256);
  switch (W) {
#define C(w) case w: hipLaunchKernelGGL(k_random_sets<w>, dim3(blocks), dim3(256), 0, st, out, count, n, thr, seed); break;
    C(1) C(2) C(3) C(4) C(5) C(6) C(7) C(8)
#undef C
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_vsub(int W, const VsubParams& p, hipStream_t st) {
  switch (W) {
    case 1: return launch_vsub_w<1>(p, st);
    case 2: return launch_vsub_w<2>(p, st);
    case 3: return launch_vsub_w<3>(p, st);
    case 4: return launch_vsub_w<4>(p, st);
    case 5: return launch_vsub_w<5>(p, st);
    case 6: return launch_vsub_w<6>(p, st);
    case 7: return launch_vsub_w<7>(p, st);
    case 8: return launch_vsub_w<8>(p, st);
  }
  return hipErrorInvalidValue;
}

}  // namespace sw
