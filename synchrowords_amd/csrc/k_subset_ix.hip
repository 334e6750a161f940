// List-vs-list subset test, indexed variant: k_vsub_ix.
//
// The same question and the same bit-transposed tile as k_vsub (k_subset.hip):
// 4096 tile sets per workgroup, row q of LDS = 64 x 64-bit masks, bit b of
// mask g says "state q is in tile set 64 g + b".  Two things differ.
//
// * The states of every stream set are listed once, by a pre-pass
//   (k_ix_pair_count, an exclusive scan, k_ix_pair_fill), as the LDS byte
//   offsets of their rows.  k_vsub finds the states of each stream set on the
//   scalar unit (s_ff1 and s_bitset0 per state, plus per-set control), once
//   per tile, and that unit is its busiest; here the walk is vector loads of
//   offsets and LDS reads.
// * A wave takes two stream sets at once — the pair (2i, 2i+1) — one per
//   half-wave.  Lane L of a half holds 128 tile sets (masks 2L and 2L+1 of the
//   row) and reads its 16 bytes of row q with one ds_read_b128, so one LDS
//   instruction serves both sets; gfx950 reads 128-bit lanes at twice the
//   rate of 64-bit ones.
//
// Index layout, per pair: G groups of 8 offsets, group k = 4 offsets of the
// first set then 4 of the second, so lane L of half h loads its group as one
// uint4 at entry 8k + 4h.  G is the larger set's state count rounded up to 4,
// over 4; the shorter set is padded with the all-ones padding row.  A pair's
// groups (up to kIxPrefetch) are loaded while the previous pair is walked.
#include <hip/hip_runtime.h>

#include <atomic>
#include <climits>

#include "sw_device.hpp"
#include "sw_kernels.hpp"

namespace sw {

namespace {

constexpr int kIxTile = 4096;
constexpr int kIxPrefetch = 8;  // uint4 groups of offsets (32 states) kept in registers

__device__ inline uint32_t and3u(uint32_t a, uint32_t b, uint32_t c) { return __builtin_amdgcn_bitop3_b32(a, b, c, 0x80); }

struct R4 {
  uint32_t v[4];
};

__device__ inline void fold2(R4& r, const uint4& a, const uint4& b) {
  r.v[0] = and3u(r.v[0], a.x, b.x);
  r.v[1] = and3u(r.v[1], a.y, b.y);
  r.v[2] = and3u(r.v[2], a.z, b.z);
  r.v[3] = and3u(r.v[3], a.w, b.w);
}

__device__ inline const uint4& row16(const unsigned char* smem, uint32_t byte) {
  return *reinterpret_cast<const uint4*>(smem + byte);
}

// Per pair i of stream sets (2i, 2i+1; the second may not exist): size[s] =
// |S_s| (complemented first when cmask is all ones) and entries[i] = 8 G, G =
// max of the two sizes rounded up to 4, over 4.
template <int W>
__global__ void k_ix_pair_count(const uint64_t* __restrict__ sets, uint64_t count, uint32_t n, uint64_t cmask,
    uint32_t* __restrict__ size, uint32_t* __restrict__ entries) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t pairs = (count + 1) / 2;
  if (i >= pairs) return;
  uint32_t c[2] = {0u, 0u};
  for (int h = 0; h < 2; ++h) {
    const uint64_t s = 2 * i + h;
    if (s >= count) break;
#pragma unroll
    for (int w = 0; w < W; ++w) c[h] += __popcll((sets[s * W + w] ^ cmask) & word_mask(n, w));
    size[s] = c[h];
  }
  const uint32_t g = (max(c[0], c[1]) + 3u) / 4u;
  entries[i] = 8u * g;
}

// list[start[i] + 8k + 4h + j] = row byte offset (q * 512) of the (4k + j)-th
// state of set 2i + h; the padding row's offset (n * 512) past its end.
template <int W>
__global__ void k_ix_pair_fill(const uint64_t* __restrict__ sets, uint64_t count, uint32_t n, uint64_t cmask,
    const uint32_t* __restrict__ start, uint32_t* __restrict__ list) {
  const uint64_t s = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t pairs = (count + 1) / 2;
  if (s >= 2 * pairs) return;
  const uint64_t i = s / 2;
  const uint32_t h = uint32_t(s & 1);
  uint32_t* out = list + start[i] + 4 * h;
  const uint32_t slots = (start[i + 1] - start[i]) / 2;  // 4 G
  uint32_t j = 0;
  if (s < count) {
#pragma unroll
    for (int w = 0; w < W; ++w) {
      uint64_t b = (sets[s * W + w] ^ cmask) & word_mask(n, w);
      while (b) {
        const uint32_t q = uint32_t(w) * 64 + __builtin_ctzll(b);
        b &= b - 1;
        out[8 * (j >> 2) + (j & 3)] = q * 512u;
        ++j;
      }
    }
  }
  for (; j < slots; ++j) out[8 * (j >> 2) + (j & 3)] = n * 512u;
}

template <int W, int MODE>
__global__ __launch_bounds__(1024) void k_vsub_ix(VsubParams p) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint64_t* rows = reinterpret_cast<uint64_t*>(smem);  // [(n + 1) rows][64 masks]
  const uint32_t n = p.n;
  uint16_t* tsize = reinterpret_cast<uint16_t*>(smem + size_t(n + 1) * 512);
  const int lane = threadIdx.x & 63;
  const int half = lane >> 5, L = lane & 31;
  const int wave = rfl(threadIdx.x >> 6);
  const int nwaves = blockDim.x >> 6;
  const uint64_t tile_base = uint64_t(blockIdx.x) * kIxTile;
  const uint64_t tile_left = p.t_count - tile_base;  // > 0 by grid construction

  // ---- transpose the tile (as k_vsub): mask g of row q = ballot over tile
  // sets [64 g, 64 g + 64) of "q is in the set".
  for (int g = wave; g < kIxTile / 64; g += nwaves) {
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
      uint64_t kept = 0;
      for (int b = 0; b < 64; ++b) {
        const uint64_t m = __builtin_amdgcn_ballot_w64((x >> b) & 1ull);
        kept = (lane == b) ? m : kept;
      }
      const uint32_t q = uint32_t(w) * 64 + lane;
      if (q < n) rows[size_t(q) * 64 + g] = kept;
    }
    if (p.proper_by_size) tsize[t] = uint16_t(valid ? pc : 0xffff);
  }
  if (wave == 0) rows[size_t(n) * 64 + lane] = ~0ull;  // padding row
  __syncthreads();

  // Lane L's 128 tile sets [128 L, 128 L + 128) that exist.
  R4 valid;
  {
    const int64_t c = int64_t(tile_left < uint64_t(kIxTile) ? tile_left : uint64_t(kIxTile)) - int64_t(L) * 128;
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      const int64_t cd = c - 32 * d;
      valid.v[d] = cd >= 32 ? ~0u : (cd <= 0 ? 0u : ((1u << cd) - 1u));
    }
  }

  const uint64_t s_begin = uint64_t(blockIdx.y) * p.s_chunk;  // even
  uint64_t s_end = s_begin + p.s_chunk;
  if (s_end > p.s_count) s_end = p.s_count;
  const uint64_t p_end = (s_end + 1) / 2;  // pairs [s_begin / 2, p_end)
  const uint32_t* __restrict__ start = p.ix_start;
  const uint4* __restrict__ list4 = reinterpret_cast<const uint4*>(p.ix_list);
  const uint32_t lane_off = uint32_t(L) * 16u;

  uint64_t cnt = 0;
  R4 hits = {{0u, 0u, 0u, 0u}};
  uint32_t iter = 0;

  // Walk one pair whose first groups are in `cur` (G groups from entry e0),
  // while the next pair's first groups load into `nxt`.
  auto walk = [&](uint64_t pi, uint32_t e0, uint32_t G, uint4 (&cur)[kIxPrefetch], uint64_t pn, uint32_t& en,
                  uint32_t& Gn, uint4 (&nxt)[kIxPrefetch]) {
    if (pn < p_end) {
      en = rfl(start[pn]);
      Gn = (rfl(start[pn + 1]) - en) / 8u;
    } else {
      en = 0;
      Gn = 0;
    }
#pragma unroll
    for (int k = 0; k < kIxPrefetch; ++k)
      if (uint32_t(k) < Gn) nxt[k] = list4[(en >> 2) + 2 * k + half];
    const uint64_t sh = 2 * pi + half;
    const bool live = sh < s_end;
    R4 r = valid;
    if (!live) r = R4{{0u, 0u, 0u, 0u}};
#pragma unroll
    for (int k = 0; k < kIxPrefetch; ++k) {
      if (uint32_t(k) >= G) break;
      const uint4 o = cur[k];
      const uint4 m0 = row16(smem, o.x + lane_off);
      const uint4 m1 = row16(smem, o.y + lane_off);
      const uint4 m2 = row16(smem, o.z + lane_off);
      const uint4 m3 = row16(smem, o.w + lane_off);
      fold2(r, m0, m1);
      fold2(r, m2, m3);
    }
    for (uint32_t k = kIxPrefetch; k < G; ++k) {  // sets of more than 4 kIxPrefetch states
      const uint4 o = list4[(e0 >> 2) + 2 * k + half];
      const uint4 m0 = row16(smem, o.x + lane_off);
      const uint4 m1 = row16(smem, o.y + lane_off);
      const uint4 m2 = row16(smem, o.z + lane_off);
      const uint4 m3 = row16(smem, o.w + lane_off);
      fold2(r, m0, m1);
      fold2(r, m2, m3);
    }

    if (live && p.diag != INT64_MIN) {
      // Skip the pair (s, t) with s - t == diag: the same element of one list.
      const int64_t t = int64_t(sh) - p.diag - int64_t(tile_base);
      if (t >= 0 && t < kIxTile && L == int(t / 128)) {
        const int d = int(t % 128) / 32;  // constant indices only: r stays in registers
#pragma unroll
        for (int e = 0; e < 4; ++e)
          if (e == d) r.v[e] &= ~(1u << (t % 32));
      }
    }
    if (live && p.proper_by_size) {
      // A stream set inside a tile set of equal cardinality is that set.
      const uint32_t sz = p.ix_size[sh];
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        uint32_t rr = r.v[d];
        while (rr) {
          const int b = __builtin_ctz(rr);
          rr &= rr - 1;
          if (tsize[L * 128 + d * 32 + b] == sz) r.v[d] &= ~(1u << b);
        }
      }
    }
    const bool any = (r.v[0] | r.v[1] | r.v[2] | r.v[3]) != 0;
    if (MODE == VSUB_COUNT) {
      cnt += __popc(r.v[0]) + __popc(r.v[1]) + __popc(r.v[2]) + __popc(r.v[3]);
    } else if (MODE == VSUB_SCOUNT) {
      uint32_t pc = __popc(r.v[0]) + __popc(r.v[1]) + __popc(r.v[2]) + __popc(r.v[3]);
      for (int off = 16; off > 0; off >>= 1) pc += __shfl_xor(pc, off, 64);  // within the half
      if (L == 0 && live && pc) atomicAdd(&p.scount[sh], pc);
    } else if (MODE == VSUB_FLAGS) {
#pragma unroll
      for (int d = 0; d < 4; ++d) hits.v[d] |= r.v[d];
    } else if (MODE == VSUB_SFLAGS) {
      const uint64_t bal = __builtin_amdgcn_ballot_w64(any);
      if (L == 0 && live && ((bal >> (32 * half)) & 0xffffffffull))
        atomicOr(reinterpret_cast<unsigned long long*>(p.flags) + (sh >> 6), 1ull << (sh & 63));
    } else if (MODE == VSUB_ANY) {
      if (any) {
        int b = 0;
        for (int d = 0; d < 4; ++d)
          if (r.v[d]) {
            b = d * 32 + __builtin_ctz(r.v[d]);
            break;
          }
        const unsigned long long wv = (sh << 32) | (tile_base + uint64_t(L) * 128 + uint64_t(b));
        atomicMin(p.witness, wv);
        __hip_atomic_store(p.found, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  };

  // Pairs of this wave: s_begin / 2 + wave, then every nwaves-th.  Two
  // register buffers take turns (no copies).
  uint4 bufA[kIxPrefetch], bufB[kIxPrefetch];
  uint64_t pi = s_begin / 2 + uint64_t(wave);
  uint32_t eA = 0, GA = 0, eB = 0, GB = 0;
  if (pi < p_end) {
    eA = rfl(start[pi]);
    GA = (rfl(start[pi + 1]) - eA) / 8u;
#pragma unroll
    for (int k = 0; k < kIxPrefetch; ++k)
      if (uint32_t(k) < GA) bufA[k] = list4[(eA >> 2) + 2 * k + half];
  }
  while (pi < p_end) {
    if (MODE == VSUB_ANY && (iter++ & 7) == 0) {
      if (rfl(__hip_atomic_load(p.found, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) break;
    }
    walk(pi, eA, GA, bufA, pi + nwaves, eB, GB, bufB);
    pi += nwaves;
    if (pi >= p_end) break;
    walk(pi, eB, GB, bufB, pi + nwaves, eA, GA, bufA);
    pi += nwaves;
  }

  if (MODE == VSUB_COUNT) {
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    if (lane == 0 && cnt) atomicAdd(p.count_out, (unsigned long long)cnt);
  } else if (MODE == VSUB_FLAGS) {
    // Both halves cover the same tile sets: merge, then lane L writes masks 2L, 2L+1.
#pragma unroll
    for (int d = 0; d < 4; ++d) hits.v[d] |= __shfl_xor(hits.v[d], 32, 64);
    if (half == 0) {
      unsigned long long* f = reinterpret_cast<unsigned long long*>(p.flags) + tile_base / 64 + 2 * L;
      const uint64_t lo = uint64_t(hits.v[0]) | (uint64_t(hits.v[1]) << 32);
      const uint64_t hi = uint64_t(hits.v[2]) | (uint64_t(hits.v[3]) << 32);
      if (lo) atomicOr(f, lo);
      if (hi) atomicOr(f + 1, hi);
    }
  }
}

template <int W, int MODE>
hipError_t launch_ix_mode(const VsubParams& p, uint64_t gx, uint64_t gy, size_t lds, hipStream_t st) {
  static std::atomic<bool> attr_set{false};
  if (!attr_set.load(std::memory_order_acquire)) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_vsub_ix<W, MODE>),
        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e != hipSuccess) return e;
    attr_set.store(true, std::memory_order_release);
  }
  hipLaunchKernelGGL((k_vsub_ix<W, MODE>), dim3(unsigned(gx), unsigned(gy)), dim3(1024), lds, st, p);
  return hipGetLastError();
}

template <int W>
hipError_t launch_ix_w(const VsubParams& p0, hipStream_t st) {
  VsubParams p = p0;
  if (p.t_count == 0 || p.s_count == 0) return hipSuccess;
  const uint64_t gx = (p.t_count + kIxTile - 1) / kIxTile;
  if (p.s_chunk == 0) {
    const uint64_t target = p.target_wgs ? p.target_wgs : 1024;
    const uint64_t gyt = (target + gx - 1) / gx;
    uint64_t chunk = (p.s_count + gyt - 1) / gyt;
    if (chunk < 128) chunk = 128;
    p.s_chunk = chunk;
  }
  p.s_chunk = (p.s_chunk + 1) & ~1ull;  // even: pairs never straddle two workgroups
  const uint64_t gy = (p.s_count + p.s_chunk - 1) / p.s_chunk;
  if (gx > 0x7fffffffull || gy > 65535ull) return hipErrorInvalidValue;
  const size_t lds = size_t(p.n + 1) * 512 + (p.proper_by_size ? size_t(kIxTile) * 2 : 0);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  switch (p.mode) {
    case VSUB_COUNT: return launch_ix_mode<W, VSUB_COUNT>(p, gx, gy, lds, st);
    case VSUB_ANY: return launch_ix_mode<W, VSUB_ANY>(p, gx, gy, lds, st);
    case VSUB_FLAGS: return launch_ix_mode<W, VSUB_FLAGS>(p, gx, gy, lds, st);
    case VSUB_SCOUNT: return launch_ix_mode<W, VSUB_SCOUNT>(p, gx, gy, lds, st);
    case VSUB_SFLAGS: return launch_ix_mode<W, VSUB_SFLAGS>(p, gx, gy, lds, st);
  }
  return hipErrorInvalidValue;
}

}  // namespace

bool vsub_ix_fits(uint32_t n, int W, bool proper_by_size) {
  return W >= 1 && W <= 5 && size_t(n + 1) * 512 + (proper_by_size ? size_t(kIxTile) * 2 : 0) <= 160 * 1024;
}

hipError_t launch_ix_pair_count(int W, const uint64_t* sets, uint64_t count, uint32_t n, int complement,
    uint32_t* size, uint32_t* entries, hipStream_t st) {
  if (!count) return hipSuccess;
  const uint64_t cmask = complement ? ~0ull : 0ull;
  const unsigned blocks = unsigned(((count + 1) / 2 + 255) / 256);
  switch (W) {
#define C(w) case w: hipLaunchKernelGGL(k_ix_pair_count<w>, dim3(blocks), dim3(256), 0, st, sets, count, n, cmask, size, entries); break;
    C(1) C(2) C(3) C(4) C(5)
#undef C
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_ix_pair_fill(int W, const uint64_t* sets, uint64_t count, uint32_t n, int complement,
    const uint32_t* start, uint32_t* list, hipStream_t st) {
  if (!count) return hipSuccess;
  const uint64_t cmask = complement ? ~0ull : 0ull;
  const unsigned blocks = unsigned(((count + 1) / 2 * 2 + 255) / 256);
  switch (W) {
#define C(w) case w: hipLaunchKernelGGL(k_ix_pair_fill<w>, dim3(blocks), dim3(256), 0, st, sets, count, n, cmask, start, list); break;
    C(1) C(2) C(3) C(4) C(5)
#undef C
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_vsub_ix(int W, const VsubParams& p, hipStream_t st) {
  if (!p.ix_start || !p.ix_list) return hipErrorInvalidValue;
  if (p.mode == VSUB_ANY && (!p.found || !p.witness)) return hipErrorInvalidValue;
  switch (W) {
    case 1: return launch_ix_w<1>(p, st);
    case 2: return launch_ix_w<2>(p, st);
    case 3: return launch_ix_w<3>(p, st);
    case 4: return launch_ix_w<4>(p, st);
    case 5: return launch_ix_w<5>(p, st);
  }
  return hipErrorInvalidValue;
}

}  // namespace sw
