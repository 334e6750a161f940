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
// over 4; the shorter set is padded with the all-ones padding row.
//
// Two walks.  The plain one (counts) walks every state of every pair, a
// pair's groups (up to kIxPrefetch) loaded while the previous pair is walked.
// The early-exit one (the exact search's launches) lists states rarest first,
// ballots after every group and drops a pair once neither half has a tile
// set left; its pairs' first kIxHead groups sit at a fixed stride (ix_head)
// and are loaded kIxRing pairs ahead.  Around either walk the host may sort
// both lists by size (a workgroup then stops at its tile's largest set) or
// bucket them by anchor (a work list of (tile slice, stream range) items);
// results are mapped back through the permutations.  Tuning constants below
// can be set at build time (-D) for sweeps; the defaults are the measured
// best.
#include <hip/hip_runtime.h>

#include <atomic>
#include <climits>

#include "sw_device.hpp"
#include "sw_kernels.hpp"

namespace sw {

namespace {

constexpr int kIxTile = 4096;
#ifndef SW_IX_PREFETCH
#define SW_IX_PREFETCH 8
#endif
constexpr int kIxPrefetch = SW_IX_PREFETCH;  // uint4 groups of offsets (32 states) kept in registers
// With early exit: groups per pair kept at a fixed stride (ix_head), and
// pairs whose head groups are loaded ahead per wave.
#ifndef SW_IX_HEAD
#define SW_IX_HEAD 2
#endif
#ifndef SW_IX_RING
#define SW_IX_RING 8
#endif
#ifndef SW_IX_REV
#define SW_IX_REV 1
#endif
#ifndef SW_IX_EE_CHUNK
#define SW_IX_EE_CHUNK 131072
#endif
#ifndef SW_IX_EE_WGS
#define SW_IX_EE_WGS 512
#endif
constexpr int kIxHead = SW_IX_HEAD;
constexpr int kIxRing = SW_IX_RING;

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
__global__ void k_ix_pair_count(const uint64_t* __restrict__ sets, const uint32_t* __restrict__ perm, uint64_t count,
    uint32_t n, uint64_t cmask, uint32_t* __restrict__ size, uint32_t* __restrict__ entries) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t pairs = (count + 1) / 2;
  if (i >= pairs) return;
  uint32_t c[2] = {0u, 0u};
  for (int h = 0; h < 2; ++h) {
    const uint64_t s = 2 * i + h;
    if (s >= count) break;
    const uint64_t* set = sets + (perm ? uint64_t(perm[s]) : s) * W;
#pragma unroll
    for (int w = 0; w < W; ++w) c[h] += __popcll((set[w] ^ cmask) & word_mask(n, w));
    size[s] = c[h];
  }
  const uint32_t g = (max(c[0], c[1]) + 3u) / 4u;
  entries[i] = 8u * g;
}

// list[start[i] + 8k + 4h + j] = row byte offset (q * 512) of the (4k + j)-th
// state of set 2i + h; the padding row's offset (n * 512) past its end.
// States are listed in ascending q, or, when RANKED, in ascending rank[q]
// (the set is first re-spelled over ranks, then walked; order[r] is the row
// offset of the state of rank r).
// When head != NULL, the first kIxHead groups of every pair are also
// written at a fixed stride: head[((i kIxHead + g) 2 + h) 4 + j], padded.
template <int W, bool RANKED>
__global__ void k_ix_pair_fill(const uint64_t* __restrict__ sets, const uint32_t* __restrict__ perm, uint64_t count,
    uint32_t n, uint64_t cmask, const uint32_t* __restrict__ start, const uint32_t* __restrict__ rank,
    const uint32_t* __restrict__ order, uint32_t* __restrict__ list, uint32_t* __restrict__ head) {
  // The rank and order tables (n <= 512 words each) in LDS: a set's walk
  // looks one up per state.
  __shared__ uint32_t s_rank[RANKED ? 512 : 1], s_order[RANKED ? 512 : 1];
  if (RANKED) {
    for (uint32_t t = threadIdx.x; t < n; t += blockDim.x) {
      s_rank[t] = rank[t];
      s_order[t] = order[t];
    }
    __syncthreads();
  }
  const uint64_t s = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t pairs = (count + 1) / 2;
  if (s >= 2 * pairs) return;
  const uint64_t i = s / 2;
  const uint32_t h = uint32_t(s & 1);
  uint32_t* out = list + start[i] + 4 * h;
  const uint32_t slots = (start[i + 1] - start[i]) / 2;  // 4 G
  auto emit = [&](uint32_t j, uint32_t v) {
    if (j < slots) out[8 * (j >> 2) + (j & 3)] = v;
    if (head && j < 4u * kIxHead) head[((i * kIxHead + (j >> 2)) * 2 + h) * 4 + (j & 3)] = v;
  };
  uint32_t j = 0;
  if (s < count) {
    const uint64_t* set = sets + (perm ? uint64_t(perm[s]) : s) * W;
    uint64_t x[W];
#pragma unroll
    for (int w = 0; w < W; ++w) x[w] = (set[w] ^ cmask) & word_mask(n, w);
    if (RANKED) {
      uint64_t y[W];
#pragma unroll
      for (int w = 0; w < W; ++w) y[w] = 0;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        uint64_t b = x[w];
        while (b) {
          const uint32_t r = s_rank[uint32_t(w) * 64 + __builtin_ctzll(b)];
          b &= b - 1;
#pragma unroll
          for (int v = 0; v < W; ++v)  // constant indices only: y stays in registers
            if (int(r >> 6) == v) y[v] |= 1ull << (r & 63);
        }
      }
#pragma unroll
      for (int w = 0; w < W; ++w) x[w] = y[w];
    }
#pragma unroll
    for (int w = 0; w < W; ++w) {
      uint64_t b = x[w];
      while (b) {
        const uint32_t q = uint32_t(w) * 64 + __builtin_ctzll(b);
        b &= b - 1;
        emit(j, RANKED ? s_order[q] : q * 512u);
        ++j;
      }
    }
  }
  const uint32_t pad_to = head ? max(slots, 4u * kIxHead) : slots;
  for (; j < pad_to; ++j) emit(j, n * 512u);
}

// freq[q] += sets holding state q.  A wave takes 64 sets at a time; ballot b
// of word w says which of them hold state 64 w + b, and lane b keeps the
// count of that state.
template <int W>
__global__ __launch_bounds__(256) void k_state_freq(const uint64_t* __restrict__ sets, uint64_t count, uint32_t n,
    uint64_t cmask, uint32_t* __restrict__ freq) {
  __shared__ uint32_t acc[W * 64];
  for (int t = threadIdx.x; t < W * 64; t += blockDim.x) acc[t] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const uint64_t waves = uint64_t(gridDim.x) * (blockDim.x >> 6);
  uint32_t c[W];
#pragma unroll
  for (int w = 0; w < W; ++w) c[w] = 0;
  for (uint64_t g = uint64_t(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6); g * 64 < count; g += waves) {
    const uint64_t t = g * 64 + lane;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const uint64_t x = t < count ? ((sets[t * W + w] ^ cmask) & word_mask(n, w)) : 0ull;
      c[w] += uint32_t(__popcll(lane_transpose64(x)));  // lane b: sets of the 64 holding state w*64 + b
    }
  }
#pragma unroll
  for (int w = 0; w < W; ++w)
    if (c[w]) atomicAdd(&acc[w * 64 + lane], c[w]);
  __syncthreads();
  for (int t = threadIdx.x; t < W * 64; t += blockDim.x)
    if (uint32_t(t) < n && acc[t]) atomicAdd(&freq[t], acc[t]);
}

// hist[c] += sets of cardinality c (0 <= c <= n); hist must start at zero.
template <int W>
__global__ __launch_bounds__(256) void k_size_hist(const uint64_t* __restrict__ sets, uint64_t count, uint32_t n,
    uint64_t cmask, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[W * 64 + 1];
  for (int t = threadIdx.x; t <= W * 64; t += blockDim.x) h[t] = 0;
  __syncthreads();
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += uint64_t(gridDim.x) * blockDim.x) {
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) c += __popcll((sets[i * W + w] ^ cmask) & word_mask(n, w));
    atomicAdd(&h[c], 1u);
  }
  __syncthreads();
  for (int t = threadIdx.x; t <= W * 64; t += blockDim.x)
    if (uint32_t(t) <= n && h[t]) atomicAdd(&hist[t], h[t]);
}

// Exclusive (cursor) and inclusive (bound) prefix sums of hist[0..n], n < 1024:
// one workgroup of 1024 threads, a scan per wave, then the earlier waves'
// totals.  (One thread walking the bins took 16 us.)
__global__ __launch_bounds__(1024) void k_size_offsets(const uint32_t* __restrict__ hist, uint32_t n,
    uint32_t* __restrict__ cursor, uint32_t* __restrict__ bound) {
  __shared__ uint32_t wsum[16];
  const uint32_t t = threadIdx.x, lane = t & 63u, wv = t >> 6;
  const uint32_t v = t <= n ? hist[t] : 0u;
  uint32_t x = v;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, 64);
    if (lane >= uint32_t(off)) x += y;
  }
  if (lane == 63) wsum[wv] = x;
  __syncthreads();
  uint32_t before = 0;
  for (uint32_t w = 0; w < wv; ++w) before += wsum[w];
  if (t <= n) {
    cursor[t] = before + x - v;
    if (bound) bound[t] = before + x;
  }
}

// A workgroup takes kScatterPer sets per thread, counts them per size in
// LDS, reserves one range per size with a single global atomic, then writes.
constexpr int kScatterPer = 8;

template <int W>
__global__ __launch_bounds__(256) void k_size_scatter(const uint64_t* __restrict__ sets, uint64_t count, uint32_t n,
    uint64_t cmask, uint32_t* __restrict__ cursor, uint32_t* __restrict__ perm) {
  __shared__ uint32_t cnt[W * 64 + 1], base[W * 64 + 1];
  for (int t = threadIdx.x; t <= W * 64; t += blockDim.x) cnt[t] = 0;
  __syncthreads();
  const uint64_t i0 = uint64_t(blockIdx.x) * blockDim.x * kScatterPer + threadIdx.x;
  uint32_t c[kScatterPer], local[kScatterPer];
#pragma unroll
  for (int e = 0; e < kScatterPer; ++e) {
    const uint64_t i = i0 + uint64_t(e) * blockDim.x;
    c[e] = 0;
    local[e] = 0;
    if (i < count) {
#pragma unroll
      for (int w = 0; w < W; ++w) c[e] += __popcll((sets[i * W + w] ^ cmask) & word_mask(n, w));
      local[e] = atomicAdd(&cnt[c[e]], 1u);
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t <= W * 64; t += blockDim.x)
    if (uint32_t(t) <= n && cnt[t]) base[t] = atomicAdd(&cursor[t], cnt[t]);
  __syncthreads();
#pragma unroll
  for (int e = 0; e < kScatterPer; ++e) {
    const uint64_t i = i0 + uint64_t(e) * blockDim.x;
    if (i < count) perm[base[c[e]] + local[e]] = uint32_t(i);
  }
}

// Anchor of each stream set: its state of lowest rank (fewest tile sets hold
// it), or n for the empty set; hist[anchor] += 1 (hist starts at zero).
template <int W>
__global__ __launch_bounds__(256) void k_anchor(const uint64_t* __restrict__ sets, uint64_t count, uint32_t n,
    uint64_t cmask, const uint32_t* __restrict__ rank, uint32_t* __restrict__ anchor, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[W * 64 + 1], s_rank[W * 64];
  for (int t = threadIdx.x; t <= W * 64; t += blockDim.x) h[t] = 0;
  for (uint32_t t = threadIdx.x; t < n; t += blockDim.x) s_rank[t] = rank[t];
  __syncthreads();
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) {
    uint32_t best = 0xffffffffu, a = n;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      uint64_t b = (sets[i * W + w] ^ cmask) & word_mask(n, w);
      while (b) {
        const uint32_t q = uint32_t(w) * 64 + __builtin_ctzll(b);
        b &= b - 1;
        const uint32_t r = s_rank[q];
        if (r < best) {
          best = r;
          a = q;
        }
      }
    }
    anchor[i] = a;
    atomicAdd(&h[a], 1u);
  }
  __syncthreads();
  for (int t = threadIdx.x; t <= W * 64; t += blockDim.x)
    if (uint32_t(t) <= n && h[t]) atomicAdd(&hist[t], h[t]);
}

// perm[cursor[key[i]]++] = i, with one global atomic per key per workgroup
// (keys < nkeys <= 513).
__global__ __launch_bounds__(256) void k_key_scatter(const uint32_t* __restrict__ key, uint64_t count, uint32_t nkeys,
    uint32_t* __restrict__ cursor, uint32_t* __restrict__ perm) {
  __shared__ uint32_t cnt[513], base[513];
  for (uint32_t t = threadIdx.x; t < nkeys; t += blockDim.x) cnt[t] = 0;
  __syncthreads();
  const uint64_t i0 = uint64_t(blockIdx.x) * blockDim.x * kScatterPer + threadIdx.x;
  uint32_t c[kScatterPer], local[kScatterPer];
#pragma unroll
  for (int e = 0; e < kScatterPer; ++e) {
    const uint64_t i = i0 + uint64_t(e) * blockDim.x;
    c[e] = 0;
    local[e] = 0;
    if (i < count) {
      c[e] = key[i];
      local[e] = atomicAdd(&cnt[c[e]], 1u);
    }
  }
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < nkeys; t += blockDim.x)
    if (cnt[t]) base[t] = atomicAdd(&cursor[t], cnt[t]);
  __syncthreads();
#pragma unroll
  for (int e = 0; e < kScatterPer; ++e) {
    const uint64_t i = i0 + uint64_t(e) * blockDim.x;
    if (i < count) perm[base[c[e]] + local[e]] = uint32_t(i);
  }
}

// Inverted lists of the tile: for every set t and state q of it,
// lists[cursor[q]++] = t (one global atomic per state per workgroup).
template <int W>
__global__ __launch_bounds__(256) void k_invert(const uint64_t* __restrict__ sets, uint64_t count, uint32_t n,
    uint64_t cmask, uint32_t* __restrict__ cursor, uint32_t* __restrict__ lists) {
  __shared__ uint32_t cnt[W * 64], base[W * 64];
  for (int t = threadIdx.x; t < W * 64; t += blockDim.x) cnt[t] = 0;
  __syncthreads();
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  uint64_t x[W];
#pragma unroll
  for (int w = 0; w < W; ++w) x[w] = i < count ? ((sets[i * W + w] ^ cmask) & word_mask(n, w)) : 0ull;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    uint64_t b = x[w];
    while (b) {
      atomicAdd(&cnt[w * 64 + __builtin_ctzll(b)], 1u);
      b &= b - 1;
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < W * 64; t += blockDim.x)
    if (uint32_t(t) < n && cnt[t]) {
      base[t] = atomicAdd(&cursor[t], cnt[t]);
      cnt[t] = 0;
    }
  __syncthreads();
#pragma unroll
  for (int w = 0; w < W; ++w) {
    uint64_t b = x[w];
    while (b) {
      const uint32_t q = uint32_t(w) * 64 + __builtin_ctzll(b);
      b &= b - 1;
      lists[base[q] + atomicAdd(&cnt[q], 1u)] = uint32_t(i);
    }
  }
}

// One workgroup: rank[q] = #{q' : (freq[q'], q') < (freq[q], q)}.
__global__ __launch_bounds__(512) void k_state_rank(const uint32_t* __restrict__ freq, uint32_t n,
    uint32_t* __restrict__ rank, uint32_t* __restrict__ order) {
  __shared__ uint32_t f[512];
  for (uint32_t t = threadIdx.x; t < n; t += blockDim.x) f[t] = freq[t];
  __syncthreads();
  for (uint32_t q = threadIdx.x; q < n; q += blockDim.x) {
    uint32_t r = 0;
    const uint32_t fq = f[q];
    for (uint32_t u = 0; u < n; ++u) r += (f[u] < fq || (f[u] == fq && u < q)) ? 1u : 0u;
    rank[q] = r;
    order[r] = q * 512u;
  }
}

template <int W, int MODE, bool EE>
__global__ __launch_bounds__(1024) void k_vsub_ix(VsubParams p) {
  extern __shared__ __align__(16) unsigned char smem[];
  uint64_t* rows = reinterpret_cast<uint64_t*>(smem);  // [(n + 1) rows][64 masks]
  const uint32_t n = p.n;
  uint16_t* tsize = reinterpret_cast<uint16_t*>(smem + size_t(n + 1) * 512);
  const int lane = threadIdx.x & 63;
  const int half = lane >> 5, L = lane & 31;
  const int wave = rfl(threadIdx.x >> 6);
  const int nwaves = blockDim.x >> 6;
  // Lists sorted by size: the tiles of the largest sets, whose workgroups
  // walk the most stream sets, are dispatched first.  With a work list
  // (anchor buckets), the workgroup's item names its tile (up to 4096 entries
  // of t_perm from x) and its stream sets [z, w).
  uint64_t tile_base, tile_left, s_begin, s_end;
  if (p.work) {
    const uint4 wi = p.work[blockIdx.x];
    tile_base = wi.x;
    tile_left = wi.y;
    s_begin = wi.z;  // even
    s_end = wi.w;
  } else {
    const uint32_t tile_ix = (SW_IX_REV && p.t_perm) ? gridDim.x - 1u - blockIdx.x : blockIdx.x;
    tile_base = uint64_t(tile_ix) * kIxTile;
    tile_left = p.t_count - tile_base;  // > 0 by grid construction
    s_begin = uint64_t(blockIdx.y) * p.s_chunk;  // even
    s_end = s_begin + p.s_chunk;
    if (s_end > p.s_count) s_end = p.s_count;
  }

  // Lists sorted by size: the tile's largest set has tmax states, the
  // smallest c with t_bound[c] past the tile's last set; no stream set past
  // the first s_bound[tmax] fits in a tile set (the pair holding the last
  // that may is walked whole).  A workgroup with no such pair in its chunk
  // leaves before loading its tile.
  uint64_t p_bound = UINT64_MAX;
  if (p.s_bound) {
    const uint64_t last = tile_base + (tile_left < uint64_t(kIxTile) ? tile_left : uint64_t(kIxTile)) - 1;
    uint32_t lo = 0, hi = n;  // t_bound[n] = t_count > last
    while (lo < hi) {
      const uint32_t mid = (lo + hi) / 2;
      if (p.t_bound[mid] > last) hi = mid;
      else lo = mid + 1;
    }
    p_bound = (uint64_t(p.s_bound[lo]) + 1) / 2;
    if (s_begin / 2 >= p_bound) return;  // the whole workgroup, before any barrier
  }

  // ---- transpose the tile (as k_vsub): mask g of row q = bits over tile
  // sets [64 g, 64 g + 64) of "q is in the set"; groups past the tile's last
  // set are skipped (read only under a zero valid mask).
  const int ngroups = tile_left < uint64_t(kIxTile) ? int((tile_left + 63) / 64) : kIxTile / 64;
  for (int g = wave; g < ngroups; g += nwaves) {
    const uint64_t t = uint64_t(g) * 64 + lane;
    const bool valid = t < tile_left;
    const uint64_t* set = p.tile + (valid && p.t_perm ? uint64_t(p.t_perm[tile_base + t]) : tile_base + t) * W;
    int pc = 0;
    for (int w = 0; w < W; ++w) {
      uint64_t x = 0;
      if (valid) {
        x = set[w];
        if (p.complement_tile) x = ~x & word_mask(n, w);
      }
      pc += __popcll(x);
      const uint64_t kept = lane_transpose64(x);
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

  uint64_t p_end = (s_end + 1) / 2;  // pairs [s_begin / 2, p_end)
  if (p_bound < p_end) p_end = p_bound;
  const uint32_t* __restrict__ start = p.ix_start;
  const uint4* __restrict__ list4 = reinterpret_cast<const uint4*>(p.ix_list);
  const uint32_t lane_off = uint32_t(L) * 16u;

  uint64_t cnt = 0;
  R4 hits = {{0u, 0u, 0u, 0u}};
  uint32_t iter = 0;

  // Fold one group of four states per half into r; false once neither half
  // has a tile set left (the accumulator only loses bits).
  auto fold_group = [&](R4& r, const uint4& o) {
    const uint4 m0 = row16(smem, o.x + lane_off);
    const uint4 m1 = row16(smem, o.y + lane_off);
    const uint4 m2 = row16(smem, o.z + lane_off);
    const uint4 m3 = row16(smem, o.w + lane_off);
    fold2(r, m0, m1);
    fold2(r, m2, m3);
    return __builtin_amdgcn_ballot_w64((r.v[0] | r.v[1] | r.v[2] | r.v[3]) != 0) != 0;
  };

  // The pair's result for stream set sh (this half), r = its tile sets.
  auto finish = [&](uint64_t sh, bool live, R4& r) {
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
      if (L == 0 && live && pc) atomicAdd(&p.scount[p.s_perm ? uint64_t(p.s_perm[sh]) : sh], pc);
    } else if (MODE == VSUB_FLAGS) {
#pragma unroll
      for (int d = 0; d < 4; ++d) hits.v[d] |= r.v[d];
    } else if (MODE == VSUB_SFLAGS) {
      const uint64_t bal = __builtin_amdgcn_ballot_w64(any);
      if (L == 0 && live && ((bal >> (32 * half)) & 0xffffffffull)) {
        const uint64_t si = p.s_perm ? uint64_t(p.s_perm[sh]) : sh;
        atomicOr(reinterpret_cast<unsigned long long*>(p.flags) + (si >> 6), 1ull << (si & 63));
      }
    } else if (MODE == VSUB_ANY) {
      if (any) {
        int b = 0;
        for (int d = 0; d < 4; ++d)
          if (r.v[d]) {
            b = d * 32 + __builtin_ctz(r.v[d]);
            break;
          }
        const uint64_t tp = tile_base + uint64_t(L) * 128 + uint64_t(b);
        const uint64_t si = p.s_perm ? uint64_t(p.s_perm[sh]) : sh;
        const uint64_t ti = p.t_perm ? uint64_t(p.t_perm[tp]) : tp;
        const unsigned long long wv = (si << 32) | ti;
        atomicMin(p.witness, wv);
        __hip_atomic_store(p.found, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  };

  const uint64_t pi0 = s_begin / 2 + uint64_t(wave);  // pairs of this wave: pi0, then every nwaves-th
  if (EE) {
    // Most stream sets are dropped within the first groups, so a pair's cost
    // is the latency of finding its states: the first kIxHead groups of
    // every pair sit at a fixed stride (ix_head), loaded kIxRing pairs ahead
    // without a lookup; only a pair still alive after them reads its start.
    // Every load of the ring is unconditional (past the last pair it reads
    // the last pair again), so the compiler can count the loads in flight and
    // wait for exactly the one it needs.
    const uint4* __restrict__ head4 = reinterpret_cast<const uint4*>(p.ix_head);
    const uint32_t pe = uint32_t(p_end);  // pairs < 2^30
    const uint32_t last = pe ? pe - 1u : 0u;
    const uint32_t stride = uint32_t(kIxRing) * uint32_t(nwaves);
    auto head_at = [&](uint32_t pk, int g) {  // the list has at least one pair
      const uint32_t pc = pk < pe ? pk : last;
      return head4[(uint64_t(pc) * kIxHead + g) * 2 + half];
    };
    uint4 ring[kIxRing][kIxHead];
    const uint32_t p0 = uint32_t(pi0);
    // Issued in slot order (the scheduler may not reorder them), as the loop
    // reloads them, so the loop's first slot waits for its own loads only.
#pragma unroll
    for (int k = 0; k < kIxRing; ++k) {
#pragma unroll
      for (int g = 0; g < kIxHead; ++g) ring[k][g] = head_at(p0 + uint32_t(k) * nwaves, g);
      __builtin_amdgcn_sched_barrier(0);
    }
    for (uint32_t base = p0; base < pe; base += stride) {
      int stop = 0;
      if (MODE == VSUB_ANY) stop = __hip_atomic_load(p.found, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
      for (int k = 0; k < kIxRing; ++k) {
        const uint32_t pk = base + uint32_t(k) * nwaves;
        // Past the wave's last pair nothing is folded, but the ring is still
        // reloaded, so every path issues the same loads.
        // A half past the end of the list holds padding only and stays alive
        // with its valid mask; it is cleared before its result is taken.
        R4 r = valid;
        bool alive = pk < pe;
        if (alive) {
#pragma unroll
          for (int g = 0; g < kIxHead; ++g) {
            if (!fold_group(r, ring[k][g])) {
              alive = false;
              break;
            }
          }
        }
#pragma unroll
        for (int g = 0; g < kIxHead; ++g) ring[k][g] = head_at(pk + stride, g);
        if (alive) {  // else r is zero in both halves: no result to take
          const uint32_t e0 = rfl(start[pk]);
          const uint32_t G = (rfl(start[pk + 1]) - e0) / 8u;
          for (uint32_t g = kIxHead; g < G; ++g)
            if (!fold_group(r, list4[(e0 >> 2) + 2 * g + half])) break;
          const uint64_t sh = 2 * uint64_t(pk) + half;
          const bool live = sh < s_end;
          if (!live) r = R4{{0u, 0u, 0u, 0u}};
          finish(sh, live, r);
        }
      }
      if (MODE == VSUB_ANY && rfl(stop)) break;
    }
  } else {
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
      finish(sh, live, r);
    };

    // Two register buffers take turns (no copies).
    uint4 bufA[kIxPrefetch], bufB[kIxPrefetch];
    uint64_t pi = pi0;
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
  }

  if (MODE == VSUB_COUNT) {
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off, 64);
    if (lane == 0 && cnt) atomicAdd(p.count_out, (unsigned long long)cnt);
  } else if (MODE == VSUB_FLAGS) {
    // Both halves cover the same tile sets: merge, then lane L writes masks 2L, 2L+1.
#pragma unroll
    for (int d = 0; d < 4; ++d) hits.v[d] |= __shfl_xor(hits.v[d], 32, 64);
    if (half == 0 && !p.t_perm) {
      unsigned long long* f = reinterpret_cast<unsigned long long*>(p.flags) + tile_base / 64 + 2 * L;
      const uint64_t lo = uint64_t(hits.v[0]) | (uint64_t(hits.v[1]) << 32);
      const uint64_t hi = uint64_t(hits.v[2]) | (uint64_t(hits.v[3]) << 32);
      if (lo) atomicOr(f, lo);
      if (hi) atomicOr(f + 1, hi);
    } else if (half == 0) {
      // Sorted tile: each hit is flagged at its set's place in the list.
      unsigned long long* f = reinterpret_cast<unsigned long long*>(p.flags);
#pragma unroll
      for (int d = 0; d < 4; ++d) {
        uint32_t m = hits.v[d];
        while (m) {
          const int b = __builtin_ctz(m);
          m &= m - 1;
          const uint64_t t = p.t_perm[tile_base + uint64_t(L) * 128 + uint64_t(d) * 32 + uint64_t(b)];
          atomicOr(f + (t >> 6), 1ull << (t & 63));
        }
      }
    }
  }
}

template <int W, int MODE, bool EE>
hipError_t launch_ix_kernel(const VsubParams& p, uint64_t gx, uint64_t gy, size_t lds, hipStream_t st) {
  static std::atomic<bool> attr_set{false};
  if (!attr_set.load(std::memory_order_acquire)) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_vsub_ix<W, MODE, EE>),
        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    if (e != hipSuccess) return e;
    attr_set.store(true, std::memory_order_release);
  }
  hipLaunchKernelGGL((k_vsub_ix<W, MODE, EE>), dim3(unsigned(gx), unsigned(gy)), dim3(1024), lds, st, p);
  return hipGetLastError();
}

template <int W, int MODE>
hipError_t launch_ix_mode(const VsubParams& p, uint64_t gx, uint64_t gy, size_t lds, hipStream_t st) {
  return p.ix_early_exit ? launch_ix_kernel<W, MODE, true>(p, gx, gy, lds, st)
                      : launch_ix_kernel<W, MODE, false>(p, gx, gy, lds, st);
}

template <int W>
hipError_t launch_ix_w(const VsubParams& p0, hipStream_t st) {
  VsubParams p = p0;
  if (p.t_count == 0 || p.s_count == 0) return hipSuccess;
  if (p.work) {
    // Anchor buckets: one workgroup per work item.
    if (p.work_count == 0) return hipSuccess;
    if (p.work_count > 0x7fffffffull || !p.t_perm || !p.s_perm || p.s_bound || p.diag != INT64_MIN ||
        p.mode == VSUB_COUNT || p.mode == VSUB_SCOUNT)
      return hipErrorInvalidValue;
    const size_t lds = size_t(p.n + 1) * 512 + (p.proper_by_size ? size_t(kIxTile) * 2 : 0);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    switch (p.mode) {
      case VSUB_ANY: return launch_ix_mode<W, VSUB_ANY>(p, p.work_count, 1, lds, st);
      case VSUB_FLAGS: return launch_ix_mode<W, VSUB_FLAGS>(p, p.work_count, 1, lds, st);
      case VSUB_SFLAGS: return launch_ix_mode<W, VSUB_SFLAGS>(p, p.work_count, 1, lds, st);
    }
    return hipErrorInvalidValue;
  }
  const uint64_t gx = (p.t_count + kIxTile - 1) / kIxTile;
  if (p.s_chunk == 0) {
    uint64_t target = p.target_wgs ? p.target_wgs : 1024;
    if (p.ix_early_exit && target < SW_IX_EE_WGS) target = SW_IX_EE_WGS;
    const uint64_t gyt = (target + gx - 1) / gx;
    uint64_t chunk = (p.s_count + gyt - 1) / gyt;
    if (chunk < 128) chunk = 128;
    p.s_chunk = chunk;
  }
  if (p.ix_early_exit && p.s_bound && p0.s_chunk == 0 && p.s_chunk > SW_IX_EE_CHUNK) {
    // Sorted lists: chunks of the stream small enough that the workgroups
    // running at once (consecutive tiles, one chunk) share its index in L2;
    // smaller ones when the workgroup target asks for them.
    p.s_chunk = SW_IX_EE_CHUNK;
  }
  p.s_chunk = (p.s_chunk + 1) & ~1ull;  // even: pairs never straddle two workgroups
  const uint64_t gy = (p.s_count + p.s_chunk - 1) / p.s_chunk;
  if (gx > 0x7fffffffull || gy > 65535ull) return hipErrorInvalidValue;
  const size_t lds = size_t(p.n + 1) * 512 + (p.proper_by_size ? size_t(kIxTile) * 2 : 0);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  if (p.s_bound && (!p.s_perm || !p.t_perm || !p.t_bound || p.diag != INT64_MIN))
    return hipErrorInvalidValue;
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

hipError_t launch_ix_pair_count(int W, const uint64_t* sets, const uint32_t* perm, uint64_t count, uint32_t n,
    int complement, uint32_t* size, uint32_t* entries, hipStream_t st) {
  if (!count) return hipSuccess;
  const uint64_t cmask = complement ? ~0ull : 0ull;
  const unsigned blocks = unsigned(((count + 1) / 2 + 255) / 256);
  switch (W) {
#define C(w) case w: hipLaunchKernelGGL(k_ix_pair_count<w>, dim3(blocks), dim3(256), 0, st, sets, perm, count, n, cmask, size, entries); break;
    C(1) C(2) C(3) C(4) C(5)
#undef C
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

int ix_head_groups() { return kIxHead; }

hipError_t launch_ix_pair_fill(int W, const uint64_t* sets, const uint32_t* perm, uint64_t count, uint32_t n,
    int complement, const uint32_t* start, const uint32_t* rank, const uint32_t* order, uint32_t* list,
    uint32_t* head, hipStream_t st) {
  if (!count) return hipSuccess;
  if ((rank == nullptr) != (order == nullptr)) return hipErrorInvalidValue;
  const uint64_t cmask = complement ? ~0ull : 0ull;
  const unsigned blocks = unsigned(((count + 1) / 2 * 2 + 255) / 256);
  switch (W) {
#define C(w)                                                                                                   \
  case w:                                                                                                      \
    if (rank)                                                                                                  \
      hipLaunchKernelGGL((k_ix_pair_fill<w, true>), dim3(blocks), dim3(256), 0, st, sets, perm, count, n,     \
          cmask, start, rank, order, list, head);                                                              \
    else                                                                                                       \
      hipLaunchKernelGGL((k_ix_pair_fill<w, false>), dim3(blocks), dim3(256), 0, st, sets, perm, count, n,    \
          cmask, start, rank, order, list, head);                                                              \
    break;
    C(1) C(2) C(3) C(4) C(5)
#undef C
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_state_freq(int W, const uint64_t* sets, uint64_t count, uint32_t n, int complement, uint32_t* freq,
    hipStream_t st) {
  if (!count) return hipSuccess;
  const uint64_t cmask = complement ? ~0ull : 0ull;
  uint64_t blocks = (count + 255) / 256;  // four waves of 64 sets each
  if (blocks > 2048) blocks = 2048;
  switch (W) {
#define C(w) case w: hipLaunchKernelGGL(k_state_freq<w>, dim3(unsigned(blocks)), dim3(256), 0, st, sets, count, n, cmask, freq); break;
    C(1) C(2) C(3) C(4) C(5)
#undef C
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_size_hist(int W, const uint64_t* sets, uint64_t count, uint32_t n, int complement, uint32_t* hist,
    hipStream_t st) {
  if (!count) return hipSuccess;
  const uint64_t cmask = complement ? ~0ull : 0ull;
  uint64_t blocks = (count + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  switch (W) {
#define C(w) case w: hipLaunchKernelGGL(k_size_hist<w>, dim3(unsigned(blocks)), dim3(256), 0, st, sets, count, n, cmask, hist); break;
    C(1) C(2) C(3) C(4) C(5)
#undef C
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_size_offsets(const uint32_t* hist, uint32_t n, uint32_t* cursor, uint32_t* bound, hipStream_t st) {
  if (n >= 1024) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_size_offsets, dim3(1), dim3(1024), 0, st, hist, n, cursor, bound);
  return hipGetLastError();
}

hipError_t launch_size_scatter(int W, const uint64_t* sets, uint64_t count, uint32_t n, int complement,
    uint32_t* cursor, uint32_t* perm, hipStream_t st) {
  if (!count) return hipSuccess;
  if (count > 0xffffffffull) return hipErrorInvalidValue;
  const uint64_t cmask = complement ? ~0ull : 0ull;
  const unsigned blocks = unsigned((count + 256 * kScatterPer - 1) / (256 * kScatterPer));
  switch (W) {
#define C(w) case w: hipLaunchKernelGGL(k_size_scatter<w>, dim3(blocks), dim3(256), 0, st, sets, count, n, cmask, cursor, perm); break;
    C(1) C(2) C(3) C(4) C(5)
#undef C
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_anchor(int W, const uint64_t* sets, uint64_t count, uint32_t n, int complement,
    const uint32_t* rank, uint32_t* anchor, uint32_t* hist, hipStream_t st) {
  if (!count) return hipSuccess;
  const uint64_t cmask = complement ? ~0ull : 0ull;
  const unsigned blocks = unsigned((count + 255) / 256);
  switch (W) {
#define C(w) case w: hipLaunchKernelGGL(k_anchor<w>, dim3(blocks), dim3(256), 0, st, sets, count, n, cmask, rank, anchor, hist); break;
    C(1) C(2) C(3) C(4) C(5)
#undef C
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_key_scatter(const uint32_t* key, uint64_t count, uint32_t nkeys, uint32_t* cursor, uint32_t* perm,
    hipStream_t st) {
  if (!count) return hipSuccess;
  if (nkeys > 513 || count > 0xffffffffull) return hipErrorInvalidValue;
  const unsigned blocks = unsigned((count + 256 * kScatterPer - 1) / (256 * kScatterPer));
  hipLaunchKernelGGL(k_key_scatter, dim3(blocks), dim3(256), 0, st, key, count, nkeys, cursor, perm);
  return hipGetLastError();
}

hipError_t launch_invert(int W, const uint64_t* sets, uint64_t count, uint32_t n, int complement, uint32_t* cursor,
    uint32_t* lists, hipStream_t st) {
  if (!count) return hipSuccess;
  if (count > 0xffffffffull) return hipErrorInvalidValue;
  const uint64_t cmask = complement ? ~0ull : 0ull;
  const unsigned blocks = unsigned((count + 255) / 256);
  switch (W) {
#define C(w) case w: hipLaunchKernelGGL(k_invert<w>, dim3(blocks), dim3(256), 0, st, sets, count, n, cmask, cursor, lists); break;
    C(1) C(2) C(3) C(4) C(5)
#undef C
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_state_rank(const uint32_t* freq, uint32_t n, uint32_t* rank, uint32_t* order, hipStream_t st) {
  if (n == 0 || n > 512) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_state_rank, dim3(1), dim3(512), 0, st, freq, n, rank, order);
  return hipGetLastError();
}

hipError_t launch_vsub_ix(int W, const VsubParams& p, hipStream_t st) {
  if (!p.ix_start || !p.ix_list) return hipErrorInvalidValue;
  if (p.mode == VSUB_ANY && (!p.found || !p.witness)) return hipErrorInvalidValue;
  if (p.ix_early_exit && !p.ix_head) return hipErrorInvalidValue;
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
