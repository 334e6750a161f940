// Device subset trie (the reference's SubsetsTrie<N,Threads>,
// synchrolib/data_structures/subsets_trie.hpp:18-468): a binary trie over a
// list of state sets, each internal node splitting its sets on a division
// bit — the state held by the most sets of the node that is not held by all
// of them, lowest state on ties (subsets_trie.hpp:299-350) — with each
// node's smallest and largest cardinality for pruning (the reference keeps
// the smallest, :22-38,355-363).  Leaves hold at most kTrieLeafSets sets, or
// any number at depth kTrieMaxDepth.
//
// Build, level-synchronous while a node holds more sets than one
// workgroup's LDS takes (host-driven, trie.cpp): per 2048-set chunk a state
// histogram in LDS added into the node's (k_trie_hist), the division bit
// per node (k_trie_choose), each chunk's sets without the bit
// (k_trie_chunk_zeros), then a stable scatter of every chunk into the next
// buffer, sets without the bit first (k_trie_scatter).  Nodes small enough
// are built whole by one workgroup each in LDS (k_trie_local): histogram,
// bit, stable partition, children, depth first with an explicit stack.
// Every partition is stable and node ids come from fixed per-node ranges, so
// the trie is the same on every run.
//
// Query (k_trie_query): persistent workgroups pull batches of 64 queries;
// each lane walks the trie for its query with a stack of pending nodes in
// LDS (at most one per level of the current path).  Two kinds:
//   SUBSET   is some trie set inside the query (the reference's
//            contains_subset_of, :353-468): the zero child always, the one
//            child only when the query holds the bit; prune a node whose
//            smallest set is larger than the query;
//   SUPERSET is some trie set containing the query: the one child always,
//            the zero child only when the query lacks the bit; prune a node
//            whose largest set is smaller than the query.
// Queries may be ordered by cardinality first (k_card_*): the lanes of a
// wave then walk alike, and once a batch is pruned at the root every later
// one is too.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "sw_device.hpp"
#include "sw_kernels.hpp"

namespace sw {

namespace {

template <int W>
__device__ __forceinline__ void trie_load(const uint64_t* __restrict__ src, uint64_t i, uint64_t cmask, uint32_t n,
    uint64_t (&x)[W]) {
#pragma unroll
  for (int w = 0; w < W; ++w) x[w] = (src[i * W + w] ^ cmask) & word_mask(n, w);
}

template <int W>
__device__ __forceinline__ uint32_t popc_set(const uint64_t (&x)[W]) {
  uint32_t c = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) c += uint32_t(__popcll(x[w]));
  return c;
}

// Largest (count, -state) over states with 0 < count < size, by one wave:
// (division bit, count) or (kTrieLeafBit, 0) when every state is constant.
__device__ __forceinline__ uint64_t wave_best_bit(const uint32_t* h, uint32_t n, uint32_t size, uint32_t lane) {
  uint64_t best = 0;
  for (uint32_t q = lane; q < n; q += 64) {
    const uint32_t c = h[q];
    if (c > 0 && c < size) {
      const uint64_t key = (uint64_t(c) << 32) | uint64_t(0xffffffffu - q);
      best = key > best ? key : best;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const uint64_t o = __shfl_xor(best, off, 64);
    best = o > best ? o : best;
  }
  return best;
}

}  // namespace

// ------------------------------------------------------------------ build

template <int W>
__global__ __launch_bounds__(256) void k_trie_hist(const uint64_t* __restrict__ src, uint64_t cmask, uint32_t n,
    const TrieChunk* __restrict__ chunks, uint32_t* __restrict__ hist, uint32_t* __restrict__ seg_min,
    uint32_t* __restrict__ seg_max) {
  __shared__ uint32_t h[kMaxW * 64];
  __shared__ uint32_t smin, smax;
  const TrieChunk c = chunks[blockIdx.x];
  for (uint32_t t = threadIdx.x; t < W * 64; t += blockDim.x) h[t] = 0;
  if (threadIdx.x == 0) {
    smin = 0xffffffffu;
    smax = 0;
  }
  __syncthreads();
  // 64-set groups per wave, counted by 64 x 64 transposes (lane j of the
  // transposed word w holds the group's column of state 64 w + j).
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t mn = 0xffffffffu, mx = 0;
  uint32_t part[W];
#pragma unroll
  for (int w = 0; w < W; ++w) part[w] = 0;
  for (uint32_t g = c.begin + wv * 64; g < c.end; g += blockDim.x) {
    const uint32_t i = g + lane;
    const bool ok = i < c.end;
    uint64_t x[W];
#pragma unroll
    for (int w = 0; w < W; ++w) x[w] = 0;
    if (ok) trie_load<W>(src, i, cmask, n, x);
    uint32_t p = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      p += uint32_t(__popcll(x[w]));
      part[w] += uint32_t(__popcll(lane_transpose64(x[w])));
    }
    if (ok) {
      mn = p < mn ? p : mn;
      mx = p > mx ? p : mx;
    }
  }
#pragma unroll
  for (int w = 0; w < W; ++w)
    if (part[w]) atomicAdd(&h[w * 64 + lane], part[w]);
  atomicMin(&smin, mn);
  atomicMax(&smax, mx);
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < n; t += blockDim.x)
    if (h[t]) atomicAdd(&hist[uint64_t(c.seg) * n + t], h[t]);
  if (threadIdx.x == 0) {
    atomicMin(&seg_min[c.seg], smin);
    atomicMax(&seg_max[c.seg], smax);
  }
}

__global__ __launch_bounds__(64) void k_trie_choose(const uint32_t* __restrict__ hist,
    const TrieSeg* __restrict__ segs, uint32_t n, uint32_t* __restrict__ bit_out, uint32_t* __restrict__ ones_out) {
  const TrieSeg s = segs[blockIdx.x];
  const uint32_t size = s.end - s.begin;
  const uint64_t best = wave_best_bit(hist + uint64_t(blockIdx.x) * n, n, size, threadIdx.x);
  if (threadIdx.x == 0) {
    bit_out[blockIdx.x] = best ? 0xffffffffu - uint32_t(best) : kTrieLeafBit;
    ones_out[blockIdx.x] = uint32_t(best >> 32);
  }
}

template <int W>
__global__ __launch_bounds__(256) void k_trie_chunk_zeros(const uint64_t* __restrict__ src, uint64_t cmask,
    const TrieChunk* __restrict__ chunks, const uint32_t* __restrict__ bits, uint32_t* __restrict__ czero) {
  __shared__ uint32_t tot;
  const TrieChunk c = chunks[blockIdx.x];
  const uint32_t b = bits[c.seg];
  if (threadIdx.x == 0) tot = 0;
  __syncthreads();
  uint32_t z = 0;
  if (b != kTrieLeafBit) {
    const uint32_t bw = b >> 6;
    const uint64_t bm = 1ull << (b & 63);
    for (uint32_t i = c.begin + threadIdx.x; i < c.end; i += blockDim.x)
      z += ((src[uint64_t(i) * W + bw] ^ cmask) & bm) ? 0u : 1u;
  }
  atomicAdd(&tot, z);
  __syncthreads();
  if (threadIdx.x == 0) czero[blockIdx.x] = tot;
}

// Stable partition of every chunk: sets without the division bit go to
// coff[c].x onwards, sets with it to coff[c].y onwards (absolute rows of
// dst), in list order.  256 threads.
template <int W>
__global__ __launch_bounds__(256) void k_trie_scatter(const uint64_t* __restrict__ src, uint64_t cmask, uint32_t n,
    const uint32_t* __restrict__ src_tags, uint64_t* __restrict__ dst, uint32_t* __restrict__ dst_tags,
    const TrieChunk* __restrict__ chunks, const uint32_t* __restrict__ bits, const uint2* __restrict__ coff) {
  __shared__ uint32_t wz[4], wo[4];
  const TrieChunk c = chunks[blockIdx.x];
  const uint32_t b = bits[c.seg];
  if (b == kTrieLeafBit) return;  // the whole workgroup leaves together
  const uint32_t bw = b >> 6;
  const uint64_t bm = 1ull << (b & 63);
  uint32_t zb = coff[blockIdx.x].x, ob = coff[blockIdx.x].y;
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  for (uint32_t i0 = c.begin; i0 < c.end; i0 += 256) {
    const uint32_t i = i0 + threadIdx.x;
    const bool ok = i < c.end;
    uint64_t x[W];
    bool one = false;
    if (ok) {
      trie_load<W>(src, i, cmask, n, x);
      one = (x[bw] & bm) != 0;
    }
    const uint64_t bz = __ballot(ok && !one), bo = __ballot(ok && one);
    if (lane == 0) {
      wz[wv] = uint32_t(__popcll(bz));
      wo[wv] = uint32_t(__popcll(bo));
    }
    __syncthreads();
    uint32_t pz = 0, po = 0, tz = 0, to = 0;
#pragma unroll
    for (uint32_t w = 0; w < 4; ++w) {
      if (w < wv) {
        pz += wz[w];
        po += wo[w];
      }
      tz += wz[w];
      to += wo[w];
    }
    if (ok) {
      const uint32_t d = one ? ob + po + uint32_t(__popcll(bo & below)) : zb + pz + uint32_t(__popcll(bz & below));
#pragma unroll
      for (int w = 0; w < W; ++w) dst[uint64_t(d) * W + w] = x[w];
      dst_tags[d] = src_tags ? src_tags[i] : i;
    }
    zb += tz;
    ob += to;
    __syncthreads();
  }
}

// A node of more sets than LDS takes that cannot split (depth cap, or equal
// sets): copied to the output as one leaf by one workgroup.
template <int W>
__global__ __launch_bounds__(256) void k_trie_bigleaf(TrieBuildArgs a, const TrieLocalJob* __restrict__ jobs) {
  __shared__ uint32_t smin, smax;
  const TrieLocalJob job = jobs[blockIdx.x];
  const uint64_t* src = job.buf == 0 ? a.in : (job.buf == 1 ? a.buf1 : a.buf2);
  const uint32_t* stags = job.buf == 0 ? a.in_tags : (job.buf == 1 ? a.tags1 : a.tags2);
  const uint64_t cmask = job.buf == 0 ? a.in_cmask : 0ull;
  if (threadIdx.x == 0) {
    smin = 0xffffffffu;
    smax = 0;
  }
  __syncthreads();
  uint32_t mn = 0xffffffffu, mx = 0;
  for (uint32_t i = job.begin + threadIdx.x; i < job.end; i += blockDim.x) {
    uint64_t x[W];
    trie_load<W>(src, i, cmask, a.n, x);
    const uint32_t p = popc_set<W>(x);
    mn = p < mn ? p : mn;
    mx = p > mx ? p : mx;
#pragma unroll
    for (int w = 0; w < W; ++w) a.out[uint64_t(i) * W + w] = x[w];
    a.out_tags[i] = stags ? stags[i] : i;
  }
  atomicMin(&smin, mn);
  atomicMax(&smax, mx);
  __syncthreads();
  if (threadIdx.x == 0) a.nodes[job.root] = make_uint4(job.begin, job.end - job.begin, kTrieLeafBit | (smin << 16), smax);
}

// Whole subtrees of at most trie_local_cap(W) sets, one per workgroup (256
// threads), in LDS: two set buffers and two tag buffers (dynamic LDS).  Nodes
// of more than 64 sets are split by the whole workgroup, one at a time
// (state histogram, division bit, stable partition); a node of at most 64
// sets goes to a queue, and then the four waves build those subtrees
// side by side, one lane per set, with no workgroup barrier.  Node ids: the
// job's root, then job.desc_base onwards in the order the workgroup meets
// nodes: 2 per split of the workgroup, and 2 (s - 1) reserved for a queued
// subtree of s sets (at most s - 1 splits below its root).  In all at most
// 2 (size - 1): the workgroup's splits number one less than the queued
// subtrees and its own leaves, and each of those leaves holds a set.
struct LocalFrame {
  uint32_t node, b, e, depth_buf;  // depth << 1 | buffer holding [b, e)
};
struct SmallFrame {
  uint32_t node, b, e, depth_buf, id_base;
};
constexpr uint32_t kTrieSmallQueue = 256;

// One wave builds the subtree of at most 64 sets of frame f: node records
// and the leaves' sets and tags written out; stack = this wave's frames.
template <int W>
__device__ void wave_subtree(const TrieBuildArgs& a, uint32_t job_begin, SmallFrame sf, uint64_t* S0, uint64_t* S1,
    uint32_t* T0, uint32_t* T1, LocalFrame* stack, uint32_t lane) {
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  uint32_t sp = 1, next = sf.id_base;
  stack[0] = LocalFrame{sf.node, sf.b, sf.e, sf.depth_buf};
  while (sp) {
    // Every lane writes and reads the same frames: each sees its own writes.
    const LocalFrame f = stack[--sp];
    const uint32_t depth = f.depth_buf >> 1, which = f.depth_buf & 1;
    const uint64_t* A = which ? S1 : S0;
    uint64_t* B = which ? S0 : S1;
    const uint32_t* TA = which ? T1 : T0;
    uint32_t* TB = which ? T0 : T1;
    const uint32_t cnt = f.e - f.b;
    const bool ok = lane < cnt;
    uint64_t x[W];
    uint32_t p = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      x[w] = ok ? A[size_t(f.b + lane) * W + w] : 0ull;
      p += uint32_t(__popcll(x[w]));
    }
    const uint32_t tag = ok ? TA[f.b + lane] : 0u;
    uint32_t mn = ok ? p : 0xffffffffu, mx = ok ? p : 0u;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      mn = min(mn, uint32_t(__shfl_xor(mn, off, 64)));
      mx = max(mx, uint32_t(__shfl_xor(mx, off, 64)));
    }
    uint64_t best = 0;
    if (cnt > kTrieLeafSets && depth < kTrieMaxDepth) {
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const uint32_t c = uint32_t(__popcll(lane_transpose64(x[w])));  // sets holding state 64 w + lane
        const uint32_t q = uint32_t(w) * 64 + lane;
        if (c > 0 && c < cnt && q < a.n) {
          const uint64_t key = (uint64_t(c) << 32) | uint64_t(0xffffffffu - q);
          best = key > best ? key : best;
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const uint64_t o = __shfl_xor(best, off, 64);
        best = o > best ? o : best;
      }
    }
    if (!best) {
      if (lane == 0) a.nodes[f.node] = make_uint4(job_begin + f.b, cnt, kTrieLeafBit | (mn << 16), mx);
      if (ok) {
#pragma unroll
        for (int w = 0; w < W; ++w) a.out[uint64_t(job_begin + f.b + lane) * W + w] = x[w];
        a.out_tags[job_begin + f.b + lane] = tag;
      }
      continue;
    }
    const uint32_t bit = 0xffffffffu - uint32_t(best);
    const bool one = ok && ((x[bit >> 6] >> (bit & 63)) & 1ull);
    const uint64_t bz = __ballot(ok && !one), bo = __ballot(one);
    const uint32_t zc = uint32_t(__popcll(bz));
    if (ok) {
      const uint32_t d = one ? f.b + zc + uint32_t(__popcll(bo & below)) : f.b + uint32_t(__popcll(bz & below));
#pragma unroll
      for (int w = 0; w < W; ++w) B[size_t(d) * W + w] = x[w];
      TB[d] = tag;
    }
    // The children's reads of B follow these writes in the wave's LDS order.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint32_t z = next, o = next + 1;
    next += 2;
    if (lane == 0) a.nodes[f.node] = make_uint4(z, o, bit | (mn << 16), mx);
    const uint32_t db = ((depth + 1) << 1) | (which ^ 1u);
    stack[sp++] = LocalFrame{o, f.b + zc, f.e, db};
    stack[sp++] = LocalFrame{z, f.b, f.b + zc, db};
  }
}

template <int W>
__global__ __launch_bounds__(256) void k_trie_local(TrieBuildArgs a, const TrieLocalJob* __restrict__ jobs) {
  constexpr uint32_t CAP = trie_local_cap(W);
  extern __shared__ uint64_t lds[];
  uint64_t* S0 = lds;
  uint64_t* S1 = lds + size_t(CAP) * W;
  uint32_t* T0 = reinterpret_cast<uint32_t*>(lds + 2 * size_t(CAP) * W);
  uint32_t* T1 = T0 + CAP;
  __shared__ uint32_t h[kMaxW * 64];
  __shared__ LocalFrame stack[kTrieMaxDepth + 2];
  __shared__ LocalFrame wstack[4][kTrieMaxDepth + 2];
  __shared__ SmallFrame queue[kTrieSmallQueue];
  __shared__ LocalFrame cur;
  __shared__ uint32_t sp, qn, qhead, next_id, smin, smax, sbit, sones;
  __shared__ uint32_t wz[4], wo[4];
  const TrieLocalJob job = jobs[blockIdx.x];
  const uint32_t size = job.end - job.begin;
  const uint64_t* src = job.buf == 0 ? a.in : (job.buf == 1 ? a.buf1 : a.buf2);
  const uint32_t* stags = job.buf == 0 ? a.in_tags : (job.buf == 1 ? a.tags1 : a.tags2);
  const uint64_t cmask = job.buf == 0 ? a.in_cmask : 0ull;
  const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  for (uint32_t i = tid; i < size; i += blockDim.x) {
    uint64_t x[W];
    trie_load<W>(src, job.begin + i, cmask, a.n, x);
#pragma unroll
    for (int w = 0; w < W; ++w) S0[size_t(i) * W + w] = x[w];
    T0[i] = stags ? stags[job.begin + i] : job.begin + i;
  }
  if (tid == 0) {
    next_id = job.desc_base;
    qn = 0;
    qhead = 0;
    if (size <= 64) {  // the whole job is one small subtree
      queue[0] = SmallFrame{job.root, 0, size, job.depth << 1, job.desc_base};
      qn = 1;
      sp = 0;
    } else {
      sp = 1;
      stack[0] = LocalFrame{job.root, 0, size, job.depth << 1};
    }
  }
  // Block phase: nodes of more than 64 sets (or small ones the queue had no room for).
  while (true) {
    __syncthreads();
    if (tid == 0) {
      cur = sp ? stack[--sp] : LocalFrame{0xffffffffu, 0, 0, 0};
      smin = 0xffffffffu;
      smax = 0;
    }
    __syncthreads();
    const LocalFrame f = cur;
    if (f.node == 0xffffffffu) break;
    const uint32_t depth = f.depth_buf >> 1;
    const uint32_t which = f.depth_buf & 1;
    uint64_t* A = which ? S1 : S0;
    uint64_t* B = which ? S0 : S1;
    uint32_t* TA = which ? T1 : T0;
    uint32_t* TB = which ? T0 : T1;
    const uint32_t cnt = f.e - f.b;
    const bool split = cnt > kTrieLeafSets && depth < kTrieMaxDepth;
    if (split)
      for (uint32_t t = tid; t < W * 64; t += blockDim.x) h[t] = 0;
    __syncthreads();
    // State counts by 64-set groups, one per wave at a time: lane l takes
    // set l of the group; a 64 x 64 transpose of each word puts the
    // group's column of state 64 w + j in lane j, whose popcount is that
    // state's count.  Each wave adds its counts to LDS once (no two lanes
    // on one address), instead of one contended atomic per set bit.
    uint32_t mn = 0xffffffffu, mx = 0;
    uint32_t part[W];
#pragma unroll
    for (int w = 0; w < W; ++w) part[w] = 0;
    for (uint32_t g = f.b + wv * 64; g < f.e; g += 256) {
      const uint32_t i = g + lane;
      const bool ok = i < f.e;
      uint32_t p = 0;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const uint64_t v = ok ? A[size_t(i) * W + w] : 0ull;
        p += uint32_t(__popcll(v));
        if (split) part[w] += uint32_t(__popcll(lane_transpose64(v)));
      }
      if (ok) {
        mn = p < mn ? p : mn;
        mx = p > mx ? p : mx;
      }
    }
    if (split)
#pragma unroll
      for (int w = 0; w < W; ++w)
        if (part[w]) atomicAdd(&h[w * 64 + lane], part[w]);
    atomicMin(&smin, mn);
    atomicMax(&smax, mx);
    __syncthreads();
    if (wv == 0) {
      const uint64_t best = split ? wave_best_bit(h, a.n, cnt, lane) : 0ull;
      if (lane == 0) {
        sbit = best ? 0xffffffffu - uint32_t(best) : kTrieLeafBit;
        sones = uint32_t(best >> 32);
      }
    }
    __syncthreads();
    const uint32_t bit = sbit;
    if (bit == kTrieLeafBit) {
      if (tid == 0) a.nodes[f.node] = make_uint4(job.begin + f.b, cnt, kTrieLeafBit | (smin << 16), smax);
      for (uint32_t i = f.b + tid; i < f.e; i += blockDim.x) {
#pragma unroll
        for (int w = 0; w < W; ++w) a.out[uint64_t(job.begin + i) * W + w] = A[size_t(i) * W + w];
        a.out_tags[job.begin + i] = TA[i];
      }
      continue;
    }
    const uint32_t zc = cnt - sones;
    const uint32_t bw = bit >> 6;
    const uint64_t bm = 1ull << (bit & 63);
    uint32_t zb = f.b, ob = f.b + zc;
    for (uint32_t i0 = f.b; i0 < f.e; i0 += 256) {
      const uint32_t i = i0 + tid;
      const bool ok = i < f.e;
      const bool one = ok && (A[size_t(i) * W + bw] & bm) != 0;
      const uint64_t bz = __ballot(ok && !one), bo = __ballot(one);
      if (lane == 0) {
        wz[wv] = uint32_t(__popcll(bz));
        wo[wv] = uint32_t(__popcll(bo));
      }
      __syncthreads();
      uint32_t pz = 0, po = 0, tz = 0, to = 0;
#pragma unroll
      for (uint32_t w = 0; w < 4; ++w) {
        if (w < wv) {
          pz += wz[w];
          po += wo[w];
        }
        tz += wz[w];
        to += wo[w];
      }
      if (ok) {
        const uint32_t d = one ? ob + po + uint32_t(__popcll(bo & below)) : zb + pz + uint32_t(__popcll(bz & below));
#pragma unroll
        for (int w = 0; w < W; ++w) B[size_t(d) * W + w] = A[size_t(i) * W + w];
        TB[d] = TA[i];
      }
      zb += tz;
      ob += to;
      __syncthreads();
    }
    if (tid == 0) {
      const uint32_t z = next_id, o = next_id + 1;
      next_id += 2;
      a.nodes[f.node] = make_uint4(z, o, bit | (smin << 16), smax);
      const uint32_t db = ((depth + 1) << 1) | (which ^ 1u);
      const LocalFrame kids[2] = {LocalFrame{o, f.b + zc, f.e, db}, LocalFrame{z, f.b, f.b + zc, db}};
      for (const LocalFrame& k : kids) {
        const uint32_t ks = k.e - k.b;
        if (ks <= 64 && qn < kTrieSmallQueue) {
          queue[qn++] = SmallFrame{k.node, k.b, k.e, k.depth_buf, next_id};
          next_id += 2 * (ks - 1);
        } else {
          stack[sp++] = k;
        }
      }
    }
  }
  // Wave phase: the queued subtrees, one wave each at a time.
  while (true) {
    uint32_t qi = 0;
    if (lane == 0) qi = atomicAdd(&qhead, 1u);
    qi = __shfl(qi, 0, 64);
    if (qi >= qn) break;
    wave_subtree<W>(a, job.begin, queue[qi], S0, S1, T0, T1, wstack[wv], lane);
  }
}

// ------------------------------------------------------------------ query

// Persistent: each wave takes 64 queries at a time from *next.  Lane stacks
// in LDS: stack[wave][level][lane], at most kTrieMaxDepth entries (a pending
// node is a sibling of a node on the current path, one per level).
template <int W, int SUPER, int ANY>
__global__ __launch_bounds__(256) void k_trie_query(TrieQueryArgs a) {
  __shared__ uint32_t stack[4][kTrieMaxDepth + 1][64];
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint4 root = a.nodes[0];
  const uint32_t root_min = root.z >> 16, root_max = root.w;
  while (true) {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(a.next, 64u);
    base = __shfl(base, 0, 64);
    if (base >= a.qcount) break;
    if (ANY) {
      int f = 0;
      if (lane == 0) f = __hip_atomic_load(a.found, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (__shfl(f, 0, 64)) break;
    }
    const uint32_t qi = base + lane;
    bool active = qi < a.qcount;
    uint32_t idx = 0, pop = 0;
    uint64_t x[W];
#pragma unroll
    for (int w = 0; w < W; ++w) x[w] = 0;
    if (active) {
      const uint32_t p = a.reverse ? a.qcount - 1 - qi : qi;
      idx = a.perm ? a.perm[p] : p;
      trie_load<W>(a.queries, idx, a.q_cmask, a.n, x);
      pop = popc_set<W>(x);
      active = SUPER ? (pop <= root_max) : (pop >= root_min);
    }
    // In cardinality order a batch pruned at the root means every later one is.
    if (a.sorted && __ballot(active) == 0) break;
    // A walk visits each node at most once: the visit count bounds it even
    // if a node record were wrong.
    uint32_t sp = 0, v = 0, iter = 0, visits = 0;
    while (__ballot(active)) {
      if (ANY && ((++iter & 15u) == 0)) {
        int f = 0;
        if (lane == 0) f = __hip_atomic_load(a.found, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (__shfl(f, 0, 64)) {
          active = false;
          break;
        }
      }
      if (!active) continue;
      if (v >= a.node_count || ++visits > a.node_count) {
        active = false;
        continue;
      }
      const uint4 nd = a.nodes[v];
      const uint32_t bit = nd.z & 0xffffu;
      bool up = SUPER ? (pop > nd.w) : (pop < (nd.z >> 16));
      if (!up) {
        if (bit == kTrieLeafBit) {
          for (uint32_t s = nd.x; s < nd.x + nd.y; ++s) {
            uint64_t miss = 0;
#pragma unroll
            for (int w = 0; w < W; ++w) {
              const uint64_t t = a.tsets[uint64_t(s) * W + w];
              miss |= SUPER ? (x[w] & ~t) : (t & ~x[w]);
            }
            if (miss == 0) {
              if (ANY) {
                const uint64_t tag = a.ttags[s];
                const uint64_t key = a.query_high ? ((uint64_t(idx) << 32) | tag) : ((tag << 32) | idx);
                atomicMin(a.witness, (unsigned long long)key);
                __hip_atomic_store(a.found, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              } else {
                a.flags[idx] = 1;
              }
              active = false;
              break;
            }
          }
          up = true;
        } else {
          const bool has = (x[bit >> 6] >> (bit & 63)) & 1ull;
          if (SUPER) {
            if (!has) stack[wv][sp++][lane] = nd.x;
            v = nd.y;
          } else {
            if (has) stack[wv][sp++][lane] = nd.y;
            v = nd.x;
          }
        }
      }
      if (active && up) {
        if (sp == 0) active = false;
        else v = stack[wv][--sp][lane];
      }
    }
  }
}

// ------------------------------------------------------ cardinality order

template <int W>
__global__ __launch_bounds__(256) void k_card_hist(const uint64_t* __restrict__ sets, uint64_t count, uint64_t cmask,
    uint32_t n, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[kMaxW * 64 + 1];
  for (uint32_t t = threadIdx.x; t <= n; t += blockDim.x) h[t] = 0;
  __syncthreads();
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += uint64_t(gridDim.x) * blockDim.x) {
    uint64_t x[W];
    trie_load<W>(sets, i, cmask, n, x);
    atomicAdd(&h[popc_set<W>(x)], 1u);
  }
  __syncthreads();
  for (uint32_t t = threadIdx.x; t <= n; t += blockDim.x)
    if (h[t]) atomicAdd(&hist[t], h[t]);
}

// A workgroup takes 8 sets per thread, counts them per cardinality in LDS
// and reserves each cardinality's range with one global atomic.
constexpr int kCardPer = 8;

template <int W>
__global__ __launch_bounds__(256) void k_card_scatter(const uint64_t* __restrict__ sets, uint64_t count,
    uint64_t cmask, uint32_t n, uint32_t* __restrict__ cursor, uint32_t* __restrict__ perm) {
  __shared__ uint32_t cnt[kMaxW * 64 + 1], base[kMaxW * 64 + 1];
  for (uint32_t t = threadIdx.x; t <= n; t += blockDim.x) cnt[t] = 0;
  __syncthreads();
  const uint64_t i0 = uint64_t(blockIdx.x) * blockDim.x * kCardPer + threadIdx.x;
  uint32_t c[kCardPer], local[kCardPer];
#pragma unroll
  for (int e = 0; e < kCardPer; ++e) {
    const uint64_t i = i0 + uint64_t(e) * blockDim.x;
    c[e] = 0;
    local[e] = 0;
    if (i < count) {
      uint64_t x[W];
      trie_load<W>(sets, i, cmask, n, x);
      c[e] = popc_set<W>(x);
      local[e] = atomicAdd(&cnt[c[e]], 1u);
    }
  }
  __syncthreads();
  for (uint32_t t = threadIdx.x; t <= n; t += blockDim.x)
    if (cnt[t]) base[t] = atomicAdd(&cursor[t], cnt[t]);
  __syncthreads();
#pragma unroll
  for (int e = 0; e < kCardPer; ++e) {
    const uint64_t i = i0 + uint64_t(e) * blockDim.x;
    if (i < count) perm[base[c[e]] + local[e]] = uint32_t(i);
  }
}

// ------------------------------------------------------------- launchers

#define SW_TRIE_W(MACRO) MACRO(1) MACRO(2) MACRO(3) MACRO(4) MACRO(5) MACRO(6) MACRO(7) MACRO(8)

hipError_t launch_trie_hist(int W, const uint64_t* src, int complement, uint32_t n, const TrieChunk* chunks,
    uint32_t nchunks, uint32_t* hist, uint32_t* seg_min, uint32_t* seg_max, hipStream_t st) {
  if (!nchunks) return hipSuccess;
  const uint64_t cm = complement ? ~0ull : 0ull;
  switch (W) {
#define C(w) case w: hipLaunchKernelGGL(k_trie_hist<w>, dim3(nchunks), dim3(256), 0, st, src, cm, n, chunks, hist, seg_min, seg_max); break;
    SW_TRIE_W(C)
#undef C
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_trie_choose(const uint32_t* hist, const TrieSeg* segs, uint32_t nsegs, uint32_t n, uint32_t* bits,
    uint32_t* ones, hipStream_t st) {
  if (!nsegs) return hipSuccess;
  hipLaunchKernelGGL(k_trie_choose, dim3(nsegs), dim3(64), 0, st, hist, segs, n, bits, ones);
  return hipGetLastError();
}

hipError_t launch_trie_chunk_zeros(int W, const uint64_t* src, int complement, const TrieChunk* chunks,
    uint32_t nchunks, const uint32_t* bits, uint32_t* czero, hipStream_t st) {
  if (!nchunks) return hipSuccess;
  const uint64_t cm = complement ? ~0ull : 0ull;
  switch (W) {
#define C(w) case w: hipLaunchKernelGGL(k_trie_chunk_zeros<w>, dim3(nchunks), dim3(256), 0, st, src, cm, chunks, bits, czero); break;
    SW_TRIE_W(C)
#undef C
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_trie_scatter(int W, const uint64_t* src, int complement, uint32_t n, const uint32_t* src_tags,
    uint64_t* dst, uint32_t* dst_tags, const TrieChunk* chunks, uint32_t nchunks, const uint32_t* bits,
    const uint2* coff, hipStream_t st) {
  if (!nchunks) return hipSuccess;
  const uint64_t cm = complement ? ~0ull : 0ull;
  switch (W) {
#define C(w) case w: hipLaunchKernelGGL(k_trie_scatter<w>, dim3(nchunks), dim3(256), 0, st, src, cm, n, src_tags, dst, dst_tags, chunks, bits, coff); break;
    SW_TRIE_W(C)
#undef C
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_trie_bigleaf(int W, const TrieBuildArgs& a, const TrieLocalJob* jobs, uint32_t njobs,
    hipStream_t st) {
  if (!njobs) return hipSuccess;
  switch (W) {
#define C(w) case w: hipLaunchKernelGGL(k_trie_bigleaf<w>, dim3(njobs), dim3(256), 0, st, a, jobs); break;
    SW_TRIE_W(C)
#undef C
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int W>
static hipError_t launch_trie_local_w(const TrieBuildArgs& a, const TrieLocalJob* jobs, uint32_t njobs,
    hipStream_t st) {
  const size_t lds = trie_local_lds(W);
  static bool attr = false;  // idempotent; a race only repeats the call
  if (!attr) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_trie_local<W>),
        hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (e != hipSuccess) return e;
    attr = true;
  }
  hipLaunchKernelGGL(k_trie_local<W>, dim3(njobs), dim3(256), lds, st, a, jobs);
  return hipGetLastError();
}

hipError_t launch_trie_local(int W, const TrieBuildArgs& a, const TrieLocalJob* jobs, uint32_t njobs,
    hipStream_t st) {
  if (!njobs) return hipSuccess;
  switch (W) {
#define C(w) case w: return launch_trie_local_w<w>(a, jobs, njobs, st);
    SW_TRIE_W(C)
#undef C
  }
  return hipErrorInvalidValue;
}

hipError_t launch_trie_query(int W, int superset, int any, const TrieQueryArgs& a, uint32_t grid, hipStream_t st) {
  if (!a.qcount || !grid) return hipSuccess;
#define C(w)                                                                                      \
  case w:                                                                                         \
    if (superset && any) hipLaunchKernelGGL((k_trie_query<w, 1, 1>), dim3(grid), dim3(256), 0, st, a); \
    else if (superset) hipLaunchKernelGGL((k_trie_query<w, 1, 0>), dim3(grid), dim3(256), 0, st, a);   \
    else if (any) hipLaunchKernelGGL((k_trie_query<w, 0, 1>), dim3(grid), dim3(256), 0, st, a);        \
    else hipLaunchKernelGGL((k_trie_query<w, 0, 0>), dim3(grid), dim3(256), 0, st, a);                 \
    break;
  switch (W) {
    SW_TRIE_W(C)
    default: return hipErrorInvalidValue;
  }
#undef C
  return hipGetLastError();
}

hipError_t launch_card_hist(int W, const uint64_t* sets, uint64_t count, int complement, uint32_t n, uint32_t* hist,
    hipStream_t st) {
  if (!count) return hipSuccess;
  const uint64_t cm = complement ? ~0ull : 0ull;
  const unsigned blocks = unsigned(std::min<uint64_t>((count + 255) / 256, 2048));
  switch (W) {
#define C(w) case w: hipLaunchKernelGGL(k_card_hist<w>, dim3(blocks), dim3(256), 0, st, sets, count, cm, n, hist); break;
    SW_TRIE_W(C)
#undef C
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_card_scatter(int W, const uint64_t* sets, uint64_t count, int complement, uint32_t n,
    uint32_t* cursor, uint32_t* perm, hipStream_t st) {
  if (!count) return hipSuccess;
  if (count > 0xffffffffull) return hipErrorInvalidValue;
  const uint64_t cm = complement ? ~0ull : 0ull;
  const unsigned blocks = unsigned((count + 256 * kCardPer - 1) / (256 * kCardPer));
  switch (W) {
#define C(w) case w: hipLaunchKernelGGL(k_card_scatter<w>, dim3(blocks), dim3(256), 0, st, sets, count, cm, n, cursor, perm); break;
    SW_TRIE_W(C)
#undef C
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace sw
