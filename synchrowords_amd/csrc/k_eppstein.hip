// Eppstein's greedy bound on the device, one workgroup per automaton (the
// reference's Eppstein, synchrolib/algorithm/eppstein/eppstein.hpp:25-114,
// over PairsTree::calculate_tree, data_structures/pairs_tree.hpp:25-65).
//
// Pair BFS.  Backwards from the pairs one letter merges, over unordered state
// pairs, recording per pair the length of a shortest merging word and its
// first letter.  The reference's queue is first-in first-out and a pair keeps
// the letter of its first discovery, so both the queue order and the letters
// depend on the order candidates are met: (parent's place in the queue,
// letter, place of each state in its preimage list).  Here a level is
// expanded in parallel, one parent per thread, in three passes:
//   1. every candidate {x, y} not reached before proposes its key, packed in
//      that order, by atomicMin on the pair;
//   2. each parent counts its candidates whose key won; a scan places them;
//   3. each parent writes its winners, in its own order, after the earlier
//      parents' winners, which is the reference's queue order.
// The first level's parents are (state, letter) in the reference's loop order.
//
// Greedy.  While the current set has two states, take the pair of the set
// with the shortest merging word (ties: lowest first state, then lowest
// second, the reference's scan order) and apply its word letter by letter.
// The bound is the sum of those lengths; past `upper` the bound is unchanged.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "sw_device.hpp"
#include "sw_kernels.hpp"

namespace sw {

namespace {

constexpr uint64_t kNoKey = ~0ull;
constexpr int kEppThreads = 1024;

__device__ inline uint32_t pair_ix(uint32_t x, uint32_t y, uint32_t n) {
  return x < y ? x * n + y : y * n + x;
}

// Exclusive block scan of one value per thread (1024 threads); returns the
// prefix and sets *total.
__device__ inline uint32_t block_scan(uint32_t v, uint32_t* wsum, uint32_t* total) {
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t x = v;
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, 64);
    if (lane >= uint32_t(off)) x += y;
  }
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  if (wave == 0) {
    uint32_t s = lane < (blockDim.x >> 6) ? wsum[lane] : 0u;
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(s, off, 64);
      if (lane >= uint32_t(off)) s += y;
    }
    if (lane < (blockDim.x >> 6)) wsum[lane] = s;  // inclusive over waves
  }
  __syncthreads();
  const uint32_t before = wave ? wsum[wave - 1] : 0u;
  *total = wsum[(blockDim.x >> 6) - 1];
  __syncthreads();  // wsum is reused by the next scan
  return before + x - v;
}

// Candidates of parent p (level 0: p = q k + a, the pairs of one preimage
// list; later: p = place in the level, the preimage pairs of queue[p] under
// every letter), in the reference's order.  f(x, y, key, letter) per candidate.
template <typename F>
__device__ inline void candidates(uint32_t level, uint32_t p, uint32_t n, uint32_t k, const uint32_t* __restrict__ beg,
    const uint16_t* __restrict__ edges, const uint32_t* __restrict__ lvl, F&& f) {
  if (level == 0) {
    const uint32_t q = p / k, a = p % k;
    const uint32_t b = beg[a * n + q], e = beg[a * n + q + 1];
    for (uint32_t i1 = b; i1 < e; ++i1)
      for (uint32_t i2 = i1 + 1; i2 < e; ++i2)
        f(uint32_t(edges[i1]), uint32_t(edges[i2]),
          (uint64_t(p) << 20) | (uint64_t(i1 - b) << 10) | uint64_t(i2 - b), a);
    return;
  }
  const uint32_t v = lvl[p];
  const uint32_t v1 = v & 0xffffu, v2 = v >> 16;
  for (uint32_t a = 0; a < k; ++a) {
    const uint32_t b1 = beg[a * n + v1], e1 = beg[a * n + v1 + 1];
    const uint32_t b2 = beg[a * n + v2], e2 = beg[a * n + v2 + 1];
    for (uint32_t i1 = b1; i1 < e1; ++i1)
      for (uint32_t i2 = b2; i2 < e2; ++i2)
        f(uint32_t(edges[i1]), uint32_t(edges[i2]),
          (uint64_t(p) << 28) | (uint64_t(a) << 20) | (uint64_t(i1 - b1) << 10) | uint64_t(i2 - b2), a);
  }
}

// Greedy merging of the closest pair, over the pair lengths in info (one
// workgroup of kEppThreads).  The current set is W 64-bit words in LDS
// (mask).  Per merge the block compacts the set's states (each thread its
// own state's place, from popcounts of the words below it) and searches the
// closest pair; then wave 0 alone applies the pair's word letter by letter:
// lane l holds states l + 64 j of the set in registers, and the image under
// a letter is one OR over the wave per word, with no barrier per letter.
// letters[x n + y] (LDS, may be NULL: then info's low byte) is the first
// letter of pair {x, y}; dl (LDS, may be NULL: then d) the transitions as
// 16-bit words.
__device__ void epp_greedy(const EppJob& job, uint32_t n, uint32_t k, const uint32_t* __restrict__ d,
    const uint32_t* __restrict__ info, uint32_t* __restrict__ words, EppOut* __restrict__ out, uint16_t* cs,
    uint64_t* mask, uint64_t* s_best, uint32_t* s_stop, const unsigned char* letters, const uint16_t* dl) {
  const uint32_t tid = threadIdx.x, nt = blockDim.x, lane = tid & 63u;
  const uint32_t W = (n + 63) / 64;
  if (tid < W) mask[tid] = word_mask(n, int(tid));
  if (tid == 0) *s_stop = 0;
  __syncthreads();
  uint64_t total = 0, wl = 0;
  uint32_t* word = words + job.word_off;
  while (true) {
    // The set's states, ascending, in cs[0 .. m).
    uint32_t m = 0;
    for (uint32_t w = 0; w < W; ++w) m += uint32_t(__popcll(mask[w]));
    if (tid < n) {
      const uint32_t w = tid >> 6, b = tid & 63u;
      const uint64_t x = mask[w];
      if ((x >> b) & 1ull) {
        uint32_t pos = uint32_t(__popcll(x & ((1ull << b) - 1ull)));
        for (uint32_t v = 0; v < w; ++v) pos += uint32_t(__popcll(mask[v]));
        cs[pos] = uint16_t(tid);
      }
    }
    __syncthreads();
    if (m <= 1) break;
    uint64_t best = ~0ull;  // (length, first, second)
    for (uint32_t t = tid; t < m * m; t += nt) {
      const uint32_t i = t / m, j = t % m;
      if (i < j) {
        const uint32_t a = cs[i], b = cs[j];
        const uint64_t key = (uint64_t(info[a * n + b] >> 8) << 32) | (uint64_t(a) << 16) | b;
        best = key < best ? key : best;
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const uint64_t o = __shfl_xor(best, off, 64);
      best = o < best ? o : best;
    }
    if (lane == 0) s_best[tid >> 6] = best;
    __syncthreads();
    if (tid < 64) {
      best = s_best[0];
      for (uint32_t w = 1; w < (nt >> 6); ++w) best = s_best[w] < best ? s_best[w] : best;
      total += best >> 32;
      if (total > job.upper) {
        if (tid == 0) {
          *out = EppOut{0, 0, job.upper, 0};
          *s_stop = 1;
        }
      } else {
        uint64_t S[8];
#pragma unroll
        for (uint32_t w = 0; w < 8; ++w) S[w] = w < W ? mask[w] : 0ull;
        uint32_t n1 = uint32_t(best >> 16) & 0xffffu, n2 = uint32_t(best) & 0xffffu;
        while (n1 != n2) {
          const uint32_t ix = pair_ix(n1, n2, n);
          const uint32_t x = letters ? uint32_t(letters[ix]) : (info[ix] & 255u);
          if (tid == 0 && wl < job.word_cap) word[wl] = x;
          ++wl;
          uint64_t acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
          for (uint32_t j = 0; j < 8; ++j) {
            const uint32_t q = lane + 64 * j;
            if (j < W && ((S[j] >> lane) & 1ull)) {
              const uint32_t t = dl ? uint32_t(dl[q * k + x]) : d[size_t(q) * k + x];
              const uint32_t tw = t >> 6;
              const uint64_t bit = 1ull << (t & 63u);
#pragma unroll
              for (uint32_t v = 0; v < 8; ++v) acc[v] |= v == tw ? bit : 0ull;
            }
          }
#pragma unroll
          for (uint32_t v = 0; v < 8; ++v) {
            uint64_t y = acc[v];
            if (v < W)
              for (int off = 32; off > 0; off >>= 1) y |= __shfl_xor(y, off, 64);
            S[v] = y;
          }
          n1 = dl ? uint32_t(dl[n1 * k + x]) : d[size_t(n1) * k + x];
          n2 = dl ? uint32_t(dl[n2 * k + x]) : d[size_t(n2) * k + x];
        }
        uint64_t mine = 0;
#pragma unroll
        for (uint32_t v = 0; v < 8; ++v) mine = v == lane ? S[v] : mine;
        if (lane < W) mask[lane] = mine;
      }
    }
    __syncthreads();
    if (*s_stop) return;
  }
  if (tid == 0) *out = EppOut{0, 1, total, wl};
}

// The preimage lists in LDS when they fit the launch's dynamic LDS (every
// candidate walk reads them; n = 300, k = 2: 3.6 KB): copies beg[k n + 1]
// and edges[k n] there and points *b, *e at the copies.  Every thread of the
// workgroup calls it.
constexpr uint32_t kEppListLds = 48 * 1024;
// Pair letters for the greedy walk (n * n bytes) after the lists in the
// dynamic LDS, 64 KB of dynamic LDS in all; transitions of up to
// kEppDeltaLds entries in static LDS.
constexpr uint32_t kEppDynLds = 64 * 1024;
constexpr uint32_t kEppDeltaLds = 4096;
// Pair tables of automata up to 64 states in static LDS (about 56 KB).
constexpr uint32_t kEppSmallPairs = 64 * 64;
constexpr uint32_t kEppSmallCnt = 2048;
__host__ __device__ inline uint32_t epp_letters_off(uint32_t list_lds) { return (list_lds + 15u) & ~15u; }
__device__ inline size_t epp_list_bytes(uint32_t kn) { return ((size_t(kn) + 1) * 4 + size_t(kn) * 2 + 7) & ~size_t(7); }
__device__ inline void epp_stage_lists(uint32_t kn, uint32_t lds_cap, unsigned char* lds, const uint32_t*& b,
    const uint16_t*& e) {
  if (epp_list_bytes(kn) > lds_cap) return;
  uint32_t* sb = reinterpret_cast<uint32_t*>(lds);
  uint16_t* se = reinterpret_cast<uint16_t*>(lds + (size_t(kn) + 1) * 4);
  for (uint32_t t = threadIdx.x; t <= kn; t += blockDim.x) sb[t] = b[t];
  for (uint32_t t = threadIdx.x; t < kn; t += blockDim.x) se[t] = e[t];
  __syncthreads();
  b = sb;
  e = se;
}

template <bool SMALL>
__global__ __launch_bounds__(kEppThreads) void k_eppstein_wg(const EppJob* __restrict__ jobs,
    const uint32_t* __restrict__ deltas, unsigned char* __restrict__ scratch, uint32_t* __restrict__ words,
    EppOut* __restrict__ outs, uint32_t lds_cap, uint32_t letter_lds) {
  extern __shared__ __align__(16) unsigned char epp_lds[];
  __shared__ uint32_t wsum[16];
  __shared__ uint64_t s_best[16];
  __shared__ uint64_t mask[8];
  __shared__ uint32_t s_stop;
  __shared__ uint16_t cs[kEppMaxN];
  __shared__ uint16_t dls[kEppDeltaLds];
  __shared__ uint32_t s_info[SMALL ? kEppSmallPairs : 1];
  __shared__ unsigned long long s_keys[SMALL ? kEppSmallPairs : 1];
  __shared__ uint32_t s_queue[SMALL ? kEppSmallCnt : 1], s_cnt[SMALL ? kEppSmallCnt : 1];
  const EppJob job = jobs[blockIdx.x];
  const uint32_t n = job.n, k = job.k;
  const uint32_t* d = deltas + job.delta_off;
  EppOut* out = outs + blockIdx.x;
  const uint32_t tid = threadIdx.x, nt = blockDim.x;
  if (n <= 1) {  // nothing to merge; the bound is left as it was
    if (tid == 0) *out = EppOut{0, 0, job.upper, 0};
    return;
  }
  // Scratch regions (epp_layout): scratch_off is a multiple of 256, every
  // region a multiple of 8.
  unsigned char* base = scratch + job.scratch_off;
  const EppLayout lay = epp_layout(n, k);
  uint32_t* beg = reinterpret_cast<uint32_t*>(base);
  uint16_t* edges = reinterpret_cast<uint16_t*>(base + lay.edges);
  uint32_t* info = reinterpret_cast<uint32_t*>(base + lay.info);
  unsigned long long* keys = reinterpret_cast<unsigned long long*>(base + lay.keys);
  const uint32_t all = n * (n - 1) / 2;
  uint32_t* queue = reinterpret_cast<uint32_t*>(base + lay.queue);
  uint32_t* cnt = reinterpret_cast<uint32_t*>(base + lay.cnt);
  // Automata of at most 64 states keep the pair tables in LDS: every
  // candidate of the pair BFS reads its pair's entry and may race for its
  // key, a chain of dependent accesses per parent that device memory makes
  // an L2 round trip each.
  if (SMALL && n * n <= kEppSmallPairs && k * n <= kEppSmallCnt && all <= kEppSmallCnt) {
    info = s_info;
    keys = s_keys;
    queue = s_queue;
    cnt = s_cnt;
  }

  // Preimage lists, sources ascending: list (a, t) at beg[a n + t].
  for (uint32_t at = tid; at < k * n; at += nt) {
    const uint32_t a = at / n, t = at % n;
    uint32_t c = 0;
    for (uint32_t q = 0; q < n; ++q) c += d[size_t(q) * k + a] == t ? 1u : 0u;
    cnt[at] = c;
  }
  for (uint32_t i = tid; i < n * n; i += nt) {
    info[i] = 0;
    keys[i] = kNoKey;
  }
  __syncthreads();
  {
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < k * n; b0 += nt) {
      const uint32_t i = b0 + tid;
      const uint32_t v = i < k * n ? cnt[i] : 0u;
      uint32_t tot;
      const uint32_t ex = block_scan(v, wsum, &tot);
      if (i < k * n) beg[i] = carry + ex;
      carry += tot;
    }
    if (tid == 0) beg[k * n] = carry;
  }
  __syncthreads();
  for (uint32_t at = tid; at < k * n; at += nt) {
    const uint32_t a = at / n, t = at % n;
    uint32_t o = beg[at];
    for (uint32_t q = 0; q < n; ++q)
      if (d[size_t(q) * k + a] == t) edges[o++] = uint16_t(q);
  }
  __syncthreads();
  const uint32_t* lbeg = beg;
  const uint16_t* ledges = edges;
  epp_stage_lists(k * n, lds_cap, epp_lds, lbeg, ledges);

  // ---- pair BFS
  uint32_t head = 0, tail = 0;  // this level's parents: queue[head, tail) (level 0: k n list pairs)
  uint32_t found = 0;
  for (uint32_t level = 0;; ++level) {
    const uint32_t parents = level == 0 ? k * n : tail - head;
    const uint32_t* lvl = queue + head;
    // 1. proposals
    for (uint32_t p = tid; p < parents; p += nt)
      candidates(level, p, n, k, lbeg, ledges, lvl, [&](uint32_t x, uint32_t y, uint64_t key, uint32_t) {
        const uint32_t ix = pair_ix(x, y, n);
        if (info[ix] == 0) atomicMin(&keys[ix], (unsigned long long)key);
      });
    __syncthreads();
    // 2. winners per parent, placed by a scan
    uint32_t carry = 0;
    const uint32_t wbase = tail;
    for (uint32_t b0 = 0; b0 < parents; b0 += nt) {
      const uint32_t p = b0 + tid;
      uint32_t c = 0;
      if (p < parents)
        candidates(level, p, n, k, lbeg, ledges, lvl, [&](uint32_t x, uint32_t y, uint64_t key, uint32_t) {
          const uint32_t ix = pair_ix(x, y, n);
          c += (info[ix] == 0 && keys[ix] == key) ? 1u : 0u;
        });
      uint32_t tot;
      const uint32_t ex = block_scan(c, wsum, &tot);
      if (p < parents) cnt[p] = carry + ex;
      carry += tot;
    }
    __syncthreads();
    // 3. winners written in order; their length and first letter recorded
    for (uint32_t p = tid; p < parents; p += nt) {
      uint32_t o = wbase + cnt[p];
      candidates(level, p, n, k, lbeg, ledges, lvl, [&](uint32_t x, uint32_t y, uint64_t key, uint32_t a) {
        const uint32_t ix = pair_ix(x, y, n);
        if (info[ix] == 0 && keys[ix] == key) {
          info[ix] = ((level + 1) << 8) | a;
          queue[o++] = x | (y << 16);
        }
      });
    }
    __syncthreads();
    found += carry;
    head = wbase;
    tail = wbase + carry;
    if (carry == 0 || found >= all) break;
  }
  if (found < all) {
    if (tid == 0) *out = EppOut{1, 0, job.upper, 0};
    return;
  }

  // The greedy walk's tables in LDS where they fit: the first letter of
  // every pair (after the list copy in the dynamic LDS) and the transitions.
  unsigned char* letters = nullptr;
  if (letter_lds && uint64_t(n) * n <= letter_lds) {
    letters = epp_lds + epp_letters_off(lds_cap);
    for (uint32_t i = tid; i < n * n; i += nt) letters[i] = (unsigned char)(info[i] & 255u);
  }
  const uint16_t* dl = nullptr;
  if (n * k <= kEppDeltaLds) {
    for (uint32_t i = tid; i < n * k; i += nt) dls[i] = uint16_t(d[i]);
    dl = dls;
  }
  __syncthreads();
  epp_greedy(job, n, k, d, info, words, out, cs, mask, s_best, &s_stop, letters, dl);
}

// ---- whole-device pair BFS of one automaton: the passes of the workgroup
// kernel above (proposals, winners per parent, a scan of the winner counts,
// winners written at the scanned offsets), one launch each over a grid sized
// for the largest level, then k_epp_advance.  Every launch reads the level
// from the device state, so the host enqueues several levels without
// waiting; launches after the last level find `done` set and return.
__device__ inline bool epp_level(const EppLevel* st, uint32_t level, uint32_t kn, EppLevel& s, uint32_t& parents) {
  s = *st;
  parents = level == 0 ? kn : s.tail - s.head;
  return !s.done;
}

__global__ __launch_bounds__(256) void k_epp_propose(uint32_t level, uint32_t n, uint32_t k,
    const uint32_t* __restrict__ beg, const uint16_t* __restrict__ edges, const uint32_t* __restrict__ queue,
    const uint32_t* __restrict__ info, unsigned long long* __restrict__ keys, const EppLevel* __restrict__ st,
    uint32_t lds_cap) {
  extern __shared__ __align__(16) unsigned char epp_lds[];
  EppLevel s;
  uint32_t parents;
  if (!epp_level(st, level, k * n, s, parents)) return;
  const uint32_t* lb = beg;
  const uint16_t* le = edges;
  epp_stage_lists(k * n, lds_cap, epp_lds, lb, le);
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= parents) return;
  candidates(level, p, n, k, lb, le, queue + s.head, [&](uint32_t x, uint32_t y, uint64_t key, uint32_t) {
    const uint32_t ix = pair_ix(x, y, n);
    if (info[ix] == 0) atomicMin(&keys[ix], (unsigned long long)key);
  });
}

__global__ __launch_bounds__(256) void k_epp_count(uint32_t level, uint32_t n, uint32_t k,
    const uint32_t* __restrict__ beg, const uint16_t* __restrict__ edges, const uint32_t* __restrict__ queue,
    const uint32_t* __restrict__ info, const unsigned long long* __restrict__ keys, uint32_t* __restrict__ cnt,
    const EppLevel* __restrict__ st, uint32_t lds_cap) {
  extern __shared__ __align__(16) unsigned char epp_lds[];
  EppLevel s;
  uint32_t parents;
  if (!epp_level(st, level, k * n, s, parents)) return;
  const uint32_t* lb = beg;
  const uint16_t* le = edges;
  epp_stage_lists(k * n, lds_cap, epp_lds, lb, le);
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= parents) return;
  uint32_t c = 0;
  candidates(level, p, n, k, lb, le, queue + s.head, [&](uint32_t x, uint32_t y, uint64_t key, uint32_t) {
    const uint32_t ix = pair_ix(x, y, n);
    c += (info[ix] == 0 && keys[ix] == key) ? 1u : 0u;
  });
  cnt[p] = c;
}

// Exclusive scan of the level's winner counts over tiles of kEppScanTile:
// tile sums, then each tile's entries offset by the earlier tiles' sums (one
// wave scan per row of 256); the last tile also writes the level's total.
constexpr uint32_t kEppScanPer = 8;
constexpr uint32_t kEppScanTile = 256 * kEppScanPer;

__device__ inline uint32_t epp_block_sum(uint32_t v, uint32_t* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63u) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  const uint32_t t = red[0] + red[1] + red[2] + red[3];
  __syncthreads();
  return t;
}

__global__ __launch_bounds__(256) void k_epp_scan_sums(uint32_t level, uint32_t kn, const uint32_t* __restrict__ cnt,
    uint32_t* __restrict__ tile_sum, const EppLevel* __restrict__ st) {
  __shared__ uint32_t red[4];
  EppLevel s;
  uint32_t parents;
  if (!epp_level(st, level, kn, s, parents)) return;
  const uint32_t base = blockIdx.x * kEppScanTile;
  if (base >= parents) return;
  uint32_t v = 0;
#pragma unroll
  for (uint32_t r = 0; r < kEppScanPer; ++r) {
    const uint32_t i = base + r * 256 + threadIdx.x;
    v += i < parents ? cnt[i] : 0u;
  }
  v = epp_block_sum(v, red);
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = v;
}

__global__ __launch_bounds__(256) void k_epp_scan_offsets(uint32_t level, uint32_t kn,
    const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ tile_sum, uint32_t* __restrict__ off,
    EppLevel* st) {
  __shared__ uint32_t red[4], wtot[4];
  EppLevel s;
  uint32_t parents;
  if (!epp_level(st, level, kn, s, parents)) return;
  const uint32_t base = blockIdx.x * kEppScanTile;
  if (base >= parents) return;
  uint32_t c = 0;
  for (uint32_t j = threadIdx.x; j < blockIdx.x; j += 256) c += tile_sum[j];
  uint32_t run = epp_block_sum(c, red);  // entries before the current row
  const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
  for (uint32_t r = 0; r < kEppScanPer; ++r) {
    const uint32_t i = base + r * 256 + threadIdx.x;
    const uint32_t v = i < parents ? cnt[i] : 0u;
    uint32_t x = v;  // inclusive scan within the wave
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = __shfl_up(x, d, 64);
      if (lane >= uint32_t(d)) x += y;
    }
    if (lane == 63) wtot[wv] = x;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t w = 0; w < wv; ++w) before += wtot[w];
    if (i < parents) off[i] = run + before + x - v;
    run += wtot[0] + wtot[1] + wtot[2] + wtot[3];
    __syncthreads();
  }
  if (threadIdx.x == 0 && base + kEppScanTile >= parents) st->total = run;
}

// Only a pair's winner writes its info, and it reads info before writing.
__global__ __launch_bounds__(256) void k_epp_write(uint32_t level, uint32_t n, uint32_t k,
    const uint32_t* __restrict__ beg, const uint16_t* __restrict__ edges, uint32_t* queue, uint32_t* info,
    const unsigned long long* __restrict__ keys, const uint32_t* __restrict__ off, const EppLevel* __restrict__ st,
    uint32_t lds_cap) {
  extern __shared__ __align__(16) unsigned char epp_lds[];
  EppLevel s;
  uint32_t parents;
  if (!epp_level(st, level, k * n, s, parents)) return;
  const uint32_t* lb = beg;
  const uint16_t* le = edges;
  epp_stage_lists(k * n, lds_cap, epp_lds, lb, le);
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= parents) return;
  uint32_t o = s.tail + off[p];
  candidates(level, p, n, k, lb, le, queue + s.head, [&](uint32_t x, uint32_t y, uint64_t key, uint32_t a) {
    const uint32_t ix = pair_ix(x, y, n);
    if (info[ix] == 0 && keys[ix] == key) {
      info[ix] = ((level + 1) << 8) | a;
      queue[o++] = x | (y << 16);
    }
  });
}

// The next level: queue[tail, tail + total); done when a level found no
// pair or every pair is found.
__global__ void k_epp_advance(EppLevel* st, uint32_t all) {
  if (st->done) return;
  const uint32_t t = st->total;
  st->found += t;
  st->head = st->tail;
  st->tail += t;
  if (t == 0 || st->found >= all) st->done = 1;
}

__global__ __launch_bounds__(kEppThreads) void k_epp_greedy(const EppJob* __restrict__ jobs,
    const uint32_t* __restrict__ deltas, const unsigned char* __restrict__ scratch, uint32_t* __restrict__ words,
    EppOut* __restrict__ outs, uint32_t letter_lds) {
  extern __shared__ __align__(16) unsigned char epp_lds[];
  __shared__ uint64_t s_best[16];
  __shared__ uint64_t mask[8];
  __shared__ uint32_t s_stop;
  __shared__ uint16_t cs[kEppMaxN];
  __shared__ uint16_t dls[kEppDeltaLds];
  const EppJob job = jobs[0];
  const uint32_t n = job.n, k = job.k;
  const EppLayout lay = epp_layout(n, k);
  const uint32_t* info = reinterpret_cast<const uint32_t*>(scratch + job.scratch_off + lay.info);
  const uint32_t* d = deltas + job.delta_off;
  unsigned char* letters = nullptr;
  if (letter_lds && uint64_t(n) * n <= letter_lds) {
    letters = epp_lds;
    for (uint32_t i = threadIdx.x; i < n * n; i += blockDim.x) letters[i] = (unsigned char)(info[i] & 255u);
  }
  const uint16_t* dl = nullptr;
  if (n * k <= kEppDeltaLds) {
    for (uint32_t i = threadIdx.x; i < n * k; i += blockDim.x) dls[i] = uint16_t(d[i]);
    dl = dls;
  }
  __syncthreads();
  epp_greedy(job, n, k, d, info, words, outs, cs, mask, s_best, &s_stop, letters, dl);
}

}  // namespace

size_t eppstein_scratch_bytes(uint32_t n, uint32_t k) { return epp_layout(n, k).total; }

uint32_t eppstein_list_lds(uint32_t n, uint32_t k) {
  const size_t need = ((size_t(k) * n + 1) * 4 + size_t(k) * n * 2 + 7) & ~size_t(7);
  return need <= kEppListLds ? uint32_t(need) : 0u;
}

uint32_t eppstein_letter_lds(uint32_t max_n, uint32_t list_lds) {
  const uint64_t need = uint64_t(max_n) * max_n;
  return epp_letters_off(list_lds) + need <= kEppDynLds ? uint32_t(need) : 0u;
}

hipError_t launch_eppstein_wg(const EppJob* jobs, uint32_t count, const uint32_t* deltas, unsigned char* scratch,
    uint32_t* words, EppOut* outs, uint32_t list_lds, uint32_t letter_lds, bool small, hipStream_t st) {
  if (!count) return hipSuccess;
  if (list_lds > kEppListLds || (letter_lds && epp_letters_off(list_lds) + letter_lds > kEppDynLds))
    return hipErrorInvalidValue;
  const uint32_t dyn = letter_lds ? epp_letters_off(list_lds) + letter_lds : list_lds;
  if (small)
    hipLaunchKernelGGL(k_eppstein_wg<true>, dim3(count), dim3(kEppThreads), dyn, st, jobs, deltas, scratch, words,
        outs, uint32_t(list_lds), letter_lds);
  else
    hipLaunchKernelGGL(k_eppstein_wg<false>, dim3(count), dim3(kEppThreads), dyn, st, jobs, deltas, scratch, words,
        outs, uint32_t(list_lds), letter_lds);
  return hipGetLastError();
}

hipError_t launch_epp_level(uint32_t level, uint32_t n, uint32_t k, unsigned char* base, uint32_t* off,
    uint32_t* tile_sum, EppLevel* st, hipStream_t st_) {
  const EppLayout lay = epp_layout(n, k);
  const uint32_t* beg = reinterpret_cast<const uint32_t*>(base);
  const uint16_t* edges = reinterpret_cast<const uint16_t*>(base + lay.edges);
  uint32_t* info = reinterpret_cast<uint32_t*>(base + lay.info);
  unsigned long long* keys = reinterpret_cast<unsigned long long*>(base + lay.keys);
  uint32_t* queue = reinterpret_cast<uint32_t*>(base + lay.queue);
  uint32_t* cnt = reinterpret_cast<uint32_t*>(base + lay.cnt);
  const uint32_t kn = k * n, most = std::max(kn, n * (n - 1) / 2);
  const dim3 grid((most + 255) / 256), block(256), tiles((most + kEppScanTile - 1) / kEppScanTile);
  const uint32_t lds = eppstein_list_lds(n, k);
  hipLaunchKernelGGL(k_epp_propose, grid, block, lds, st_, level, n, k, beg, edges, queue, info, keys, st, lds);
  hipLaunchKernelGGL(k_epp_count, grid, block, lds, st_, level, n, k, beg, edges, queue, info, keys, cnt, st, lds);
  hipLaunchKernelGGL(k_epp_scan_sums, tiles, block, 0, st_, level, kn, cnt, tile_sum, st);
  hipLaunchKernelGGL(k_epp_scan_offsets, tiles, block, 0, st_, level, kn, cnt, tile_sum, off, st);
  hipLaunchKernelGGL(k_epp_write, grid, block, lds, st_, level, n, k, beg, edges, queue, info, keys, off, st, lds);
  hipLaunchKernelGGL(k_epp_advance, dim3(1), dim3(1), 0, st_, st, n * (n - 1) / 2);
  return hipGetLastError();
}

uint32_t epp_scan_tiles(uint32_t n, uint32_t k) {
  const uint32_t most = std::max(k * n, n * (n - 1) / 2);
  return (most + kEppScanTile - 1) / kEppScanTile;
}

hipError_t launch_epp_greedy(const EppJob* job, uint32_t n, const uint32_t* deltas, const unsigned char* scratch,
    uint32_t* words, EppOut* out, hipStream_t st) {
  const uint32_t letters = eppstein_letter_lds(n, 0);
  hipLaunchKernelGGL(k_epp_greedy, dim3(1), dim3(kEppThreads), letters, st, job, deltas, scratch, words, out,
      letters);
  return hipGetLastError();
}

}  // namespace sw
