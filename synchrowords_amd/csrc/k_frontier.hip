// Frontier kernels: expansion through slice tables, exact dedup in an
// open-addressing hash set, stream compaction, per-state counts.
#include "sw_device.hpp"
#include "sw_kernels.hpp"

namespace sw {

// ---------------------------------------------------------------- expansion
// One thread per child.  The child is the OR over slices s of table row
// [letter][s][byte s of the parent] (the reference's
// PreprocessedTransition::apply, preprocessed_transition.hpp:31-49); a zero
// byte selects the empty row.  Large launches skip it (a branch per byte);
// small ones (ALL) read every row, zero bytes included, with no branch, so
// that a thread's loads are all in flight together: a launch of a few
// thousand children is bound by the latency of its row reads, not by their
// number.
constexpr uint64_t kExpandAllMax = 1ull << 17;  // children

template <int W, bool ALL>
__global__ __launch_bounds__(256) void k_expand(const uint64_t* __restrict__ parents, uint64_t m,
    const uint64_t* __restrict__ tables, uint32_t slices, uint32_t letters, const uint32_t* __restrict__ letter_list,
    uint32_t n, int drop_mode, uint64_t* __restrict__ out, uint32_t* __restrict__ valid) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m * letters) return;
  const uint64_t ai = i / m, j = i - ai * m;
  const uint32_t a = letter_list ? letter_list[ai] : uint32_t(ai);
  uint64_t x[W], r[W];
#pragma unroll
  for (int w = 0; w < W; ++w) {
    x[w] = parents[j * W + w];
    r[w] = 0;
  }
  const uint64_t* tab = tables + uint64_t(a) * slices * 256 * W;
  if (ALL) {
    // One parent word at a time: its eight rows' loads in flight together.
    // Slices past the last read the empty row of slice 0.
#pragma unroll 1
    for (int w = 0; w < W; ++w) {
      const uint64_t xw = parents[j * W + w];
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const uint32_t s = uint32_t(w) * 8 + b;
        const bool in = s < slices;
        const uint32_t byte = in ? (uint32_t(xw >> (b * 8)) & 255u) : 0u;
        const uint64_t* row = tab + (uint64_t(in ? s : 0u) * 256 + byte) * W;
#pragma unroll
        for (int v = 0; v < W; ++v) r[v] |= row[v];
      }
    }
  } else {
#pragma unroll
    for (int w = 0; w < W; ++w) {
      for (int b = 0; b < 8; ++b) {
        const uint32_t s = uint32_t(w) * 8 + b;
        if (s >= slices) break;
        const uint32_t byte = uint32_t(x[w] >> (b * 8)) & 255u;
        if (!byte) continue;
        const uint64_t* row = tab + (uint64_t(s) * 256 + byte) * W;
#pragma unroll
        for (int v = 0; v < W; ++v) r[v] |= row[v];
      }
    }
  }
#pragma unroll
  for (int w = 0; w < W; ++w) out[i * W + w] = r[w];
  if (valid) {
    uint32_t ok = 1;
    if (drop_mode == 1) ok = set_popc<W>(r) + 1 < int(n);
    valid[i] = ok;
  }
}

template <int W>
static hipError_t expand_w(const uint64_t* parents, uint64_t m, const uint64_t* tables, uint32_t slices,
    uint32_t letters, const uint32_t* letter_list, uint32_t n, int drop_mode, uint64_t* out, uint32_t* valid,
    hipStream_t st) {
  const uint64_t total = m * letters;
  if (!total) return hipSuccess;
  const uint64_t blocks = (total + 255) / 256;
  if (total <= kExpandAllMax)
    hipLaunchKernelGGL((k_expand<W, true>), dim3(unsigned(blocks)), dim3(256), 0, st, parents, m, tables, slices,
        letters, letter_list, n, drop_mode, out, valid);
  else
    hipLaunchKernelGGL((k_expand<W, false>), dim3(unsigned(blocks)), dim3(256), 0, st, parents, m, tables, slices,
        letters, letter_list, n, drop_mode, out, valid);
  return hipGetLastError();
}

// Expansion of large launches, bit-sliced: a wave takes 64 parents, loads
// them coalesced through LDS and transposes each word across the lanes
// (lane_transpose64), so that lane l holds, per word w, the 64-bit mask of
// the parents that contain state q = 64 w + l.  Then per letter a:
//   image:    the child mask of state t is the OR of the masks of the states
//             q with delta(q, a) = t, one 64-bit LDS atomic OR per state;
//   preimage: the child mask of state q is the parent mask of delta(q, a),
//             one 64-bit LDS read per state
// (the preimage of a set commutes with the complement, so the complemented
// rows of the backward side expand the same way), and the masks are
// transposed back and written coalesced through LDS.  About W transposes per
// letter and per 64 children, against ceil(n/8) slice-table gathers per
// child in k_expand (1.77 ms for 2^24 children at n = 300, VALU-bound by set
// bits at 1.68 ms).  targets[a * n + q] = delta(q, a).
constexpr uint32_t kSlicedWaves = 4;

template <int W, bool BACK>
__global__ __launch_bounds__(64 * kSlicedWaves) void k_expand_sliced(const uint64_t* __restrict__ parents, uint64_t m,
    const uint16_t* __restrict__ targets, uint32_t letters, const uint32_t* __restrict__ letter_list, uint32_t n,
    int drop_mode, uint64_t* __restrict__ out, uint32_t* __restrict__ valid) {
  __shared__ uint64_t stage[kSlicedWaves][64 * W];
  __shared__ uint64_t pmask[BACK ? kSlicedWaves : 1][BACK ? 64 * W : 1];
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint64_t j0 = (uint64_t(blockIdx.x) * kSlicedWaves + wv) * 64;
  const uint64_t words_here = j0 < m ? (m - j0 < 64 ? (m - j0) * W : uint64_t(64) * W) : 0;
  uint64_t* my = stage[wv];
  // Parents in, coalesced: word e of the wave's block is my[e].
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const uint32_t e = uint32_t(w) * 64 + lane;
    my[e] = e < words_here ? parents[j0 * W + e] : 0ull;
  }
  __syncthreads();
  uint64_t pres[W];
#pragma unroll
  for (int w = 0; w < W; ++w) pres[w] = lane_transpose64(my[lane * W + w]);
  if (BACK) {
#pragma unroll
    for (int w = 0; w < W; ++w) pmask[wv][w * 64 + lane] = pres[w];
  }
  __syncthreads();
  for (uint32_t ai = 0; ai < letters; ++ai) {
    const uint32_t a = letter_list ? letter_list[ai] : ai;
    const uint16_t* ta = targets + uint64_t(a) * n;
    uint64_t y[W];
    if (BACK) {
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const uint32_t q = uint32_t(w) * 64 + lane;
        y[w] = q < n ? pmask[wv][ta[q]] : 0ull;
      }
    } else {
#pragma unroll
      for (int w = 0; w < W; ++w) my[w * 64 + lane] = 0ull;
      __syncthreads();
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const uint32_t q = uint32_t(w) * 64 + lane;
        if (q < n && pres[w]) atomicOr(reinterpret_cast<unsigned long long*>(&my[ta[q]]), pres[w]);
      }
      __syncthreads();
#pragma unroll
      for (int w = 0; w < W; ++w) y[w] = my[w * 64 + lane];
      __syncthreads();
    }
    // Back to one child per lane, then out through LDS.
    int pc = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      y[w] = lane_transpose64(y[w]);
      pc += __popcll(y[w]);
      my[lane * W + w] = y[w];
    }
    const uint64_t base = uint64_t(ai) * m + j0;
    if (valid && j0 + lane < m) valid[base + lane] = drop_mode == 1 ? (pc + 1 < int(n) ? 1u : 0u) : 1u;
    __syncthreads();
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const uint32_t e = uint32_t(w) * 64 + lane;
      if (e < words_here) out[base * W + e] = my[e];
    }
    __syncthreads();
  }
}

template <int W>
static hipError_t expand_sliced_w(const uint64_t* parents, uint64_t m, const uint16_t* targets, int dir,
    uint32_t letters, const uint32_t* letter_list, uint32_t n, int drop_mode, uint64_t* out, uint32_t* valid,
    hipStream_t st) {
  if (!m || !letters) return hipSuccess;
  const uint64_t blocks = (m + 64 * kSlicedWaves - 1) / (64 * kSlicedWaves);
  if (dir)
    hipLaunchKernelGGL((k_expand_sliced<W, true>), dim3(unsigned(blocks)), dim3(64 * kSlicedWaves), 0, st, parents, m,
        targets, letters, letter_list, n, drop_mode, out, valid);
  else
    hipLaunchKernelGGL((k_expand_sliced<W, false>), dim3(unsigned(blocks)), dim3(64 * kSlicedWaves), 0, st, parents,
        m, targets, letters, letter_list, n, drop_mode, out, valid);
  return hipGetLastError();
}

// Slice tables from per-state masks: row (a, s, b) = OR of masks[a][8 s + j]
// over the set bits j of b (masks: the image singleton of each state, or its
// preimage set).  One thread per row.
template <int W>
__global__ __launch_bounds__(256) void k_tables(const uint64_t* __restrict__ masks, uint32_t n, uint32_t slices,
    uint64_t rows, uint64_t* __restrict__ out) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  const uint32_t b = uint32_t(i & 255u);
  const uint64_t as = i >> 8;
  const uint32_t s = uint32_t(as % slices), a = uint32_t(as / slices);
  uint64_t r[W];
#pragma unroll
  for (int w = 0; w < W; ++w) r[w] = 0;
  for (uint32_t j = 0; j < 8; ++j) {
    const uint32_t q = s * 8 + j;
    if (!((b >> j) & 1u) || q >= n) continue;
    const uint64_t* m = masks + (uint64_t(a) * n + q) * W;
#pragma unroll
    for (int w = 0; w < W; ++w) r[w] |= m[w];
  }
#pragma unroll
  for (int w = 0; w < W; ++w) out[i * W + w] = r[w];
}

template <int W>
static hipError_t tables_w(const uint64_t* masks, uint32_t n, uint32_t k, uint64_t* out, hipStream_t st) {
  const uint32_t slices = (n + 7) / 8;
  const uint64_t rows = uint64_t(k) * slices * 256;
  hipLaunchKernelGGL(k_tables<W>, dim3(unsigned((rows + 255) / 256)), dim3(256), 0, st, masks, n, slices, rows, out);
  return hipGetLastError();
}

// ------------------------------------------------------------------- hashing
// Linear probing.  A candidate claims an empty slot with CAS; on meeting an
// equal set it is a duplicate, unless that set is a candidate of this step
// with a larger index, which it then displaces — so the survivor of a group
// of equal new sets is the one with the smallest index, independent of
// timing.
template <int W>
__global__ __launch_bounds__(256) void k_hash_insert(const uint64_t* __restrict__ cand,
    const uint32_t* __restrict__ valid, uint64_t count, uint32_t* table, uint64_t mask,
    const uint64_t* __restrict__ arena, uint32_t* __restrict__ slot_of, unsigned long long* probes) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  uint32_t nprobe = 0;
  if (i < count) {
    uint32_t res = kEmpty;
    if (!valid || valid[i]) {
      uint64_t s[W];
#pragma unroll
      for (int w = 0; w < W; ++w) s[w] = cand[i * W + w];
      const uint32_t mine = kCand | uint32_t(i);
      uint64_t pos = set_hash<W>(s) & mask;
      while (true) {
        ++nprobe;
        // Claim first: at most half the slots are taken, so the first CAS
        // usually succeeds; a failed one returns the slot's holder.
        uint32_t v = atomicCAS(&table[pos], kEmpty, mine);
        if (v == kEmpty) {
          res = uint32_t(pos);
          goto done;
        }
        bool advance = false;
        while (!advance) {
          if (v == kEmpty) {
            uint32_t prev = atomicCAS(&table[pos], kEmpty, mine);
            if (prev == kEmpty) {
              res = uint32_t(pos);
              goto done;
            }
            v = prev;
            continue;
          }
          const uint64_t* other = (v & kCand) ? cand + uint64_t(v & ~kCand) * W : arena + uint64_t(v) * W;
          if (!set_eq<W>(s, other)) {
            advance = true;
            break;
          }
          if ((v & kCand) && (v & ~kCand) > uint32_t(i)) {
            uint32_t prev = atomicCAS(&table[pos], v, mine);
            if (prev == v) {
              res = uint32_t(pos);
              goto done;
            }
            v = prev;
            continue;
          }
          goto done;  // equal to an older row or a smaller candidate
        }
        pos = (pos + 1) & mask;
      }
    done:
      slot_of[i] = res;
    } else {
      slot_of[i] = kEmpty;
    }
  }
  if (probes) {
    uint64_t c = nprobe;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(probes, (unsigned long long)c);
  }
}

// valid[i] = 0 for a candidate equal to a row the table holds (the table
// holds rows of `arena`; marks of candidates of an unfinished step are
// skipped).  Read only: the backward-deepening phase drops sets the side
// already produced at a smaller or equal depth.
template <int W>
__global__ __launch_bounds__(256) void k_hash_lookup(const uint64_t* __restrict__ cand, uint32_t* __restrict__ valid,
    uint64_t count, const uint32_t* __restrict__ table, uint64_t mask, const uint64_t* __restrict__ arena) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count || !valid[i]) return;
  uint64_t s[W];
#pragma unroll
  for (int w = 0; w < W; ++w) s[w] = cand[i * W + w];
  uint64_t pos = set_hash<W>(s) & mask;
  while (true) {
    const uint32_t v = table[pos];
    if (v == kEmpty) return;
    if (!(v & kCand) && set_eq<W>(s, arena + uint64_t(v) * W)) {
      valid[i] = 0;
      return;
    }
    pos = (pos + 1) & mask;
  }
}

__global__ __launch_bounds__(256) void k_hash_resolve(const uint32_t* __restrict__ slot_of,
    const uint32_t* __restrict__ table, uint64_t count, uint32_t* __restrict__ keep) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t s = slot_of[i];
  keep[i] = (s != kEmpty && table[s] == (kCand | uint32_t(i))) ? 1u : 0u;
}

template <int W>
__global__ __launch_bounds__(256) void k_commit(const uint64_t* __restrict__ cand, const uint32_t* __restrict__ keep,
    const uint32_t* __restrict__ pos, const uint32_t* __restrict__ slot_of, uint64_t count, uint64_t m,
    const uint32_t* __restrict__ parent_idx, uint64_t parent_base, const uint32_t* __restrict__ letter_list,
    uint64_t base, uint64_t* __restrict__ arena, uint32_t* __restrict__ parent, uint8_t* __restrict__ letter,
    uint32_t* table, uint64_t* __restrict__ front, uint32_t* __restrict__ front_idx, int* found,
    unsigned long long* witness) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i == 0 && found) {  // arm the goal check that follows
    *found = 0;
    *witness = ~0ull;
  }
  if (i >= count || !keep[i]) return;
  const uint64_t dst = base + pos[i];
#pragma unroll
  for (int w = 0; w < W; ++w) arena[dst * W + w] = cand[i * W + w];
  const uint64_t ai = i / m;
  if (parent) parent[dst] = parent_idx ? parent_idx[i - ai * m] : uint32_t(parent_base + (i - ai * m));
  if (letter) letter[dst] = uint8_t(letter_list ? letter_list[ai] : ai);
  if (table) table[slot_of[i]] = uint32_t(dst);
  if (front) {  // the new level as the next frontier, in row order
#pragma unroll
    for (int w = 0; w < W; ++w) front[uint64_t(pos[i]) * W + w] = cand[i * W + w];
    front_idx[pos[i]] = uint32_t(dst);
  }
}

__global__ __launch_bounds__(256) void k_alive(const uint8_t* __restrict__ dead, uint64_t count,
    uint32_t* __restrict__ keep) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  keep[i] = ((dead[i >> 3] >> (i & 7)) & 1) ? 0u : 1u;
}

template <int W>
__global__ __launch_bounds__(256) void k_select(const uint64_t* __restrict__ rows, uint64_t lo,
    const uint32_t* __restrict__ keep, const uint32_t* __restrict__ pos, uint64_t count, uint64_t* __restrict__ out,
    uint32_t* __restrict__ idx_out) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count || (keep && !keep[i])) return;
  const uint64_t j = pos ? pos[i] : i;
#pragma unroll
  for (int w = 0; w < W; ++w) out[j * W + w] = rows[(lo + i) * W + w];
  if (idx_out) idx_out[j] = uint32_t(lo + i);
}

template <int W>
__global__ __launch_bounds__(256) void k_beam_select(const uint64_t* __restrict__ cand,
    const uint32_t* __restrict__ keep, const uint32_t* __restrict__ pos, uint64_t count, uint32_t n,
    uint64_t* __restrict__ out, uint32_t* __restrict__ idx_out, uint32_t* __restrict__ idx_copy,
    unsigned long long* first, uint32_t* __restrict__ table, const uint32_t* __restrict__ slot_of) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count || !keep[i]) return;
  table[slot_of[i]] = kEmpty;
  const uint64_t j = pos[i];
  bool full = true;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const uint64_t x = cand[i * W + w];
    out[j * W + w] = x;
    full &= x == word_mask(n, w);
  }
  idx_out[j] = uint32_t(i);
  idx_copy[j] = uint32_t(i);
  if (full) *first = i;  // one kept candidate at most is Q
}

template <int W>
__global__ __launch_bounds__(256) void k_count_bits(const uint64_t* __restrict__ sets, uint64_t count, uint32_t n,
    unsigned long long* acc) {
  __shared__ uint32_t hist[kMaxW * 64];
  for (uint32_t q = threadIdx.x; q < n; q += blockDim.x) hist[q] = 0;
  __syncthreads();
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += uint64_t(gridDim.x) * blockDim.x) {
#pragma unroll
    for (int w = 0; w < W; ++w) {
      uint64_t x = sets[i * W + w];
      while (x) {
        atomicAdd(&hist[w * 64 + __builtin_ctzll(x)], 1u);
        x &= x - 1;
      }
    }
  }
  __syncthreads();
  for (uint32_t q = threadIdx.x; q < n; q += blockDim.x)
    if (hist[q]) atomicAdd(&acc[q], (unsigned long long)hist[q]);
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t count) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += uint64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

// Plain streaming copy (16-byte vectors): the calibration kernel whose
// counter readings scale the HBM traffic of the other kernels.
__global__ __launch_bounds__(256) void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n16) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n16; i += uint64_t(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

// Keep flags and their exclusive scan (launch_keep_scan):
//   mode 0: keep[i] given;
//   mode 1: keep[i] = candidate i owns its hash slot (as k_hash_resolve);
//   mode 2: keep[i] = bit i of `dead` is clear (as k_alive).
// pos[i] = sum of keep[j], j < i; *total = the sum of all.
//
// The list is cut into tiles of kScanTile entries, kScanPer per thread at a
// stride of kScanThreads so that every load and store is coalesced and a
// thread's kScanPer loads are independent.  A list of one tile takes one
// launch (k_keep_tile_scan with no tile sums); longer ones take two:
// k_keep_tile_sums writes the flags and each tile's total, then each tile of
// k_keep_tile_scan adds up the totals of the tiles before it and scans its own
// entries with one ballot per row of kScanThreads.
constexpr uint32_t kScanThreads = 256;
constexpr uint32_t kScanPer = 8;
constexpr uint32_t kScanTile = kScanThreads * kScanPer;
constexpr uint32_t kScanWaves = kScanThreads / 64;
constexpr uint64_t kScanTwoPassMax = uint64_t(kScanTile) * 1024;  // at most 1024 tiles

__device__ __forceinline__ void keep_flags(int mode, const uint32_t* __restrict__ slot_of,
    const uint32_t* __restrict__ table, const uint8_t* __restrict__ dead, uint64_t base, uint64_t count,
    uint32_t* __restrict__ keep, uint32_t (&k)[kScanPer]) {
  const uint32_t t = threadIdx.x;
  if (mode == 1) {
    uint32_t s[kScanPer];
#pragma unroll
    for (uint32_t r = 0; r < kScanPer; ++r) {
      const uint64_t i = base + r * kScanThreads + t;
      s[r] = i < count ? slot_of[i] : kEmpty;
    }
#pragma unroll
    for (uint32_t r = 0; r < kScanPer; ++r) {
      const uint64_t i = base + r * kScanThreads + t;
      k[r] = (s[r] != kEmpty && table[s[r]] == (kCand | uint32_t(i))) ? 1u : 0u;
    }
  } else if (mode == 2) {
#pragma unroll
    for (uint32_t r = 0; r < kScanPer; ++r) {
      const uint64_t i = base + r * kScanThreads + t;
      k[r] = (i < count && !((dead[i >> 3] >> (i & 7)) & 1)) ? 1u : 0u;
    }
  } else {
#pragma unroll
    for (uint32_t r = 0; r < kScanPer; ++r) {
      const uint64_t i = base + r * kScanThreads + t;
      k[r] = i < count ? keep[i] : 0u;
    }
    return;
  }
#pragma unroll
  for (uint32_t r = 0; r < kScanPer; ++r) {
    const uint64_t i = base + r * kScanThreads + t;
    if (i < count) keep[i] = k[r];
  }
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

__global__ __launch_bounds__(kScanThreads) void k_keep_tile_sums(int mode, const uint32_t* __restrict__ slot_of,
    const uint32_t* __restrict__ table, const uint8_t* __restrict__ dead, uint64_t count,
    uint32_t* __restrict__ keep, uint32_t* __restrict__ tile_sum) {
  __shared__ uint32_t wsum[kScanWaves];
  uint32_t k[kScanPer];
  keep_flags(mode, slot_of, table, dead, uint64_t(blockIdx.x) * kScanTile, count, keep, k);
  uint32_t s = 0;
#pragma unroll
  for (uint32_t r = 0; r < kScanPer; ++r) s += k[r];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
#pragma unroll
    for (uint32_t w = 0; w < kScanWaves; ++w) t += wsum[w];
    tile_sum[blockIdx.x] = t;
  }
}

// tile_sum == NULL: one tile (gridDim.x == 1), flags computed here.
// tile_off != NULL: tile_off[t] already holds the entries before tile t (the
// scanned tile sums of a list of more than 1024 tiles).
__global__ __launch_bounds__(kScanThreads) void k_keep_tile_scan(int mode, const uint32_t* __restrict__ slot_of,
    const uint32_t* __restrict__ table, const uint8_t* __restrict__ dead, uint64_t count,
    uint32_t* __restrict__ keep, const uint32_t* __restrict__ tile_sum, uint32_t* __restrict__ pos,
    uint32_t* __restrict__ total, const uint32_t* __restrict__ tile_off) {
  __shared__ uint32_t wcnt[kScanPer][kScanWaves];
  __shared__ uint32_t wcarry[kScanWaves];
  const uint32_t t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const uint64_t base = uint64_t(blockIdx.x) * kScanTile;
  uint32_t k[kScanPer];
  uint32_t c = 0;
  if (tile_sum) {
#pragma unroll
    for (uint32_t r = 0; r < kScanPer; ++r) {
      const uint64_t i = base + r * kScanThreads + t;
      k[r] = i < count ? keep[i] : 0u;
    }
    if (tile_off) {
      c = t == 0 ? tile_off[blockIdx.x] : 0u;
    } else {
      for (uint32_t j = t; j < blockIdx.x; j += kScanThreads) c += tile_sum[j];
    }
    c = wave_sum(c);
  } else {
    keep_flags(mode, slot_of, table, dead, base, count, keep, k);
  }
  uint64_t m[kScanPer];
#pragma unroll
  for (uint32_t r = 0; r < kScanPer; ++r) m[r] = __ballot(k[r]);
  if (lane == 0) {
    wcarry[wv] = c;
#pragma unroll
    for (uint32_t r = 0; r < kScanPer; ++r) wcnt[r][wv] = uint32_t(__popcll(m[r]));
  }
  __syncthreads();
  uint32_t run = 0;  // entries before the current row: earlier tiles, then earlier rows
#pragma unroll
  for (uint32_t w = 0; w < kScanWaves; ++w) run += wcarry[w];
#pragma unroll
  for (uint32_t r = 0; r < kScanPer; ++r) {
    uint32_t before = 0, row = 0;
#pragma unroll
    for (uint32_t w = 0; w < kScanWaves; ++w) {
      const uint32_t x = wcnt[r][w];
      before += w < wv ? x : 0u;
      row += x;
    }
    const uint64_t i = base + r * kScanThreads + t;
    const uint32_t below = __builtin_amdgcn_mbcnt_hi(uint32_t(m[r] >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m[r]), 0u));
    if (i < count) pos[i] = run + before + below;
    run += row;
  }
  if (t == 0 && blockIdx.x == gridDim.x - 1) *total = run;
}

// Resets the goal flag and witness of a meet check.
__global__ void k_goal_reset(int* found, unsigned long long* witness) {
  *found = 0;
  *witness = ~0ull;
}


// Exclusive scan of arbitrary 32-bit values (sums wrap mod 2^32): tiles of
// kScanTile entries, eight consecutive entries per thread (two 16-byte loads),
// a block scan of the 256 thread totals (wave shuffles, then the waves'
// totals in LDS).  Longer lists scan their tile sums first, with the same
// kernels, and add them back in k_scan_apply.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t lane) {
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, 64);
    if (lane >= uint32_t(off)) x += y;
  }
  return x - v;
}

__device__ __forceinline__ void load8(const uint32_t* __restrict__ in, uint64_t i0, uint64_t count, uint32_t (&v)[8]) {
  if (i0 + 8 <= count && ((reinterpret_cast<uintptr_t>(in + i0) & 15) == 0)) {
    const uint4 a = *reinterpret_cast<const uint4*>(in + i0);
    const uint4 b = *reinterpret_cast<const uint4*>(in + i0 + 4);
    v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w, v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
  } else {
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = i0 + r < count ? in[i0 + r] : 0u;
  }
}

__global__ __launch_bounds__(kScanThreads) void k_scan_tile_sums(const uint32_t* __restrict__ in, uint64_t count,
    uint32_t* __restrict__ tile_sum) {
  __shared__ uint32_t wsum[kScanWaves];
  uint32_t v[8];
  load8(in, uint64_t(blockIdx.x) * kScanTile + uint64_t(threadIdx.x) * 8, count, v);
  uint32_t s = 0;
#pragma unroll
  for (int r = 0; r < 8; ++r) s += v[r];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
#pragma unroll
    for (uint32_t w = 0; w < kScanWaves; ++w) t += wsum[w];
    tile_sum[blockIdx.x] = t;
  }
}

// out may alias in.  tile_off == NULL: one tile.  The last tile writes *total.
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const uint32_t* in, uint64_t count,
    const uint32_t* __restrict__ tile_off, uint32_t* out, uint32_t* __restrict__ total) {
  __shared__ uint32_t wsum[kScanWaves];
  const uint32_t t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const uint64_t i0 = uint64_t(blockIdx.x) * kScanTile + uint64_t(t) * 8;
  uint32_t v[8];
  load8(in, i0, count, v);
  uint32_t s = 0;
#pragma unroll
  for (int r = 0; r < 8; ++r) s += v[r];
  const uint32_t ex = wave_excl_scan(s, lane);
  if (lane == 63) wsum[wv] = ex + s;
  __syncthreads();
  uint32_t run = tile_off ? tile_off[blockIdx.x] : 0u, all = run;
#pragma unroll
  for (uint32_t w = 0; w < kScanWaves; ++w) {
    run += w < wv ? wsum[w] : 0u;
    all += wsum[w];
  }
  run += ex;
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    if (i0 + r < count) out[i0 + r] = run;
    run += v[r];
  }
  if (t == 0 && blockIdx.x == gridDim.x - 1 && total) *total = all;
}

// ------------------------------------------------------------------ launchers
#define SW_DISPATCH_W(W, CALL)      \
  switch (W) {                      \
    case 1: { constexpr int w_ = 1; return CALL; } \
    case 2: { constexpr int w_ = 2; return CALL; } \
    case 3: { constexpr int w_ = 3; return CALL; } \
    case 4: { constexpr int w_ = 4; return CALL; } \
    case 5: { constexpr int w_ = 5; return CALL; } \
    case 6: { constexpr int w_ = 6; return CALL; } \
    case 7: { constexpr int w_ = 7; return CALL; } \
    case 8: { constexpr int w_ = 8; return CALL; } \
  }                                 \
  return hipErrorInvalidValue;

static inline unsigned blocks_for(uint64_t count, unsigned bs = 256) { return unsigned((count + bs - 1) / bs); }

hipError_t launch_expand(int W, const uint64_t* parents, uint64_t m, const uint64_t* tables, uint32_t slices,
    uint32_t letters, const uint32_t* letter_list, uint32_t n, int drop_mode, uint64_t* out, uint32_t* valid,
    hipStream_t st) {
  SW_DISPATCH_W(W, expand_w<w_>(parents, m, tables, slices, letters, letter_list, n, drop_mode, out, valid, st))
}

hipError_t launch_expand_sets(int W, const uint64_t* parents, uint64_t m, const uint64_t* tables, uint32_t slices,
    const uint16_t* targets, int dir, uint32_t letters, const uint32_t* letter_list, uint32_t n, int drop_mode,
    uint64_t* out, uint32_t* valid, hipStream_t st) {
  if (m * letters <= kExpandAllMax || !targets)
    return launch_expand(W, parents, m, tables, slices, letters, letter_list, n, drop_mode, out, valid, st);
  SW_DISPATCH_W(W, expand_sliced_w<w_>(parents, m, targets, dir, letters, letter_list, n, drop_mode, out, valid, st))
}

hipError_t launch_tables(int W, const uint64_t* masks, uint32_t n, uint32_t k, uint64_t* out, hipStream_t st) {
  SW_DISPATCH_W(W, tables_w<w_>(masks, n, k, out, st))
}

template <int W>
static hipError_t hash_insert_w(const uint64_t* cand, const uint32_t* valid, uint64_t count, uint32_t* table,
    uint64_t mask, const uint64_t* arena, uint32_t* slot_of, unsigned long long* probes, hipStream_t st) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(k_hash_insert<W>, dim3(blocks_for(count)), dim3(256), 0, st, cand, valid, count, table, mask,
      arena, slot_of, probes);
  return hipGetLastError();
}

hipError_t launch_hash_insert(int W, const uint64_t* cand, const uint32_t* valid, uint64_t count, uint32_t* table,
    uint64_t mask, const uint64_t* arena, uint32_t* slot_of, unsigned long long* probes, hipStream_t st) {
  SW_DISPATCH_W(W, hash_insert_w<w_>(cand, valid, count, table, mask, arena, slot_of, probes, st))
}

hipError_t launch_hash_resolve(const uint32_t* slot_of, const uint32_t* table, uint64_t count, uint32_t* keep,
    hipStream_t st) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(k_hash_resolve, dim3(blocks_for(count)), dim3(256), 0, st, slot_of, table, count, keep);
  return hipGetLastError();
}

// Words of scratch the scan of `count` values needs: the tile sums of every
// level above the first.
static uint64_t scan_temp_words(uint64_t count) {
  const uint64_t tiles = (count + kScanTile - 1) / kScanTile;
  return tiles <= 1 ? 0 : tiles + scan_temp_words(tiles);
}

static hipError_t scan_levels(const uint32_t* in, uint32_t* out, uint64_t count, uint32_t* total, uint32_t* temp,
    hipStream_t st) {
  const uint64_t tiles = (count + kScanTile - 1) / kScanTile;
  if (tiles <= 1) {
    hipLaunchKernelGGL(k_scan_apply, dim3(1), dim3(kScanThreads), 0, st, in, count, nullptr, out, total);
    return hipGetLastError();
  }
  uint32_t* sums = temp;
  hipLaunchKernelGGL(k_scan_tile_sums, dim3(unsigned(tiles)), dim3(kScanThreads), 0, st, in, count, sums);
  const hipError_t e = scan_levels(sums, sums, tiles, nullptr, temp + tiles, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_scan_apply, dim3(unsigned(tiles)), dim3(kScanThreads), 0, st, in, count, sums, out, total);
  return hipGetLastError();
}

hipError_t keep_scan_temp_bytes(uint64_t count, size_t* temp_bytes) {
  const uint64_t tiles = (count + kScanTile - 1) / kScanTile;
  if (count > kScanTwoPassMax) *temp_bytes = size_t(tiles + scan_temp_words(tiles)) * 4 + 4;
  else *temp_bytes = size_t(tiles) * 4 + 4;
  return hipSuccess;
}

hipError_t launch_keep_scan(int mode, const uint32_t* slot_of, const uint32_t* table, const void* dead_bits,
    uint64_t count, uint32_t* keep, uint32_t* pos, uint32_t* total, void* temp, size_t* temp_bytes, hipStream_t st) {
  const uint8_t* dead = static_cast<const uint8_t*>(dead_bits);
  if (count > 0x7fffffffull) return hipErrorInvalidValue;
  if (count <= kScanTile) {
    hipLaunchKernelGGL(k_keep_tile_scan, dim3(1), dim3(kScanThreads), 0, st, mode, slot_of, table, dead, count, keep,
        nullptr, pos, total, nullptr);
    return hipGetLastError();
  }
  const uint32_t tiles = uint32_t((count + kScanTile - 1) / kScanTile);
  size_t need = 0;
  keep_scan_temp_bytes(count, &need);
  if (!temp || !temp_bytes || *temp_bytes < need) return hipErrorInvalidValue;
  uint32_t* sums = static_cast<uint32_t*>(temp);
  hipLaunchKernelGGL(k_keep_tile_sums, dim3(tiles), dim3(kScanThreads), 0, st, mode, slot_of, table, dead, count,
      keep, sums);
  const uint32_t* off = nullptr;
  if (count > kScanTwoPassMax) {
    // More tiles than one workgroup adds up: scan the tile sums first.
    const hipError_t e = scan_levels(sums, sums, tiles, nullptr, sums + tiles, st);
    if (e != hipSuccess) return e;
    off = sums;
  }
  hipLaunchKernelGGL(k_keep_tile_scan, dim3(tiles), dim3(kScanThreads), 0, st, mode, slot_of, table, dead, count,
      keep, sums, pos, total, off);
  return hipGetLastError();
}

hipError_t launch_goal_reset(int* found, unsigned long long* witness, hipStream_t st) {
  hipLaunchKernelGGL(k_goal_reset, dim3(1), dim3(1), 0, st, found, witness);
  return hipGetLastError();
}

hipError_t launch_scan(const uint32_t* keep, uint32_t* pos, uint64_t count, uint32_t* total, void* temp,
    size_t* temp_bytes, hipStream_t st) {
  if (count > 0x7fffffffull) return hipErrorInvalidValue;
  const size_t need = size_t(scan_temp_words(count)) * 4 + 4;
  if (!temp) {
    *temp_bytes = need;
    return hipSuccess;
  }
  if (!temp_bytes || *temp_bytes < need) return hipErrorInvalidValue;
  return scan_levels(keep, pos, count, total, static_cast<uint32_t*>(temp), st);
}

template <int W>
static hipError_t commit_w(const uint64_t* cand, const uint32_t* keep, const uint32_t* pos, const uint32_t* slot_of,
    uint64_t count, uint64_t m, const uint32_t* parent_idx, uint64_t parent_base, const uint32_t* letter_list,
    uint64_t base, uint64_t* arena, uint32_t* parent, uint8_t* letter, uint32_t* table, uint64_t* front,
    uint32_t* front_idx, int* found, unsigned long long* witness, hipStream_t st) {
  if (!count) return found ? launch_goal_reset(found, witness, st) : hipSuccess;
  hipLaunchKernelGGL(k_commit<W>, dim3(blocks_for(count)), dim3(256), 0, st, cand, keep, pos, slot_of, count, m,
      parent_idx, parent_base, letter_list, base, arena, parent, letter, table, front, front_idx, found, witness);
  return hipGetLastError();
}

hipError_t launch_commit(int W, const uint64_t* cand, const uint32_t* keep, const uint32_t* pos,
    const uint32_t* slot_of, uint64_t count, uint64_t m, const uint32_t* parent_idx, uint64_t parent_base,
    const uint32_t* letter_list, uint64_t base, uint64_t* arena, uint32_t* parent, uint8_t* letter, uint32_t* table,
    hipStream_t st, uint64_t* front, uint32_t* front_idx, int* found, unsigned long long* witness) {
  SW_DISPATCH_W(W, commit_w<w_>(cand, keep, pos, slot_of, count, m, parent_idx, parent_base, letter_list, base,
      arena, parent, letter, table, front, front_idx, found, witness, st))
}

template <int W>
__global__ __launch_bounds__(256) void k_hash_rebuild(const uint64_t* __restrict__ arena, uint64_t rows,
    uint32_t* table, uint64_t mask) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  uint64_t pos = set_hash<W>(arena + i * W) & mask;
  while (atomicCAS(&table[pos], kEmpty, uint32_t(i)) != kEmpty) pos = (pos + 1) & mask;
}

template <int W>
static hipError_t rebuild_w(const uint64_t* arena, uint64_t rows, uint32_t* table, uint64_t mask, hipStream_t st) {
  if (!rows) return hipSuccess;
  hipLaunchKernelGGL(k_hash_rebuild<W>, dim3(blocks_for(rows)), dim3(256), 0, st, arena, rows, table, mask);
  return hipGetLastError();
}

hipError_t launch_hash_rebuild(int W, const uint64_t* arena, uint64_t rows, uint32_t* table, uint64_t mask,
    hipStream_t st) {
  SW_DISPATCH_W(W, rebuild_w<w_>(arena, rows, table, mask, st))
}

hipError_t launch_alive(int W, const void* dead_bits, uint64_t count, uint32_t* keep, hipStream_t st) {
  (void)W;
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(k_alive, dim3(blocks_for(count)), dim3(256), 0, st, static_cast<const uint8_t*>(dead_bits),
      count, keep);
  return hipGetLastError();
}

template <int W>
static hipError_t select_w(const uint64_t* rows, uint64_t lo, const uint32_t* keep, const uint32_t* pos,
    uint64_t count, uint64_t* out, uint32_t* idx_out, hipStream_t st) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(k_select<W>, dim3(blocks_for(count)), dim3(256), 0, st, rows, lo, keep, pos, count, out,
      idx_out);
  return hipGetLastError();
}

hipError_t launch_select(int W, const uint64_t* rows, uint64_t lo, const uint32_t* keep, const uint32_t* pos,
    uint64_t count, uint64_t* out, uint32_t* idx_out, hipStream_t st) {
  SW_DISPATCH_W(W, select_w<w_>(rows, lo, keep, pos, count, out, idx_out, st))
}

template <int W>
static hipError_t beam_select_w(const uint64_t* cand, const uint32_t* keep, const uint32_t* pos, uint64_t count,
    uint32_t n, uint64_t* out, uint32_t* idx_out, uint32_t* idx_copy, unsigned long long* first, uint32_t* table,
    const uint32_t* slot_of, hipStream_t st) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(k_beam_select<W>, dim3(blocks_for(count)), dim3(256), 0, st, cand, keep, pos, count, n, out,
      idx_out, idx_copy, first, table, slot_of);
  return hipGetLastError();
}

hipError_t launch_beam_select(int W, const uint64_t* cand, const uint32_t* keep, const uint32_t* pos,
    uint64_t count, uint32_t n, uint64_t* out, uint32_t* idx_out, uint32_t* idx_copy, unsigned long long* first,
    uint32_t* table, const uint32_t* slot_of, hipStream_t st) {
  SW_DISPATCH_W(W, beam_select_w<w_>(cand, keep, pos, count, n, out, idx_out, idx_copy, first, table, slot_of, st))
}

template <int W>
static hipError_t count_bits_w(const uint64_t* sets, uint64_t count, uint32_t n, unsigned long long* acc,
    hipStream_t st) {
  if (!count) return hipSuccess;
  unsigned blocks = blocks_for(count);
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(k_count_bits<W>, dim3(blocks), dim3(256), 0, st, sets, count, n, acc);
  return hipGetLastError();
}

hipError_t launch_count_bits(int W, const uint64_t* sets, uint64_t count, uint32_t n, unsigned long long* acc,
    hipStream_t st) {
  SW_DISPATCH_W(W, count_bits_w<w_>(sets, count, n, acc, st))
}

hipError_t launch_copy(const void* src, void* dst, uint64_t bytes, hipStream_t st) {
  const uint64_t n16 = bytes / 16;
  if (!n16) return hipSuccess;
  unsigned blocks = unsigned((n16 + 255) / 256);
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, st, static_cast<const uint4*>(src), static_cast<uint4*>(dst),
      n16);
  return hipGetLastError();
}

hipError_t launch_hash_lookup(int W, const uint64_t* cand, uint32_t* valid, uint64_t count, const uint32_t* table,
    uint64_t mask, const uint64_t* arena, hipStream_t st) {
  if (!count) return hipSuccess;
  const unsigned blocks = unsigned((count + 255) / 256);
  switch (W) {
#define C(w) case w: hipLaunchKernelGGL(k_hash_lookup<w>, dim3(blocks), dim3(256), 0, st, cand, valid, count, table, mask, arena); break;
    C(1) C(2) C(3) C(4) C(5) C(6) C(7) C(8)
#undef C
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, uint64_t count, hipStream_t st) {
  if (!count) return hipSuccess;
  unsigned blocks = blocks_for(count);
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, st, p, v, count);
  return hipGetLastError();
}

}  // namespace sw
