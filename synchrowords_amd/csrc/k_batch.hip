// Batched exact search: one persistent launch, one workgroup per instance.
//
// Workgroups pull instance indices from a device queue (an atomic counter
// over a cost-ordered list) and run the whole meet in the middle of that
// instance on their own: forward images of Q, backward preimages of the
// sink singletons (kept complemented), exact dedup per side in a hash table in
// the workgroup's scratch, and after every step a check for a forward set
// disjoint from a backward complement.  Instances that outgrow the
// workgroup's scratch are reported as overflowed and finished by the
// whole-device search.
//
// The search is the one of exact.cpp (the reference's meet in the middle,
// synchrolib/algorithm/exact/meet_in_the_middle.cpp:57-344) without
// dominance pruning: dedup alone keeps it exact.
#include "sw_batch.hpp"
#include "sw_device.hpp"
#include "sw_kernels.hpp"

namespace sw {

namespace {

constexpr uint32_t kGoalWords = 16;               // 64-set words per tile row
constexpr uint32_t kGoalTile = kGoalWords * 64;   // backward sets per tile

template <int W>
struct SetW {
  uint64_t v[W];
};

// Scratch of one workgroup, carved from a base pointer.
template <int W>
struct Scratch {
  uint64_t* arena[2];   // [side][row][W]
  uint32_t* parent[2];  // [side][row]
  uint8_t* letter[2];   // [side][row]
  uint32_t* table[2];   // [side][slot]
  uint64_t* cand;       // [cand][W]
  uint32_t* slot;       // [cand]
};

template <int W>
__device__ Scratch<W> carve(uint8_t* base, const BatchLimits& L) {
  Scratch<W> s;
  uint8_t* p = base;
  for (int side = 0; side < 2; ++side) {
    s.arena[side] = reinterpret_cast<uint64_t*>(p);
    p += size_t(L.cap_rows) * W * 8;
    s.parent[side] = reinterpret_cast<uint32_t*>(p);
    p += size_t(L.cap_rows) * 4;
    s.table[side] = reinterpret_cast<uint32_t*>(p);
    p += size_t(L.hash_slots) * 4;
    s.letter[side] = p;
    p += (size_t(L.cap_rows) + 255) & ~size_t(255);
  }
  s.cand = reinterpret_cast<uint64_t*>(p);
  p += size_t(L.cap_cand) * W * 8;
  s.slot = reinterpret_cast<uint32_t*>(p);
  return s;
}

}  // namespace

size_t batch_scratch_bytes(int W, const BatchLimits& L) {
  size_t b = 0;
  for (int side = 0; side < 2; ++side) {
    b += size_t(L.cap_rows) * W * 8 + size_t(L.cap_rows) * 4 + size_t(L.hash_slots) * 4;
    b += (size_t(L.cap_rows) + 255) & ~size_t(255);
  }
  b += size_t(L.cap_cand) * W * 8 + size_t(L.cap_cand) * 4;
  return (b + 4095) & ~size_t(4095);
}

// Exclusive scan of 0/1 flags over the workgroup of T threads: a ballot per
// wave, the lanes below in it, and the totals of the waves before (two
// barriers).
template <int T>
__device__ inline void block_scan(bool v, uint32_t* sh, uint32_t& excl, uint32_t& total) {
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint64_t b = __ballot(v);
  const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
  if (lane == 0) sh[wv] = uint32_t(__popcll(b));
  __syncthreads();
  uint32_t before = 0, tot = 0;
#pragma unroll
  for (uint32_t w = 0; w < T / 64; ++w) {
    before += w < wv ? sh[w] : 0u;
    tot += sh[w];
  }
  excl = before + uint32_t(__popcll(b & below));
  total = tot;
  __syncthreads();
}

// T threads per workgroup: 256 in the batch's persistent launch (several
// instances per CU at a time), 1024 for a lone automaton (one workgroup has
// the whole device to itself).
template <int W, int T>
__global__ __launch_bounds__(T) void k_batch(const BatchInstance* __restrict__ inst,
    const uint8_t* __restrict__ deltas, const uint32_t* __restrict__ order, uint32_t count, uint32_t* head,
    BatchResult* __restrict__ out, uint8_t* __restrict__ words, uint32_t word_cap, uint8_t* scratch_base,
    size_t scratch_stride, BatchLimits L) {
  constexpr int MAXN = W * 64;
  __shared__ uint8_t sd[MAXN * kBatchMaxK];             // delta
  __shared__ uint64_t pre[kBatchMaxK][MAXN][W];         // preimage of each state per letter
  __shared__ uint32_t sh[T];
  __shared__ uint32_t s_inst, s_found, s_stop;
  __shared__ unsigned long long s_wit;
  __shared__ uint32_t s_tcap[2];
  // Goal-check tile: kGoalTile backward sets transposed, row q = which of
  // them hold state q (64 sets per word, kGoalWords words per state).
  __shared__ uint64_t gtile[MAXN * kGoalWords];

  Scratch<W> S = carve<W>(scratch_base + size_t(blockIdx.x) * scratch_stride, L);
  const int tid = threadIdx.x;

  while (true) {
    if (tid == 0) s_inst = atomicAdd(head, 1u);
    __syncthreads();
    const uint32_t qi = s_inst;
    if (qi >= count) return;
    const uint32_t id = order[qi];
    const BatchInstance I = inst[id];
    const uint32_t n = I.n, K = I.k;
    if (n > MAXN || K > kBatchMaxK || n < 2) {
      if (tid == 0) out[id].status = kBatchUnsupported;
      __syncthreads();
      continue;
    }
    // ---- per-instance tables in LDS
    for (uint32_t i = tid; i < n * K; i += T) sd[i] = deltas[I.delta_off + i];
    for (uint32_t i = tid; i < K * MAXN * W; i += T) (&pre[0][0][0])[i] = 0;
    __syncthreads();
    for (uint32_t i = tid; i < n * K; i += T) {
      const uint32_t q = i / K, a = i - q * K, t = sd[i];
      atomicOr(reinterpret_cast<unsigned long long*>(&pre[a][t][q >> 6]), 1ull << (q & 63));
    }
    uint64_t mask[W];
#pragma unroll
    for (int w = 0; w < W; ++w) mask[w] = word_mask(n, w);

    // ---- seed both sides
    uint32_t rows[2] = {0, 0}, lo[2] = {0, 0}, hi[2] = {0, 0}, level0_end_b = 0;
    // hash tables start small and grow with the side
    uint32_t tcap[2] = {1024, 1024};
    for (int side = 0; side < 2; ++side)
      for (uint32_t i = tid; i < tcap[side]; i += T) S.table[side][i] = kEmpty;
    __syncthreads();
    if (tid == 0) {
#pragma unroll
      for (int w = 0; w < W; ++w) S.arena[0][w] = mask[w];
      S.parent[0][0] = 0xffffffffu;
      S.letter[0][0] = 0xff;
      S.table[0][set_hash<W>(S.arena[0]) & (tcap[0] - 1)] = 0;
      uint32_t nb = 0;
      for (uint32_t q = 0; q < n; ++q) {
        if (!((I.b0[q >> 6] >> (q & 63)) & 1)) continue;
        uint64_t* r = S.arena[1] + size_t(nb) * W;
#pragma unroll
        for (int w = 0; w < W; ++w) r[w] = mask[w];
        r[q >> 6] &= ~(1ull << (q & 63));
        S.parent[1][nb] = 0xffffffffu;
        S.letter[1][nb] = 0xff;
        uint64_t pos = set_hash<W>(r) & (tcap[1] - 1);
        while (S.table[1][pos] != kEmpty) pos = (pos + 1) & (tcap[1] - 1);
        S.table[1][pos] = nb;
        ++nb;
      }
      sh[0] = nb;
    }
    __syncthreads();
    rows[0] = 1;
    hi[0] = 1;
    rows[1] = sh[0];
    hi[1] = rows[1];
    level0_end_b = rows[1];
    __syncthreads();

    uint32_t status = kBatchBound, length = I.upper, detail = kOverflowNone;
    uint32_t t = 0;
    unsigned long long witness = 0;
    bool found = false;
    while (t + 1 < I.upper) {
      const uint32_t fcount = hi[0] - lo[0], bcount = hi[1] - lo[1];
      if (!fcount || !bcount) break;  // all continuations pruned: the bound is the length
      const int side = fcount <= bcount ? 0 : 1;
      const uint32_t m = hi[side] - lo[side];
      const uint64_t ncand = uint64_t(m) * K;
      if (ncand > L.cap_cand || rows[side] + ncand > L.cap_rows || t + 1 >= word_cap) {
        status = kBatchOverflow;
        detail = ncand > L.cap_cand ? kOverflowCand : (rows[side] + ncand > L.cap_rows ? kOverflowRows : kOverflowWord);
        break;
      }
      ++t;
      // ---- grow the hash table when it would pass half load
      const uint64_t need = 2 * (uint64_t(rows[side]) + ncand);
      if (need > tcap[side]) {
        uint32_t nc = tcap[side];
        while (nc < need) nc <<= 1;
        if (nc > L.hash_slots) {
          status = kBatchOverflow;
          detail = kOverflowHash;
          break;
        }
        tcap[side] = nc;
        for (uint32_t i = tid; i < nc; i += T) S.table[side][i] = kEmpty;
        __syncthreads();
        for (uint32_t r = tid; r < rows[side]; r += T) {
          uint64_t pos = set_hash<W>(S.arena[side] + size_t(r) * W) & (nc - 1);
          while (atomicCAS(&S.table[side][pos], kEmpty, r) != kEmpty) pos = (pos + 1) & (nc - 1);
        }
        __syncthreads();
      }
      const uint32_t tmask = tcap[side] - 1;
      // ---- expand and insert
      for (uint32_t c = tid; c < ncand; c += T) {
        const uint32_t a = c / m, j = lo[side] + (c - a * m);
        const uint64_t* x = S.arena[side] + size_t(j) * W;
        uint64_t r[W];
#pragma unroll
        for (int w = 0; w < W; ++w) r[w] = 0;
        bool ok = true;
        if (side == 0) {
#pragma unroll
          for (int w = 0; w < W; ++w) {
            uint64_t bits = x[w];
            while (bits) {
              const uint32_t q = w * 64 + __builtin_ctzll(bits);
              bits &= bits - 1;
              const uint32_t tq = sd[q * K + a];
#pragma unroll
              for (int v = 0; v < W; ++v) r[v] |= (int(tq >> 6) == v) ? (1ull << (tq & 63)) : 0ull;
            }
          }
        } else {
          // Complemented set C: preimage(C) = ~preimage(~C); walk the smaller.
          int pc = 0;
#pragma unroll
          for (int w = 0; w < W; ++w) pc += __popcll(x[w]);
          const bool inv = 2 * uint32_t(pc) > n;
#pragma unroll
          for (int w = 0; w < W; ++w) {
            uint64_t bits = inv ? (~x[w] & mask[w]) : x[w];
            while (bits) {
              const uint32_t q = w * 64 + __builtin_ctzll(bits);
              bits &= bits - 1;
#pragma unroll
              for (int v = 0; v < W; ++v) r[v] |= pre[a][q][v];
            }
          }
          if (inv) {
#pragma unroll
            for (int w = 0; w < W; ++w) r[w] = ~r[w] & mask[w];
          }
          int cc = 0;
#pragma unroll
          for (int w = 0; w < W; ++w) cc += __popcll(r[w]);
          ok = uint32_t(cc) + 1 < n;  // drop preimages of at most one state
        }
        uint64_t* dst = S.cand + size_t(c) * W;
#pragma unroll
        for (int w = 0; w < W; ++w) dst[w] = r[w];
        __threadfence_block();  // the row is visible before its slot claim
        uint32_t res = kEmpty;
        if (ok) {
          const uint32_t mine = kCand | c;
          uint64_t pos = set_hash<W>(r) & tmask;
          while (true) {
            uint32_t v = __hip_atomic_load(&S.table[side][pos], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            bool next = false;
            while (!next) {
              if (v == kEmpty) {
                const uint32_t prev = atomicCAS(&S.table[side][pos], kEmpty, mine);
                if (prev == kEmpty) {
                  res = uint32_t(pos);
                  goto inserted;
                }
                v = prev;
                continue;
              }
              const uint64_t* o = (v & kCand) ? S.cand + size_t(v & ~kCand) * W : S.arena[side] + size_t(v) * W;
              bool eq = true;
#pragma unroll
              for (int w = 0; w < W; ++w) eq &= (o[w] == r[w]);
              if (!eq) {
                next = true;
                break;
              }
              if ((v & kCand) && (v & ~kCand) > c) {
                const uint32_t prev = atomicCAS(&S.table[side][pos], v, mine);
                if (prev == v) {
                  res = uint32_t(pos);
                  goto inserted;
                }
                v = prev;
                continue;
              }
              goto inserted;  // duplicate
            }
            pos = (pos + 1) & tmask;
          }
        }
      inserted:
        S.slot[c] = res;
      }
      __threadfence_block();
      __syncthreads();
      // ---- resolve, scan, commit (candidate order, so rows are deterministic)
      uint32_t base = rows[side];
      for (uint32_t c0 = 0; c0 < ncand; c0 += T) {
        const uint32_t c = c0 + tid;
        uint32_t keep = 0, slot = kEmpty;
        if (c < ncand) {
          slot = S.slot[c];
          keep = (slot != kEmpty && S.table[side][slot] == (kCand | c)) ? 1u : 0u;
        }
        uint32_t excl, total;
        block_scan<T>(keep != 0, sh, excl, total);
        if (keep) {
          const uint32_t dst = base + excl;
          const uint64_t* src = S.cand + size_t(c) * W;
#pragma unroll
          for (int w = 0; w < W; ++w) S.arena[side][size_t(dst) * W + w] = src[w];
          const uint32_t a = c / m;
          S.parent[side][dst] = lo[side] + (c - a * m);
          S.letter[side][dst] = uint8_t(a);
          S.table[side][slot] = dst;
        }
        base += total;
      }
      __threadfence_block();
      __syncthreads();
      lo[side] = rows[side];
      hi[side] = base;
      rows[side] = base;
      // ---- goal: forward row A disjoint from backward complement C
      const uint32_t nf = hi[0] - lo[0], nbk = hi[1] - lo[1];
      if (uint64_t(nf) * nbk > L.max_pairs) {
        status = kBatchOverflow;
        detail = kOverflowPairs;
        break;
      }
      if (tid == 0) {
        s_found = 0;
        s_wit = ~0ull;
      }
      __syncthreads();
      if (nf && nbk) {
        // Backward sets B = Q \ C by tiles of kGoalTile, transposed into LDS
        // with 64 x 64 lane transposes; forward sets streamed by groups of
        // kGoalWords lanes (lane l holds tile sets 64 l .. 64 l + 63): the
        // AND of the rows of a forward set's states keeps the tile sets
        // that contain it, and the group stops once none can.
        const uint32_t lane = tid & 63, wv = tid >> 6;
        const uint32_t gi = tid / kGoalWords, li = tid % kGoalWords;
        const uint32_t gshift = (tid & 63) & ~(kGoalWords - 1);
        const uint64_t gmask = ((1ull << kGoalWords) - 1) << gshift;
        const uint64_t* fbase = S.arena[0] + size_t(lo[0]) * W;
        const uint64_t* bbase = S.arena[1] + size_t(lo[1]) * W;
        for (uint32_t t0 = 0; t0 < nbk; t0 += kGoalTile) {
          const uint32_t tn = nbk - t0 < kGoalTile ? nbk - t0 : kGoalTile;
          const uint32_t ng = (tn + 63) / 64;
          for (uint32_t g = wv; g < ng; g += T / 64) {
            const uint32_t j = g * 64 + lane;
            const bool ok = j < tn;
#pragma unroll
            for (int w = 0; w < W; ++w) {
              const uint64_t v = ok ? (~bbase[size_t(t0 + j) * W + w] & mask[w]) : 0ull;
              const uint64_t tr = lane_transpose64(v);
              if (uint32_t(w) * 64 + lane < n) gtile[(w * 64 + lane) * kGoalWords + g] = tr;
            }
          }
          __syncthreads();
          const uint64_t valid = li < ng ? (li + 1 < ng || (tn & 63) == 0 ? ~0ull : ((1ull << (tn & 63)) - 1)) : 0ull;
          const uint32_t groups = T / kGoalWords;
          for (uint32_t s0 = 0; s0 < nf; s0 += groups) {
            if (__hip_atomic_load(&s_found, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
            const uint32_t si = s0 + gi;
            uint64_t acc = si < nf ? valid : 0ull;
            if (si < nf) {
#pragma unroll
              for (int w = 0; w < W; ++w) {
                uint64_t bits = fbase[size_t(si) * W + w];
                while (bits && (__ballot(acc != 0) & gmask)) {
                  const uint32_t q = uint32_t(w) * 64 + __builtin_ctzll(bits);
                  bits &= bits - 1;
                  acc &= gtile[q * kGoalWords + li];
                }
              }
            }
            if (acc) {
              const uint32_t bi = t0 + li * 64 + uint32_t(__builtin_ctzll(acc));
              atomicMin(&s_wit, (unsigned long long)(uint64_t(si) << 32 | bi));
              s_found = 1;
            }
          }
          __syncthreads();
          if (s_found) break;
        }
      }
      __syncthreads();
      if (s_found) {
        found = true;
        witness = s_wit;
        break;
      }
    }
    // ---- result and word
    if (tid == 0) {
      BatchResult R;
      R.status = status;
      R.length = length;
      R.word_len = 0;
      R.steps = t;
      R.detail = detail;
      if (found) {
        R.status = kBatchFound;
        R.length = t;
        uint8_t* wout = words + size_t(id) * word_cap;
        uint32_t fr = lo[0] + uint32_t(witness >> 32), br = lo[1] + uint32_t(witness & 0xffffffffu);
        // forward letters, root first: collect backwards then reverse
        uint32_t len = 0;
        for (uint32_t r = fr; r != 0; r = S.parent[0][r]) wout[len++] = S.letter[0][r];
        for (uint32_t i = 0; i < len / 2; ++i) {
          uint8_t tmp = wout[i];
          wout[i] = wout[len - 1 - i];
          wout[len - 1 - i] = tmp;
        }
        for (uint32_t r = br; r >= level0_end_b; r = S.parent[1][r]) wout[len++] = S.letter[1][r];
        R.word_len = len;
      }
      out[id] = R;
    }
    __syncthreads();
  }
}

hipError_t launch_batch(int W, const BatchInstance* inst, const uint8_t* deltas, const uint32_t* order,
    uint32_t count, uint32_t* head, BatchResult* out, uint8_t* words, uint32_t word_cap, uint8_t* scratch,
    size_t scratch_stride, const BatchLimits& L, uint32_t grid, hipStream_t st, uint32_t threads) {
  if (threads != 256 && threads != 1024) return hipErrorInvalidValue;
#define CASE(w)                                                                                              \
  case w:                                                                                                    \
    if (threads == 256)                                                                                      \
      hipLaunchKernelGGL((k_batch<w, 256>), dim3(grid), dim3(256), 0, st, inst, deltas, order, count, head, out, \
          words, word_cap, scratch, scratch_stride, L);                                                      \
    else                                                                                                     \
      hipLaunchKernelGGL((k_batch<w, 1024>), dim3(grid), dim3(1024), 0, st, inst, deltas, order, count, head,    \
          out, words, word_cap, scratch, scratch_stride, L);                                                 \
    return hipGetLastError();
  switch (W) {
    CASE(1) CASE(2) CASE(3)
  }
#undef CASE
  return hipErrorInvalidValue;
}

}  // namespace sw
